"""Learned warm start for the Uzawa solve and learned coarse-level correction
for its multigrid (no counterpart in the reference).

A small 3-D convolutional network (4 layers of 3x3x3 convolutions, 16
channels, ReLU) maps (temperature, log10 viscosity) on a coarse multigrid
level to (velocity, nodal pressure) / Ra on that level; the prediction is
prolonged to the finest level with the solver's own interp_vector and the
pressure averaged to elements.  It is used as the initial U, P of the Uzawa
solve (the reference starts from E->U, E->P, Stokes_flow_Incomp.c:354); the
solve still iterates to the same criterion, so the result is judged exactly
like the plain one.

Training data come from the plain solver itself (random temperature fields
and viscosity contrasts); training runs on the GPU through the BF16 MFMA
kernels of csrc/learn_kernels.hip (include/citcom_learn.h).  PyTorch holds
device buffers only.

The coarse-level correction is a second network of the same kind on the
coarsest multigrid level: (right-hand side / its RMS, log10 viscosity) ->
the level's exact solve / that RMS, trained on the coarsest operators of
random temperature fields.  Installed with Citcom.set_coarse_net, it takes
the place of the coarsest level's smoother call inside every V-cycle
(General_matrix_functions.c:574, :613); each inner solve still iterates to
its own residual criterion.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import time

import numpy as np

from . import deck as _deck
from .hip import Citcom, lib

CH = 16
NIN = 2    # T, log10 eta
NOUT = 4   # u, v, w, p (scaled by 1/Ra)


def _cl():
    L = lib()
    if not getattr(L, "_cl_ready", False):
        P, i, l, f = C.c_void_p, C.c_int, C.c_long, C.c_float
        L.cl_conv3d_fwd.argtypes = [P, i, i, i, i, P, P, i, P, P, P]
        L.cl_relu_bwd.argtypes = [P, P, l, P]
        L.cl_conv3d_bwd_data.argtypes = [P, i, i, i, i, P, P, P]
        L.cl_conv3d_bwd_weight_workspace.argtypes = [i, i, i, i]
        L.cl_conv3d_bwd_weight_workspace.restype = l
        L.cl_conv3d_bwd_weight.argtypes = [P, P, i, i, i, i, P, P, P]
        L.cl_mse.argtypes = [P, P, l, i, P, P, P]
        L.cl_adam.argtypes = [P, P, P, P, l, f, f, f, f, i, P, P, P]
        L.cl_to_bf16.argtypes = [P, P, l, P]
        L.cl_conv3d_bwd_weight_noalloc.argtypes = [P, P, i, i, i, i, P, P]
        L.cl_scratch_bytes.restype = l
        L.cl_scratch_bytes.argtypes = []
        L.cl_last_error.restype = C.c_char_p
        L.cl_last_error.argtypes = []
        L._cl_ready = True
    return L


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed")


def _p(t):
    return C.c_void_p(t.data_ptr())


class ConvNet:
    """Weights (FP32 masters + BF16 copies) and activations on the device."""

    def __init__(self, layers: int = 4, seed: int = 0, nin: int = NIN, nout: int = NOUT):
        import torch
        self.torch = torch
        g = torch.Generator().manual_seed(seed)
        self.layers = layers
        self.nout = nout
        self.w, self.b = [], []
        for li in range(layers):
            cin = nin if li == 0 else CH
            cout = nout if li == layers - 1 else CH
            w = torch.zeros(CH, 27, CH)
            w[:cout, :, :cin] = torch.randn(cout, 27, cin, generator=g) * (2.0 / (27 * cin)) ** 0.5
            self.w.append(w.cuda())
            self.b.append(torch.zeros(CH).cuda())
        self.wbf = [w.to(torch.bfloat16) for w in self.w]
        self.m = [[torch.zeros_like(t) for t in (w, b)] for w, b in zip(self.w, self.b)]
        self.v = [[torch.zeros_like(t) for t in (w, b)] for w, b in zip(self.w, self.b)]
        self.steps = 0

    def forward(self, x, dims):
        """x: bf16 [nb][nv][16] on the device -> FP32 [nb][nv][16]; keeps the
        activations for backward."""
        torch = self.torch
        nx, ny, nz = dims
        nb = x.shape[0]
        L = _cl()
        self.acts = [x]
        out_f = torch.empty(x.shape, dtype=torch.float32, device=x.device)
        for li in range(self.layers):
            last = li == self.layers - 1
            y = torch.empty_like(x)
            _ok(L.cl_conv3d_fwd(_p(self.acts[-1]), nb, nx, ny, nz, _p(self.wbf[li]), _p(self.b[li]), int(not last),
                                _p(y), _p(out_f) if last else None, None), "conv3d_fwd")
            self.acts.append(y)
        self.dims = dims
        return out_f

    def backward(self, dy):
        """dy: bf16 d loss / d output -> per-layer FP32 gradients [dW (16*27*16) | db (16)]"""
        torch = self.torch
        nx, ny, nz = self.dims
        nb = dy.shape[0]
        L = _cl()
        ws = torch.empty(L.cl_conv3d_bwd_weight_workspace(nb, nx, ny, nz) // 4, dtype=torch.float32,
                         device=dy.device)
        grads = [None] * self.layers
        for li in reversed(range(self.layers)):
            if li != self.layers - 1:
                _ok(L.cl_relu_bwd(_p(self.acts[li + 1]), _p(dy), dy.numel(), None), "relu_bwd")
            g = torch.empty(CH * 27 * CH + CH, dtype=torch.float32, device=dy.device)
            _ok(L.cl_conv3d_bwd_weight(_p(self.acts[li]), _p(dy), nb, nx, ny, nz, _p(g), _p(ws), None),
                "conv3d_bwd_weight")
            grads[li] = g
            if li > 0:
                dx = torch.empty_like(dy)
                _ok(L.cl_conv3d_bwd_data(_p(dy), nb, nx, ny, nz, _p(self.wbf[li]), _p(dx), None), "conv3d_bwd_data")
                dy = dx
        return grads

    def adam(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = _cl()
        self.steps += 1
        for li, g in enumerate(grads):
            gw, gb = g[:CH * 27 * CH].view(CH, 27, CH), g[CH * 27 * CH:]
            _ok(L.cl_adam(_p(self.w[li]), _p(gw), _p(self.m[li][0]), _p(self.v[li][0]), self.w[li].numel(), lr, b1, b2,
                          eps, self.steps, _p(self.wbf[li]), None, None), "adam")
            _ok(L.cl_adam(_p(self.b[li]), _p(gb), _p(self.m[li][1]), _p(self.v[li][1]), CH, lr, b1, b2, eps,
                          self.steps, None, None, None), "adam")

    def loss_grad(self, yf, target):
        """MSE over the NOUT channels; returns (loss, dy bf16)"""
        torch = self.torch
        L = _cl()
        nvox = yf.shape[0] * yf.shape[1]
        nblk = (nvox * CH + 255) // 256
        part = torch.empty(nblk, dtype=torch.float32, device=yf.device)
        dy = torch.empty(yf.shape, dtype=torch.bfloat16, device=yf.device)
        _ok(L.cl_mse(_p(yf), _p(target), nvox, self.nout, _p(dy), _p(part), None), "mse")
        return float(part.sum()) / (nvox * self.nout), dy


# ---------------------------------------------------------------------------
# data: the plain solver on random temperature fields


def random_temperature(p: _deck.Problem, rng: np.random.Generator) -> np.ndarray:
    """T = 1 - z + a few random Fourier modes, boundary rows exact."""
    nox, noy, noz = p.dims()
    x, y, z = (c.astype(np.float64) for c in p.coords)
    X, Y, Z = np.meshgrid(x, y, z, indexing="xy")   # [noy][nox][noz]
    T = 1.0 - Z
    Lx, Ly = x[-1], y[-1]
    for _ in range(4):
        a, b = rng.integers(0, 3), rng.integers(0, 3)
        amp = rng.uniform(-0.08, 0.08)
        T = T + amp * np.cos(a * np.pi * X / Lx + rng.uniform(0, 6.3)) * np.cos(b * np.pi * Y / Ly) * np.sin(np.pi * Z)
    T = np.clip(T, 0.0, 1.0)
    T[..., 0] = 1.0
    T[..., -1] = 0.0
    return T.reshape(-1).astype(np.float32)   # node order k + noz*(i + nox*j)


def _nodal_from_elements(v_el: np.ndarray, dims) -> np.ndarray:
    """average of the (up to 8) elements around each node"""
    nox, noy, noz = dims
    e = v_el.reshape(noy - 1, nox - 1, noz - 1)
    acc = np.zeros((noy, nox, noz))
    cnt = np.zeros((noy, nox, noz))
    for dj in (0, 1):
        for di in (0, 1):
            for dk in (0, 1):
                acc[dj:dj + noy - 1, di:di + nox - 1, dk:dk + noz - 1] += e
                cnt[dj:dj + noy - 1, di:di + nox - 1, dk:dk + noz - 1] += 1
    return (acc / cnt).reshape(-1)


def _inject(v_node: np.ndarray, dims, s: int) -> np.ndarray:
    nox, noy, noz = dims
    return v_node.reshape(noy, nox, noz)[::s, ::s, ::s].reshape(-1)


@dataclasses.dataclass
class Sample:
    x: np.ndarray   # [nv][16] inputs on the net level
    y: np.ndarray   # [nv][16] targets on the net level
    T: np.ndarray   # finest temperature
    contrast: float


def problem_with_contrast(base: str, contrast: float, levels: int | None = None) -> _deck.Problem:
    """the workload with viscosity contrast exp(viscE) = contrast, optionally
    with fewer levels (the same coarsest grid under a smaller finest one)"""
    keys = _deck.workload(base)
    keys["viscE"] = repr(float(np.log(contrast)))
    if levels is not None:
        keys["levels"] = levels
    return _deck.problem_from_deck({k: str(v) for k, v in keys.items()})


def net_level(p: _deck.Problem) -> int:
    return 1


def features(c: Citcom, p: _deck.Problem, T: np.ndarray) -> np.ndarray:
    """inputs on the net level: T and log10 eta (element mean of the Gauss
    point viscosity, averaged to nodes), injected"""
    dims = p.dims()
    s = 2 ** (p.levels - 1 - net_level(p))
    c.upload("T", T)
    c.get_system_viscosity()
    evi = c.download("EVI").reshape(-1, 8).astype(np.float64)
    logeta = _nodal_from_elements(np.log10(evi).mean(axis=1), dims)
    x = np.zeros((len(_inject(T, dims, s)), CH), np.float32)
    x[:, 0] = _inject(T.astype(np.float64), dims, s)
    x[:, 1] = _inject(logeta, dims, s)
    return x


def make_samples(base: str, n: int, seed: int, log=None) -> list[Sample]:
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        contrast = float(10 ** rng.uniform(2, 4))
        p = problem_with_contrast(base, contrast)
        T = random_temperature(p, rng)
        c = Citcom(p)
        c.initialize(T)
        x = features(c, p, T)
        c.general_stokes_solver()
        U, P = c.download("U"), c.download("P")
        c.close()
        dims = p.dims()
        s = 2 ** (p.levels - 1 - net_level(p))
        y = np.zeros_like(x)
        for d in range(3):
            y[:, d] = _inject(U[d::3], dims, s) / p.rayleigh
        y[:, 3] = _inject(_nodal_from_elements(P, dims), dims, s) / p.rayleigh
        out.append(Sample(x, y, T, contrast))
        if log and (k + 1) % 16 == 0:
            log(f"samples {k + 1}/{n}")
    return out


def train(samples: list[Sample], dims, epochs: int = 400, batch: int = 16, lr: float = 2e-3, seed: int = 0,
          log=None, nin: int = NIN, nout: int = NOUT, layers: int = 4) -> tuple[ConvNet, list[float]]:
    """Adam on the mean squared error of the targets scaled per channel to
    unit RMS (the scales are kept with the network)."""
    import torch
    net = ConvNet(layers=layers, seed=seed, nin=nin, nout=nout)
    ys = np.stack([s.y for s in samples])
    net.yscale = np.ones(CH, np.float32)
    net.yscale[:nout] = np.sqrt((ys[..., :nout].astype(np.float64) ** 2).mean(axis=(0, 1))) + 1e-30
    X = torch.tensor(np.stack([s.x for s in samples])).to(torch.bfloat16).cuda()
    Y = torch.tensor(ys / net.yscale).cuda()
    n = len(samples)
    g = torch.Generator().manual_seed(seed)
    losses = []
    t0 = time.time()
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g)
        tot = 0.0
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch].cuda()
            yf = net.forward(X[idx].contiguous(), dims)
            loss, dy = net.loss_grad(yf, Y[idx].contiguous())
            net.adam(net.backward(dy), lr=lr)
            tot += loss * len(idx)
        losses.append(tot / n)
        if log and (ep + 1) % 100 == 0:
            log(f"epoch {ep + 1}: loss {losses[-1]:.3e} ({time.time() - t0:.1f} s)")
    return net, losses


def predict(net: ConvNet, x: np.ndarray, dims) -> np.ndarray:
    import torch
    xt = torch.tensor(x[None]).to(torch.bfloat16).cuda()
    return net.forward(xt, dims)[0].cpu().numpy() * net.yscale


def warm_start(net: ConvNet, c: Citcom, p: _deck.Problem, T: np.ndarray) -> None:
    """set the context's U, P to the network's prediction for temperature T"""
    lev = net_level(p)
    x = features(c, p, T)
    nd = p.dims(lev)
    y = predict(net, x, nd).astype(np.float64) * p.rayleigh
    u = np.zeros(3 * len(y))
    u[0::3], u[1::3], u[2::3] = y[:, 0], y[:, 1], y[:, 2]
    pn = np.zeros(3 * len(y))
    pn[0::3] = y[:, 3]
    for L in range(lev + 1, p.levels):
        u = c.apply("interp", L, u)
        pn = c.apply("interp", L, pn)
    nox, noy, noz = p.dims()
    pnode = pn[0::3].reshape(noy, nox, noz)
    pel = 0.125 * sum(pnode[dj:dj + noy - 1, di:di + nox - 1, dk:dk + noz - 1]
                      for dj in (0, 1) for di in (0, 1) for dk in (0, 1))
    node = c.download("NODE")
    for d, bit in enumerate((0x2, 0x8, 0x4)):   # VBX, VBY, VBZ
        u[d::3][(node & bit) != 0] = 0.0
    c.upload("U", u)
    c.upload("P", pel.reshape(-1))


# ---------------------------------------------------------------------------
# learned coarse-level correction


def coarse_logeta(c: Citcom, p: _deck.Problem) -> np.ndarray:
    """log10 viscosity per coarsest-level node: the mean over each adjacent
    element's Gauss points, averaged over those elements (k_cnet_logeta)"""
    evi = c.download("EVI", 0).reshape(-1, 8).astype(np.float64)
    return _nodal_from_elements(np.log10(evi).mean(axis=1), p.dims(0))


def coarse_free(c: Citcom) -> np.ndarray:
    """dofs of the coarsest level without a velocity boundary condition"""
    node = c.download("NODE", 0)
    free = np.ones(3 * len(node), bool)
    for d, bit in enumerate((0x2, 0x8, 0x4)):   # VBX, VBY, VBZ
        free[d::3][(node & bit) != 0] = False
    return free


def coarse_operator(c: Citcom) -> tuple[np.ndarray, np.ndarray]:
    """the coarsest level's stiffness on its free dofs, column by column
    from the device matvec"""
    free = coarse_free(c)
    idx = np.flatnonzero(free)
    K = np.zeros((len(idx), len(idx)))
    e = np.zeros(len(free))
    for col, j in enumerate(idx):
        e[j] = 1.0
        K[:, col] = c.apply("matvec", 0, e)[idx]
        e[j] = 0.0
    return K, free


def random_coarse_rhs(dims, rng: np.random.Generator) -> np.ndarray:
    """white or smoothed (as restricted residuals are) random right-hand side
    in node order, three components per node"""
    nox, noy, noz = dims
    F = rng.standard_normal((noy, nox, noz, 3))
    if rng.random() < 0.5:
        from scipy.ndimage import gaussian_filter
        F = gaussian_filter(F, sigma=(1.0, 1.0, 1.0, 0.0), mode="nearest")
    return F.reshape(-1)


def coarse_features(F: np.ndarray, logeta: np.ndarray) -> tuple[np.ndarray, float]:
    """network input for right-hand side F: (F / rms F, log10 eta); and rms F"""
    s = float(np.sqrt(np.mean(F * F)))
    x = np.zeros((len(logeta), CH), np.float32)
    x[:, :3] = (F / (s if s > 0 else 1.0)).reshape(-1, 3)
    x[:, 3] = logeta
    return x, s


@dataclasses.dataclass
class CoarseSample:
    x: np.ndarray    # [nv][16] inputs
    y: np.ndarray    # [nv][16] exact coarse solve / rms F
    F: np.ndarray    # right-hand side (stripped)
    d: np.ndarray    # exact coarse solve
    mc: np.ndarray   # the multigrid's coarsest-level smoother call on F


def make_coarse_samples(base: str, levels: int, n_fields: int, n_rhs: int, seed: int,
                        log=None) -> list[CoarseSample]:
    """coarsest-level problems of the workload's family: random temperature
    fields under the workload's own viscosity law (every material layer as
    the deck has it) with `levels` levels, n_rhs random right-hand sides
    each, solved exactly on the free dofs"""
    rng = np.random.default_rng(seed)
    keys = _deck.workload(base)
    keys["levels"] = levels
    out = []
    for k in range(n_fields):
        p = _deck.problem_from_deck({kk: str(v) for kk, v in keys.items()})
        T = random_temperature(p, rng)
        c = Citcom(p)
        c.initialize(T)
        c.get_system_viscosity()
        c.construct_stiffness_B_matrix()
        leta = coarse_logeta(c, p)
        K, free = coarse_operator(c)
        for _ in range(n_rhs):
            F = random_coarse_rhs(p.dims(0), rng) * free
            d = np.zeros_like(F)
            d[free] = np.linalg.solve(K, F[free])
            x, s = coarse_features(F, leta)
            y = np.zeros_like(x)
            y[:, :3] = (d / s).reshape(-1, 3)
            mc = c.apply("gs0", 0, None, F, sweeps=p.v_steps_low)[0]   # what the coarse matrix encodes
            out.append(CoarseSample(x, y, F, d, mc))
        c.close()
        if log and (k + 1) % 8 == 0:
            log(f"coarse fields {k + 1}/{n_fields}")
    return out


def train_coarse(samples: list[CoarseSample], dims, epochs: int = 300, seed: int = 0, layers: int = 4,
                 log=None) -> tuple[ConvNet, list[float]]:
    return train(samples, dims, epochs=epochs, seed=seed, log=log, nin=4, nout=3, layers=layers)


def coarse_predict(net: ConvNet, x: np.ndarray, s: float, dims) -> np.ndarray:
    """host-side restatement of the device correction for one input (tests):
    the same network kernels, scaled back by rms F (BC rows not stripped)"""
    return (predict(net, x, dims)[:, :3].astype(np.float64) * s).reshape(-1)


def coarse_relative_errors(net: ConvNet, samples: list[CoarseSample], dims) -> tuple[float, float]:
    """mean relative L2 error against the exact coarse solve: the network's,
    and the coarsest-level smoother call's"""
    en, es = [], []
    for smp in samples:
        s = float(np.sqrt(np.mean(smp.F * smp.F)))
        dn = coarse_predict(net, smp.x, s, dims) * (smp.F != 0)
        nd = np.linalg.norm(smp.d)
        en.append(np.linalg.norm(dn - smp.d) / nd)
        es.append(np.linalg.norm(smp.mc - smp.d) / nd)
    return float(np.mean(en)), float(np.mean(es))


def install_coarse_net(net: ConvNet, c: Citcom, on: bool = True) -> None:
    """hand the network to the context and switch the correction on"""
    import torch
    w = np.stack([wb.view(torch.int16).cpu().numpy().view(np.uint16) for wb in net.wbf])
    b = np.stack([bb.cpu().numpy() for bb in net.b])
    c.set_coarse_net(w, b, net.yscale[:3])
    c.set_option("coarse_net", int(on))
