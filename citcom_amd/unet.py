"""The learned warm start's U-Net: training on the GPU with the project's own
BF16 MFMA kernels (citcom_learn.h: cl_conv3d_*_c, cl_restrict, cl_interp,
cl_wmse, cl_adam), samples produced by the plain solver itself, and the
trained weights handed to the library (chip_set_warm_net), whose time loop
runs the same network (cl_unet_forward) as the Uzawa solve's initial guess.
No counterpart in the reference: the network only sets the initial guess
(before Stokes_flow_Incomp.c:354), and the solve iterates to its own
criterion.

Architecture (cl_unet_forward): 4 levels of node-coincident grids, 16/32/48/
64 channels, 15 3x3x3 convolutions, additive skips.  The input and output
are described at chip_set_warm_net (include/citcom_hip.h): features of the
extrapolated guess in time, and the correction of that extrapolation.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time

import numpy as np

from . import deck as _deck
from .hip import Citcom
from .learn import _cl, _ok, _p

# (ci, co) of the 15 convolutions in layer order (cl_unet_forward)
LAYERS = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 48), (48, 48), (48, 64), (64, 64),
          (64, 48), (48, 48), (48, 32), (32, 32), (32, 16), (16, 16), (16, 16)]
NOUT = 4          # u, v, w, p corrections
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warm_unet.npz")


def _lib():
    L = _cl()
    if not getattr(L, "_unet_ready", False):
        P, i, l, f = C.c_void_p, C.c_int, C.c_long, C.c_float
        L.cl_conv3d_fwd_c.argtypes = [P, i, i, i, i, i, i, P, P, P, i, i, P, P, P]
        L.cl_conv3d_transpose_w.argtypes = [P, i, i, P, P]
        L.cl_conv3d_bwd_data_c.argtypes = [P, i, i, i, i, i, i, P, P, P]
        L.cl_conv3d_bwd_weight_c_workspace.argtypes = [i, i, i, i, i, i]
        L.cl_conv3d_bwd_weight_c_workspace.restype = l
        L.cl_conv3d_bwd_weight_c.argtypes = [P, P, i, i, i, i, i, i, P, P, P]
        L.cl_restrict.argtypes = [P, i, i, i, i, i, f, P, P]
        L.cl_interp.argtypes = [P, i, i, i, i, i, f, P, P]
        L.cl_add_bf16.argtypes = [P, P, l, P]
        L.cl_wmse.argtypes = [P, i, P, l, i, P, P, P, P]
        L.cl_unet_params.argtypes = [C.POINTER(l), C.POINTER(l)]
        L.cl_unet_workspace.argtypes = [i, i, i, i]
        L.cl_unet_workspace.restype = l
        L.cl_unet_forward.argtypes = [P, i, i, i, i, P, P, P, P, P]
        L._unet_ready = True
    return L


def level_dims(dims, lev):
    s = 1 << lev
    return tuple((n - 1) // s + 1 for n in dims)


def param_counts():
    return sum(co * 27 * ci for ci, co in LAYERS), sum(co for _, co in LAYERS)


class UNet:
    """Weights (FP32 masters + BF16 copies), Adam moments and the forward
    activations of one batch, all on the device."""

    def __init__(self, seed: int = 0):
        import torch
        g = torch.Generator().manual_seed(seed)
        self.w, self.b = [], []
        for k, (ci, co) in enumerate(LAYERS):
            # He initialisation over the 27 ci fan-in; the output layer small,
            # so training starts from the plain extrapolation
            std = math.sqrt(2.0 / (27 * ci)) * (0.1 if k == len(LAYERS) - 1 else 1.0)
            self.w.append((torch.randn(co, 27, ci, generator=g) * std).cuda())
            self.b.append(torch.zeros(co).cuda())
        self.wbf = [w.to(torch.bfloat16) for w in self.w]
        self.m = [torch.zeros_like(t) for t in self.w + self.b]
        self.v = [torch.zeros_like(t) for t in self.w + self.b]
        self.step = 0
        self.acts = None

    # ---- parameters as the library takes them
    def blob(self):
        import torch
        w = torch.cat([t.reshape(-1) for t in self.wbf]).view(torch.int16).cpu().numpy().view(np.uint16)
        b = torch.cat([t.reshape(-1) for t in self.b]).cpu().numpy().astype(np.float32)
        return w, b

    def load_blob(self, w: np.ndarray, b: np.ndarray):
        import torch
        ow = ob = 0
        for k, (ci, co) in enumerate(LAYERS):
            nw = co * 27 * ci
            wb = torch.from_numpy(np.ascontiguousarray(w[ow:ow + nw]).view(np.int16)).view(torch.bfloat16)
            self.wbf[k] = wb.reshape(co, 27, ci).cuda()
            self.w[k] = self.wbf[k].float()
            self.b[k] = torch.from_numpy(np.ascontiguousarray(b[ob:ob + co])).cuda()
            ow += nw
            ob += co

    # ---- forward with every activation kept (the order of cl_unet_forward)
    def _conv(self, k, x, dims, nb, res=None, relu=True, yf=None):
        import torch
        ci, co = LAYERS[k]
        y = torch.empty((nb, x.shape[1], co), dtype=torch.bfloat16, device="cuda")
        _ok(_lib().cl_conv3d_fwd_c(_p(x), nb, *dims, ci, co, _p(self.wbf[k]), _p(self.b[k]),
                                   _p(res) if res is not None else None, int(relu), 0, _p(y),
                                   _p(yf) if yf is not None else None, None), f"conv {k}")
        return y

    def _restrict(self, x, dims, nb, scale=1.0):
        import torch
        cd = level_dims(dims, 1)
        y = torch.empty((nb, cd[0] * cd[1] * cd[2], x.shape[2]), dtype=torch.bfloat16, device="cuda")
        _ok(_lib().cl_restrict(_p(x), nb, *dims, x.shape[2], scale, _p(y), None), "restrict")
        return y

    def _interp(self, x, fdims, nb, scale=1.0):
        import torch
        y = torch.empty((nb, fdims[0] * fdims[1] * fdims[2], x.shape[2]), dtype=torch.bfloat16, device="cuda")
        _ok(_lib().cl_interp(_p(x), nb, *fdims, x.shape[2], scale, _p(y), None), "interp")
        return y

    def forward(self, x, dims):
        """x: bf16 [nb][nv][16] on the finest of the four grids -> FP32
        [nb][nv][16]"""
        import torch
        nb = x.shape[0]
        d = [level_dims(dims, l) for l in range(4)]
        a = {"x0": x}
        a["y1"] = self._conv(0, x, d[0], nb)
        a["s0"] = self._conv(1, a["y1"], d[0], nb)
        a["r1"] = self._restrict(a["s0"], d[0], nb)
        a["y3"] = self._conv(2, a["r1"], d[1], nb)
        a["s1"] = self._conv(3, a["y3"], d[1], nb)
        a["r2"] = self._restrict(a["s1"], d[1], nb)
        a["y5"] = self._conv(4, a["r2"], d[2], nb)
        a["s2"] = self._conv(5, a["y5"], d[2], nb)
        a["r3"] = self._restrict(a["s2"], d[2], nb)
        a["y7"] = self._conv(6, a["r3"], d[3], nb)
        a["b3"] = self._conv(7, a["y7"], d[3], nb)
        a["i2"] = self._interp(a["b3"], d[2], nb)
        a["c2"] = self._conv(8, a["i2"], d[2], nb, res=a["s2"])
        a["d2"] = self._conv(9, a["c2"], d[2], nb)
        a["i1"] = self._interp(a["d2"], d[1], nb)
        a["c1"] = self._conv(10, a["i1"], d[1], nb, res=a["s1"])
        a["d1"] = self._conv(11, a["c1"], d[1], nb)
        a["i0"] = self._interp(a["d1"], d[0], nb)
        a["c0"] = self._conv(12, a["i0"], d[0], nb, res=a["s0"])
        a["d0"] = self._conv(13, a["c0"], d[0], nb)
        yf = torch.empty((nb, x.shape[1], 16), dtype=torch.float32, device="cuda")
        a["out"] = self._conv(14, a["d0"], d[0], nb, relu=False, yf=yf)
        self.acts, self.dims, self.nb = a, d, nb
        return yf

    # ---- backward: gradients of every weight and bias from d loss / d out
    def backward(self, g):
        import torch
        L = _lib()
        a, d, nb = self.acts, self.dims, self.nb
        grads = [None] * len(LAYERS)

        def relu_bwd(y, gy):
            gy = gy.clone()
            _ok(L.cl_relu_bwd(_p(y), _p(gy), gy.numel(), None), "relu_bwd")
            return gy

        def conv_bwd(k, x, gy, dims, need_dx=True):
            ci, co = LAYERS[k]
            ws = torch.empty(L.cl_conv3d_bwd_weight_c_workspace(nb, *dims, ci, co) // 4, dtype=torch.float32,
                             device="cuda")
            gr = torch.empty(co * 27 * ci + co, dtype=torch.float32, device="cuda")
            _ok(L.cl_conv3d_bwd_weight_c(_p(x), _p(gy), nb, *dims, ci, co, _p(gr), _p(ws), None), f"bwd_w {k}")
            grads[k] = gr
            if not need_dx:
                return None
            wt = torch.empty((ci, 27, co), dtype=torch.bfloat16, device="cuda")
            _ok(L.cl_conv3d_transpose_w(_p(self.wbf[k]), ci, co, _p(wt), None), "transpose")
            dx = torch.empty((nb, x.shape[1], ci), dtype=torch.bfloat16, device="cuda")
            _ok(L.cl_conv3d_bwd_data_c(_p(gy), nb, *dims, ci, co, _p(wt), _p(dx), None), f"bwd_d {k}")
            return dx

        def add(x, y):
            _ok(L.cl_add_bf16(_p(x), _p(y), x.numel(), None), "add")
            return x

        def interp_T(gfine, fdims):      # transpose of the interpolation: 8 x the restriction
            return self._restrict(gfine, fdims, nb, scale=8.0)

        def restrict_T(gcoarse, fdims):  # transpose of the restriction: the interpolation / 8
            return self._interp(gcoarse, fdims, nb, scale=0.125)

        gd0 = conv_bwd(14, a["d0"], g, d[0])
        gc0 = conv_bwd(13, a["c0"], relu_bwd(a["d0"], gd0), d[0])
        gp = relu_bwd(a["c0"], gc0)
        gi0 = conv_bwd(12, a["i0"], gp, d[0])
        gs0 = gp                                   # the skip s0 receives the same gradient
        gd1 = interp_T(gi0, d[0])
        gc1 = conv_bwd(11, a["c1"], relu_bwd(a["d1"], gd1), d[1])
        gp = relu_bwd(a["c1"], gc1)
        gi1 = conv_bwd(10, a["i1"], gp, d[1])
        gs1 = gp
        gd2 = interp_T(gi1, d[1])
        gc2 = conv_bwd(9, a["c2"], relu_bwd(a["d2"], gd2), d[2])
        gp = relu_bwd(a["c2"], gc2)
        gi2 = conv_bwd(8, a["i2"], gp, d[2])
        gs2 = gp
        gb3 = interp_T(gi2, d[2])
        gy7 = conv_bwd(7, a["y7"], relu_bwd(a["b3"], gb3), d[3])
        gr3 = conv_bwd(6, a["r3"], relu_bwd(a["y7"], gy7), d[3])
        gs2 = add(restrict_T(gr3, d[2]), gs2)
        gy5 = conv_bwd(5, a["y5"], relu_bwd(a["s2"], gs2), d[2])
        gr2 = conv_bwd(4, a["r2"], relu_bwd(a["y5"], gy5), d[2])
        gs1 = add(restrict_T(gr2, d[1]), gs1)
        gy3 = conv_bwd(3, a["y3"], relu_bwd(a["s1"], gs1), d[1])
        gr1 = conv_bwd(2, a["r1"], relu_bwd(a["y3"], gy3), d[1])
        gs0 = add(restrict_T(gr1, d[0]), gs0)
        gy1 = conv_bwd(1, a["y1"], relu_bwd(a["s0"], gs0), d[0])
        conv_bwd(0, a["x0"], relu_bwd(a["y1"], gy1), d[0], need_dx=False)
        return grads

    def loss_grad(self, yf, target, wch):
        """weighted MSE over the NOUT output channels; (loss, d loss / d out)"""
        import torch
        nvox = yf.shape[0] * yf.shape[1]
        dy = torch.empty(yf.shape, dtype=torch.bfloat16, device="cuda")
        part = torch.empty((nvox * 16 + 255) // 256, dtype=torch.float32, device="cuda")
        _ok(_lib().cl_wmse(_p(yf), 16, _p(target), nvox, NOUT, _p(wch), _p(dy), _p(part), None), "wmse")
        return float(part.double().sum()) / (nvox * NOUT), dy

    def adam(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = _lib()
        self.step += 1
        for k, (ci, co) in enumerate(LAYERS):
            nw = co * 27 * ci
            g = grads[k]
            _ok(L.cl_adam(_p(self.w[k]), _p(g), _p(self.m[k]), _p(self.v[k]), nw, lr, b1, b2, eps, self.step,
                          _p(self.wbf[k]), None, None), "adam w")
            gb = g[nw:]
            j = len(LAYERS) + k
            _ok(L.cl_adam(_p(self.b[k]), _p(gb), _p(self.m[j]), _p(self.v[j]), co, lr, b1, b2, eps, self.step,
                          None, None, None), "adam b")


# ---------------------------------------------------------------------------
# samples from the solver, training, evaluation

def problem_for(base: str, rayleigh: float, contrast: float) -> _deck.Problem:
    """the workload `base` at Rayleigh number `rayleigh` and viscosity
    contrast `contrast` (eta = exp(-ln(contrast) T))"""
    keys = _deck.workload(base)
    keys["rayleigh"] = rayleigh
    nm = int(keys.get("num_mat", 1))
    keys["viscE"] = ",".join([repr(math.log(contrast))] * nm)
    return _deck.problem_from_deck({k: str(v) for k, v in keys.items()})


def draw_cases(n: int, seed: int, ra=(1e5, 1e7), contrast=(1e2, 1e5)):
    """n (Rayleigh number, contrast, temperature seed) drawn log-uniformly"""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        r = 10 ** rng.uniform(math.log10(ra[0]), math.log10(ra[1]))
        c = 10 ** rng.uniform(math.log10(contrast[0]), math.log10(contrast[1]))
        out.append((float(r), float(c), int(rng.integers(1 << 30))))
    return out


def make_samples(base: str, cases, steps: int, mixed: int = 2, log=print):
    """Run each case's time loop with the extrapolated warm start (each solve
    to the workload's own criterion) and record every solve that has three
    earlier ones: features bf16 [N][nno][16], targets FP32 [N][nno][4], all
    on the device.  `steps` recorded samples per case."""
    import torch
    feats, targs, info = None, None, []
    n_total = len(cases) * steps
    k = 0
    for ci, (ra, contrast, seed) in enumerate(cases):
        p = problem_for(base, ra, contrast)
        nno = int(np.prod(p.dims()))
        if feats is None:
            feats = torch.empty((n_total, nno, 16), dtype=torch.bfloat16, device="cuda")
            targs = torch.empty((n_total, nno, NOUT), dtype=torch.float32, device="cuda")
        t0 = time.perf_counter()
        c = Citcom(p, mixed=mixed)
        c.set_option("warm_start", 1)
        c.initialize(_deck.seeded_temperature(p, seed=seed))
        c.general_stokes_solver()
        c.timestep()     # the history then holds three solves, the
        c.timestep()     # least the network's input needs
        its = []
        for s in range(steps):
            c.warm_record(feats[k].data_ptr(), targs[k].data_ptr())
            r = c.timestep()
            sc = c.warm_scales()
            its.append(r["stokes"]["iterations"])
            info.append({"case": ci, "rayleigh": ra, "contrast": contrast, "step": s + 3,
                         "iterations": r["stokes"]["iterations"], "scales": sc[:7].tolist()})
            k += 1
        c.warm_record(None, None)
        c.close()
        log(f"case {ci}: Ra {ra:.3g} contrast {contrast:.3g}: {steps} samples, Uzawa {min(its)}-{max(its)} "
            f"iterations, {time.perf_counter() - t0:.1f} s")
    return feats, targs, info


def train(net: UNet, feats, targs, dims, epochs: int, batch: int = 4, lr: float = 1e-3, wch=(1.0, 1.0, 1.0, 2.0),
          seed: int = 0, log=print, holdout=None, cosine: bool = True):
    """Adam on shuffled minibatches; with cosine the learning rate decays by
    cosine to lr/20 over the epochs, else it stays lr.  Returns the
    per-epoch mean losses (and held-out losses)."""
    import torch
    n = feats.shape[0]
    w = torch.tensor(wch, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(seed)
    steps_total = epochs * (n // batch)
    hist = []
    it = 0
    for ep in range(epochs):
        perm = rng.permutation(n)
        tot = 0.0
        t0 = time.perf_counter()
        for b0 in range(0, n - batch + 1, batch):
            idx = torch.from_numpy(perm[b0:b0 + batch]).cuda()
            x = feats.index_select(0, idx)
            y = targs.index_select(0, idx)
            out = net.forward(x, dims)
            loss, g = net.loss_grad(out, y, w)
            grads = net.backward(g)
            cur = lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * it / max(1, steps_total)))) if cosine else lr
            net.adam(grads, lr=cur)
            tot += loss
            it += 1
        torch.cuda.synchronize()
        rec = {"epoch": ep, "loss": tot / (n // batch), "s": time.perf_counter() - t0}
        if holdout is not None:
            rec["holdout"] = evaluate(net, holdout[0], holdout[1], dims, wch)
        hist.append(rec)
        log(f"epoch {ep}: loss {rec['loss']:.4f}" + (f" held-out {rec['holdout']:.4f}" if holdout else "") +
            f" ({rec['s']:.1f} s)")
    return hist


def evaluate(net: UNet, feats, targs, dims, wch=(1.0, 1.0, 1.0, 2.0), batch: int = 4):
    """mean weighted squared error of the network's correction against the
    targets (the extrapolation alone scores the mean weighted square of the
    targets)"""
    import torch
    w = torch.tensor(wch, dtype=torch.float32, device="cuda")
    tot, cnt = 0.0, 0
    for b0 in range(0, feats.shape[0], batch):
        x = feats[b0:b0 + batch]
        y = targs[b0:b0 + batch]
        out = net.forward(x, dims)
        loss, _ = net.loss_grad(out, y, w)
        tot += loss * x.shape[0]
        cnt += x.shape[0]
    return tot / cnt


def save(net: UNet, path: str, meta: dict) -> None:
    w, b = net.blob()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, weights=w, bias=b, meta=np.array(json.dumps(meta)))


def load(path: str = WEIGHTS_PATH):
    """(weights bf16 bits, biases, metadata) of a trained network"""
    with np.load(path, allow_pickle=False) as z:
        return z["weights"], z["bias"], json.loads(str(z["meta"]))


def install(c: Citcom, path: str = WEIGHTS_PATH) -> dict:
    """give a context the trained network and switch its warm start to the
    learned form; returns the network's metadata"""
    w, b, meta = load(path)
    c.set_warm_net(w, b)
    c.set_option("warm_start", 2)
    return meta
