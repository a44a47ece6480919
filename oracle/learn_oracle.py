"""CPU restatement of the learned warm start's network kernels (numpy).

TEST INFRASTRUCTURE ONLY: the checker for citcom_amd/csrc/learn_kernels.hip
(no counterpart in the reference).  It follows the same definitions:
activations [nb][nx*ny*nz][16] rounded to bf16 (round to nearest even) where
the kernels store bf16, weights [16 out][27 taps][16 in], tap =
(dz+1) + 3(dx+1) + 9(dy+1), node order k + nz*(i + nx*j), zero padding.
Sums run in float64 here; the kernels accumulate in FP32 on MFMA, so the two
agree to FP32 accumulation rounding plus at most one bf16 step where a
stored value lands near a rounding boundary.
"""
from __future__ import annotations

import numpy as np

CH = 16


def to_bf16(a: np.ndarray) -> np.ndarray:
    """float32 -> bf16 (as float32 values), round to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32)


def _grid(x, nx, ny, nz):
    # [nb][nv][CH] -> [nb][ny][nx][nz][CH]
    return x.reshape(x.shape[0], ny, nx, nz, CH)


def _shift(xg, dz, dx, dy):
    """out[..., j, i, k, :] = xg[..., j+dy, i+dx, k+dz, :] (zero outside)."""
    out = np.zeros_like(xg)
    nb, ny, nx, nz, _ = xg.shape

    def sl(d, n):
        return (slice(max(0, -d), n - max(0, d)), slice(max(0, d), n - max(0, -d)))
    (oy, iy), (ox, ix), (oz, iz) = sl(dy, ny), sl(dx, nx), sl(dz, nz)
    out[:, oy, ox, oz, :] = xg[:, iy, ix, iz, :]
    return out


def taps():
    for t in range(27):
        yield t, t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1


def conv_fwd(x, w, bias, relu, dims):
    """x [nb][nv][16] (bf16 values), w [16][27][16], bias [16] -> (y_bf16, y_f)"""
    nx, ny, nz = dims
    xg = _grid(x.astype(np.float64), nx, ny, nz)
    acc = np.zeros_like(xg)
    for t, dz, dx, dy in taps():
        acc += _shift(xg, dz, dx, dy) @ w[:, t, :].T.astype(np.float64)
    y = acc + bias.astype(np.float64)
    if relu:
        y = np.maximum(y, 0.0)
    yf = y.astype(np.float32).reshape(x.shape)
    return to_bf16(yf), yf


def conv_bwd_data(dy, w, dims):
    nx, ny, nz = dims
    dg = _grid(dy.astype(np.float64), nx, ny, nz)
    acc = np.zeros_like(dg)
    for t, dz, dx, dy_ in taps():
        # dX[v] += dY[v - off_t] W[:, t, :]
        acc += _shift(dg, -dz, -dx, -dy_) @ w[:, t, :].astype(np.float64)
    return to_bf16(acc.astype(np.float32).reshape(dy.shape))


def conv_bwd_weight(x, dy, dims):
    nx, ny, nz = dims
    xg = _grid(x.astype(np.float64), nx, ny, nz)
    dg = dy.astype(np.float64).reshape(-1, CH)
    dW = np.zeros((CH, 27, CH))
    for t, dz, dx, dy_ in taps():
        dW[:, t, :] = dg.T @ _shift(xg, dz, dx, dy_).reshape(-1, CH)
    db = dg.sum(axis=0)
    return dW.astype(np.float32), db.astype(np.float32)


def relu_bwd(y, dy):
    return np.where(y > 0, dy, 0.0).astype(np.float32)


def mse_grad(y, target, nout):
    nvox = y.shape[0] * y.shape[1]
    e = (y - target)[..., :nout].astype(np.float64)
    g = np.zeros_like(y, dtype=np.float64)
    g[..., :nout] = 2.0 * e / (nvox * nout)
    return float((e * e).sum()), to_bf16(g.astype(np.float32))


# ---- the U-Net's kernels (citcom_amd/csrc/unet_kernels.hip): any channel
# counts, activations [nb][nv][C], weights [co][27][ci]

# storage type of the U-Net functions' results: float32 as the kernels
# store them (then rounded to bf16); a check of the chain's calculus alone
# (finite differences) sets CAST = np.float64 and BF16 = False
CAST = np.float32
BF16 = True


def _bf(a):
    return to_bf16(a) if BF16 else np.asarray(a, dtype=CAST)


def _grid_c(x, nx, ny, nz):
    return x.reshape(x.shape[0], ny, nx, nz, x.shape[-1])


def conv_fwd_c(x, w, bias, res, relu, dims):
    """y = act(conv(x) + bias + res): x [nb][nv][ci] (bf16 values),
    w [co][27][ci], bias [co] or None, res [nb][nv][co] or None
    -> (y_bf16, y_f32)"""
    nx, ny, nz = dims
    xg = _grid_c(x.astype(np.float64), nx, ny, nz)
    co = w.shape[0]
    acc = np.zeros(xg.shape[:-1] + (co,))
    for t, dz, dx, dy in taps():
        acc += _shift(xg, dz, dx, dy) @ w[:, t, :].T.astype(np.float64)
    y = acc.reshape(x.shape[0], -1, co)
    if bias is not None:
        y = y + bias.astype(np.float64)
    if res is not None:
        y = y + res.astype(np.float64)
    if relu:
        y = np.maximum(y, 0.0)
    yf = y.astype(CAST)
    return _bf(yf), yf


def conv_bwd_data_c(dy, w, dims):
    """dX[v][c] = sum_t sum_o dY[v - off_t][o] W[o][t][c] -> bf16 [nb][nv][ci]"""
    nx, ny, nz = dims
    dg = _grid_c(dy.astype(np.float64), nx, ny, nz)
    acc = np.zeros(dg.shape[:-1] + (w.shape[2],))
    for t, dz, dx, dy_ in taps():
        acc += _shift(dg, -dz, -dx, -dy_) @ w[:, t, :].astype(np.float64)
    return _bf(acc.astype(CAST).reshape(dy.shape[0], -1, w.shape[2]))


def conv_bwd_weight_c(x, dy, dims):
    """dW [co][27][ci], db [co] (float32)"""
    nx, ny, nz = dims
    ci, co = x.shape[-1], dy.shape[-1]
    xg = _grid_c(x.astype(np.float64), nx, ny, nz)
    dg = dy.astype(np.float64).reshape(-1, co)
    dW = np.zeros((co, 27, ci))
    for t, dz, dx, dy_ in taps():
        dW[:, t, :] = dg.T @ _shift(xg, dz, dx, dy_).reshape(-1, ci)
    return dW.astype(CAST), dg.sum(axis=0).astype(CAST)


def coarse_dims(dims):
    return tuple((n - 1) // 2 + 1 for n in dims)


def restrict(x, dims, scale=1.0):
    """full weighting onto the node-coincident coarse grid: coarse[c] =
    scale * sum_d w(d) fine[2c+d], w = prod(1/2 at d=0, 1/4 at d=+-1),
    outside neighbours left out -> bf16 [nb][nvc][C]"""
    nx, ny, nz = dims
    cx, cy, cz = coarse_dims(dims)
    xg = _grid_c(x.astype(np.float64), nx, ny, nz)
    out = np.zeros((x.shape[0], cy, cx, cz, x.shape[-1]))
    for _, dz, dx, dy in taps():
        w = (0.25 if dz else 0.5) * (0.25 if dx else 0.5) * (0.25 if dy else 0.5)
        sh = _shift(xg, dz, dx, dy)   # sh[j, i, k] = fine[j+dy, i+dx, k+dz]
        out += w * sh[:, ::2, ::2, ::2, :]
    return _bf((scale * out).astype(CAST).reshape(x.shape[0], -1, x.shape[-1]))


def interp(xc, dims, scale=1.0):
    """trilinear interpolation from the coarse grid to the fine grid of
    dims (fine[f] = scale * sum_c p(f-2c) coarse[c], p = prod(1 at 0,
    1/2 at +-1)) -> bf16 [nb][nv][C]"""
    nx, ny, nz = dims
    cx, cy, cz = coarse_dims(dims)
    cg = _grid_c(xc.astype(np.float64), cx, cy, cz)
    out = np.zeros((xc.shape[0], ny, nx, nz, xc.shape[-1]))
    out[:, ::2, ::2, ::2, :] = cg
    for ax in (1, 2, 3):   # fill odd indices along y, x, z in turn
        idx = [slice(None)] * 5
        lo, hi, mid = list(idx), list(idx), list(idx)
        lo[ax], hi[ax], mid[ax] = slice(0, -1, 2), slice(2, None, 2), slice(1, None, 2)
        out[tuple(mid)] = 0.5 * (out[tuple(lo)] + out[tuple(hi)])
    return _bf((scale * out).astype(CAST).reshape(xc.shape[0], -1, xc.shape[-1]))


def wmse_grad(y, ystride, target, nout, wch):
    """loss sum (not yet divided) and dY (bf16) of the weighted MSE over the
    first nout of ystride channels"""
    nvox = y.shape[0] * y.shape[1]
    e = (y[..., :nout] - target).astype(np.float64)
    g = np.zeros(y.shape, dtype=np.float64)
    g[..., :nout] = 2.0 * np.asarray(wch, np.float64) * e / (nvox * nout)
    return float((np.asarray(wch) * e * e).sum()), _bf(g.astype(CAST))


UNET_LAYERS = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 48), (48, 48), (48, 64), (64, 64),
               (64, 48), (48, 48), (48, 32), (32, 32), (32, 16), (16, 16), (16, 16)]


def unet_forward(x, W, b, dims):
    """the U-Net of the warm start (cl_unet_forward) with every activation
    kept: x bf16 values [nb][nv][16], W[k] [co][27][ci] (bf16 values),
    b[k] [co] -> (acts, out FP32)"""
    d = [dims]
    for _ in range(3):
        d.append(coarse_dims(d[-1]))
    a = {"x0": x}

    def conv(k, xin, lev, res=None, relu=True):
        yb, yf = conv_fwd_c(xin, W[k], b[k], res, relu, d[lev])
        return yb, yf
    a["y1"], _ = conv(0, x, 0)
    a["s0"], _ = conv(1, a["y1"], 0)
    a["r1"] = restrict(a["s0"], d[0])
    a["y3"], _ = conv(2, a["r1"], 1)
    a["s1"], _ = conv(3, a["y3"], 1)
    a["r2"] = restrict(a["s1"], d[1])
    a["y5"], _ = conv(4, a["r2"], 2)
    a["s2"], _ = conv(5, a["y5"], 2)
    a["r3"] = restrict(a["s2"], d[2])
    a["y7"], _ = conv(6, a["r3"], 3)
    a["b3"], _ = conv(7, a["y7"], 3)
    a["i2"] = interp(a["b3"], d[2])
    a["c2"], _ = conv(8, a["i2"], 2, res=a["s2"])
    a["d2"], _ = conv(9, a["c2"], 2)
    a["i1"] = interp(a["d2"], d[1])
    a["c1"], _ = conv(10, a["i1"], 1, res=a["s1"])
    a["d1"], _ = conv(11, a["c1"], 1)
    a["i0"] = interp(a["d1"], d[0])
    a["c0"], _ = conv(12, a["i0"], 0, res=a["s0"])
    a["d0"], _ = conv(13, a["c0"], 0)
    _, out = conv(14, a["d0"], 0, relu=False)
    return a, out, d


def unet_backward(a, g, W, d):
    """gradients [dW_k, db_k] of every layer from d loss / d out (bf16
    values), with the kernels' bf16 rounding points"""
    grads = [None] * len(UNET_LAYERS)

    def relu_bwd(y, dy):
        return np.where(y > 0, dy, 0.0).astype(CAST)

    def cb(k, x, gy, lev, need=True):
        grads[k] = conv_bwd_weight_c(x, gy, d[lev])
        return conv_bwd_data_c(gy, W[k], d[lev]) if need else None

    def add(x, y):
        return _bf((x.astype(CAST) + y.astype(CAST)).astype(CAST))
    gd0 = cb(14, a["d0"], g, 0)
    gc0 = cb(13, a["c0"], relu_bwd(a["d0"], gd0), 0)
    gp = relu_bwd(a["c0"], gc0)
    gi0 = cb(12, a["i0"], gp, 0)
    gs0 = gp
    gd1 = restrict(gi0, d[0], 8.0)
    gc1 = cb(11, a["c1"], relu_bwd(a["d1"], gd1), 1)
    gp = relu_bwd(a["c1"], gc1)
    gi1 = cb(10, a["i1"], gp, 1)
    gs1 = gp
    gd2 = restrict(gi1, d[1], 8.0)
    gc2 = cb(9, a["c2"], relu_bwd(a["d2"], gd2), 2)
    gp = relu_bwd(a["c2"], gc2)
    gi2 = cb(8, a["i2"], gp, 2)
    gs2 = gp
    gb3 = restrict(gi2, d[2], 8.0)
    gy7 = cb(7, a["y7"], relu_bwd(a["b3"], gb3), 3)
    gr3 = cb(6, a["r3"], relu_bwd(a["y7"], gy7), 3)
    gs2 = add(interp(gr3, d[2], 0.125), gs2)
    gy5 = cb(5, a["y5"], relu_bwd(a["s2"], gs2), 2)
    gr2 = cb(4, a["r2"], relu_bwd(a["y5"], gy5), 2)
    gs1 = add(interp(gr2, d[1], 0.125), gs1)
    gy3 = cb(3, a["y3"], relu_bwd(a["s1"], gs1), 1)
    gr1 = cb(2, a["r1"], relu_bwd(a["y3"], gy3), 1)
    gs0 = add(interp(gr1, d[0], 0.125), gs0)
    gy1 = cb(1, a["y1"], relu_bwd(a["s0"], gs0), 0)
    cb(0, a["x0"], relu_bwd(a["y1"], gy1), 0, need=False)
    return grads
