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
