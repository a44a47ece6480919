"""Readers for the raw arrays written by oracle/_ref/ref_harness and their
conversion to the 0-based layout of the oracle and the HIP path."""
from __future__ import annotations

import os
import re

import numpy as np

_SUFFIX = {"f32": np.float32, "f64": np.float64, "i32": np.int32, "u32": np.uint32}


def load(dirpath: str, name: str, lev: int | None = None) -> np.ndarray:
    fn = name if lev is None else f"{name}_lev{lev}"
    base = re.sub(r"\.r\d+$", "", name)   # per-rank dumps of a run under MPI: <name>.r<rank>
    dt = _SUFFIX[base.rsplit(".", 1)[1]]
    return np.fromfile(os.path.join(dirpath, fn), dtype=dt)


def full_rows_from_half(K1, K2, K3, MAP, nox, noy, noz):
    """Expand the reference's symmetric half storage (Eqn_k1..3 + Node_map,
    Construct_arrays.c:280-535) to full 27-neighbour rows [nno][3][27][3]."""
    nno = nox * noy * noz
    neq = 3 * nno
    full = np.zeros((nno, 3, 27, 3), dtype=np.float32)
    Kd = np.stack([K1[:nno * 42], K2[:nno * 42], K3[:nno * 42]], axis=0).reshape(3, nno, 42).transpose(1, 0, 2)
    M = MAP.reshape(nno, 42)
    n = np.arange(nno)
    k, i, j = n % noz, (n // noz) % nox, n // (noz * nox)
    for s in range(42):
        eq = M[:, s]
        ok = eq < neq
        m = eq[ok] // 3
        c = eq[ok] % 3
        nn = n[ok]
        mk, mi, mj = m % noz, (m // noz) % nox, m // (noz * nox)
        slot = (mk - k[ok] + 1) + 3 * (mi - i[ok] + 1) + 9 * (mj - j[ok] + 1)
        for r in range(3):
            full[nn, r, slot, c] = Kd[ok, r, s]
            # transpose into the row of the lower node m
            tslot = 26 - slot
            full[m, c, tslot, r] = Kd[ok, r, s]
    # a node's own block as n_assemble_del2_u and gauss_seidel read it: through
    # the column loop (Element_calculations.c:604-605), i.e. transposed
    for s in range(42):
        eq = M[:, s]
        self_ = (eq < neq) & (eq // 3 == n)
        a = eq[self_] % 3
        nn = n[self_]
        for b in range(3):
            full[nn, a, 13, b] = Kd[self_, b, s]
    return full.reshape(nno, 243)


# slots of the reference's per-element structs that it writes (the others
# hold whatever the allocation held): Shape_function_dA vpt[1..8], ppt[1];
# COORD centre[1..3], size[1..3], area (global_defs.h:240-244, 316-322)
GDA_USED = [1, 2, 3, 4, 5, 6, 7, 8, 10]
ECO_USED = [1, 2, 3, 5, 6, 7, 12]


class XorShift:
    """The seeded generator of ref_harness.c (xorshift64, uniform(-1,1))."""

    M = (1 << 64) - 1

    def __init__(self, seed: int):
        self.s = 88172645463325252 ^ ((seed * 0x9E3779B97F4A7C15) & self.M)

    def draw(self, n: int) -> np.ndarray:
        out = np.empty(n)
        s, M = self.s, self.M
        for i in range(n):
            s ^= (s << 13) & M
            s ^= s >> 7
            s ^= (s << 17) & M
            out[i] = ((s >> 11) / 9007199254740992.0) * 2.0 - 1.0
        self.s = s
        return out


def harness_op_inputs(problem, seed: int, strip):
    """Inputs ref_harness draws in ``ops`` mode, per level (finest first):
    u, F (stripped), d0 (stripped), p (one zero entry), coarse vector."""
    rng = XorShift(seed)
    out = {}
    for lev in range(problem.levels - 1, -1, -1):
        nox, noy, noz = problem.dims(lev)
        nno = nox * noy * noz
        nel = (nox - 1) * (noy - 1) * (noz - 1)
        u = rng.draw(3 * nno)
        F = strip(lev, rng.draw(3 * nno))
        d0 = strip(lev, rng.draw(3 * nno))
        p = rng.draw(nel)
        p[nel // 3] = 0.0
        c = None
        if lev > 0:
            cx, cy, cz = problem.dims(lev - 1)
            c = rng.draw(3 * cx * cy * cz)
        out[lev] = {"u": u, "F": F, "d0": d0, "p": p, "c": c}
    return out
