"""CPU checks of the box smoother's NumPy statement (tests/box_model.py),
which the GPU test of the box smoother compares the device with: the
partition's invariants, and with one box for the whole level the sweep is
the oracle's 8-colour Gauss-Seidel sweep (General_matrix_functions.c:
1160-1361, restated in oracle/citcom_oracle.c), so the model is pinned to
the restatement that the reference's golden vectors pin."""
import numpy as np
import pytest

from box_model import BOX, axis_split, box_ids, sweep_model
from conftest import problem


@pytest.mark.parametrize("n", [2, 5, 8, 9, 16, 17, 33, 65, 129, 257])
def test_partition(n):
    s = axis_split(n)
    seg = np.diff(s)
    assert s[0] == 0 and s[-1] == n
    assert seg.min() >= 1 and seg.max() <= BOX and seg.max() - seg.min() <= 1
    # at most 4 nodes of one parity per segment: at most 64 of a colour per box
    for a, b in zip(s[:-1], s[1:]):
        for par in (0, 1):
            assert sum(1 for k in range(a, b) if k % 2 == par) <= BOX // 2
    ids = box_ids(n)
    assert np.all(np.diff(ids) >= 0) and ids[-1] == len(seg) - 1


def test_one_box_is_the_colour_sweep(pyoracle):
    from citcom_amd import deck
    p = problem("tdep64")
    T = deck.perturbed_temperature(p)
    o = pyoracle.Oracle(p, smoother=1)
    o.set_viscosity(o.visc_from_T(T))
    lev = 2
    nox, noy, noz = p.dims(lev)
    dims = (noz, nox, noy)
    n3 = 3 * nox * noy * noz
    rng = np.random.default_rng(0)
    F = rng.standard_normal(n3)
    d0 = rng.standard_normal(n3) * 1e-2
    K, BI = o.get("K", lev), o.get("BI", lev)
    ref = o.gauss_seidel(lev, d0.copy(), F, 1, 1)[0]
    one = sweep_model(K, BI, F, d0, dims, boxes=False)
    boxed = sweep_model(K, BI, F, d0, dims, boxes=True)
    scale = np.max(np.abs(ref))
    # the oracle rounds each increment to FP32, as the reference's temp is
    assert np.max(np.abs(one - ref)) / scale < 1e-6
    # the boxes change the sweep (their faces see the previous sweep)
    assert np.max(np.abs(boxed - ref)) / scale > 1e-3
