"""The box smoother of the FP32 cycles (mixed forms) against a NumPy
statement of it.

The box smoother cuts a level into boxes of at most 8 nodes a side; a sweep
runs the 8 colour passes of the node-block Gauss-Seidel update
(General_matrix_functions.c:1160-1361, in the 8-colour order of the
restatement) inside each box, where the colours before see their update,
and reads the nodes of the other boxes as they were before the sweep.  The
model (tests/box_model.py) states that on the oracle's stiffness (bitwise
the device's) in FP64; the device runs it in FP32, so the comparison is to
a relative 2e-5.
The same model with one box for the whole level is the plain 8-colour sweep,
which the device result must NOT match to that tolerance: the test would
otherwise not see whether the boxes are there.
"""
import numpy as np
import pytest

from box_model import sweep_model
from conftest import problem

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module", params=[1, 2])
def tdep64_mixed(request, pyoracle):
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    p = problem("tdep64")
    T = deck.perturbed_temperature(p)
    c = Citcom(p, mixed=request.param)
    c.initialize(T)
    c.get_system_viscosity()
    c.construct_stiffness_B_matrix()
    o = pyoracle.Oracle(p, smoother=1)
    o.set_viscosity(o.visc_from_T(T))
    yield request.param, p, c, o
    c.close()


@pytest.mark.parametrize("lev", [2, 3])
@pytest.mark.parametrize("guess,sweeps", [(0, 2), (0, 3), (1, 3), (1, 2)])
def test_box_sweeps_match_model(tdep64_mixed, lev, guess, sweeps):
    """mixed = 1: FP32 stiffness on both levels; mixed = 2: level 3 (the
    second finest of 65^3 nodes) has the BI-scaled FP16 rows"""
    mixed, p, c, o = tdep64_mixed
    half = mixed == 2 and lev == p.levels - 2
    if mixed == 2 and not half:
        pytest.skip("the FP32 level is covered by mixed = 1")
    nox, noy, noz = p.dims(lev)
    dims = (noz, nox, noy)
    nno = c.nno(lev)
    rng = np.random.default_rng(100 * lev + 10 * guess + sweeps)
    K = o.get("K", lev)
    BI = o.get("BI", lev)
    F = rng.standard_normal(3 * nno)
    d0 = rng.standard_normal(3 * nno) * 1e-2 if guess else np.zeros(3 * nno)
    say(f"box smoother mixed {mixed} level {lev} {dims} guess {guess} sweeps {sweeps}")
    d, Ad = c.apply(f"box{guess}", lev, d0 if guess else None, F, sweeps=sweeps)
    f32 = lambda v: np.asarray(v, dtype=np.float32).astype(np.float64)
    ref = f32(d0)
    plain = f32(d0)
    for _ in range(sweeps):
        ref = sweep_model(K, f32(BI), f32(F), ref, dims, boxes=True, half=half)
        plain = sweep_model(K, f32(BI), f32(F), plain, dims, boxes=False, half=half)
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(d - ref)) / scale
    err_plain = np.max(np.abs(d - plain)) / scale
    say(f"  against the box model {err:.2e}, against the plain 8-colour sweep {err_plain:.2e}")
    assert err < 2e-5
    assert err_plain > 100 * err
    # Ad = K d (the matvec of the level after the sweeps; the FP16 rows'
    # matvec is K d to their rounding)
    Kd = o.matvec(lev, d)
    assert np.max(np.abs(Ad - Kd)) / np.max(np.abs(Kd)) < (2e-3 if half else 2e-5)
