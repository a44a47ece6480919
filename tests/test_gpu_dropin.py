"""The drop-in itself: the unmodified reference objects linked with the wrap
shim (citcom_amd/wrap/citcom_wrap.c) through GNU ld --wrap, so that the
reference's own driver calls the GPU for solve_constrained_flow_iterative and
PG_timestep.  The run is compared with the pristine reference's run of the
same deck (tests/golden/*_run.npz): per-step dt and Nu, final T, U, P.

The binary is oracle/_ref/ref_harness_hip (``make -C oracle wrap``); the test
is skipped where it was not built.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, golden, rel_l2
import refdump

pytestmark = pytest.mark.gpu

HARNESS_HIP = os.path.join(ROOT, "oracle", "_ref", "ref_harness_hip")


def run_wrapped(name: str, steps: int, tmp: str) -> str:
    from citcom_amd import deck
    keys = deck.workload(name)
    keys["datafile"] = os.path.join(tmp, "out", "run")
    keys["oldfile"] = keys["datafile"]
    os.makedirs(os.path.join(tmp, "out"), exist_ok=True)
    path = os.path.join(tmp, "deck.input")
    deck.write_deck(path, keys)
    out = os.path.join(tmp, "dump")
    print(f"[gpu-test] drop-in run {name} {steps} steps", flush=True)
    r = subprocess.run([HARNESS_HIP, path, out, "run", str(steps)], cwd=tmp, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


@pytest.mark.skipif(not os.path.exists(HARNESS_HIP), reason="oracle/_ref/ref_harness_hip not built")
@pytest.mark.parametrize("name,fixture,steps,field_tol,nu_tol", [
    ("tiny_tdep@1e-9", "tiny_tdep_at_1e-9_run.npz", 3, 1e-6, 1e-6),
    ("cfg1@1e-8", "cfg1_at_1e-8_run.npz", 200, 1e-6, 1e-5),
])
def test_wrapped_reference_run(name, fixture, steps, field_tol, nu_tol):
    g = golden(fixture)
    with tempfile.TemporaryDirectory() as tmp:
        d = run_wrapped(name, steps, tmp)
        st = np.loadtxt(os.path.join(d, "steps.txt"), ndmin=2)
        fields = {k: refdump.load(d, f"final_{k}.{t}") for k, t in (("T", "f32"), ("U", "f64"), ("P", "f64"))}
    ref = g["steps"]
    assert st.shape == ref.shape
    ddt = np.max(np.abs(st[1:, 2] / ref[1:, 2] - 1))
    dnu = np.max(np.abs(st[:, 3] / ref[:, 3] - 1))
    errs = {k: rel_l2(v, g[f"final_{k}"]) for k, v in fields.items()}
    print(f"[gpu-test] drop-in {name}: max rel ddt {ddt:.2e}, max rel dNu {dnu:.2e}, final rel L2 {errs}", flush=True)
    assert ddt < nu_tol and dnu < nu_tol
    for k, e in errs.items():
        assert e < field_tol, k
