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
    ("iso16@1e-8", "iso16_at_1e-8_run.npz", 200, 1e-6, 1e-5),
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


PROGRAM_REF = os.path.join(ROOT, "oracle", "_ref", "citcom.mpi")
PROGRAM_HIP = os.path.join(ROOT, "oracle", "_ref", "citcom_hip.mpi")


def run_program(binary: str, name: str, steps: int, tmp: str, form: str = "plain") -> str:
    from citcom_amd import deck
    keys = deck.workload(name)
    out = os.path.join(tmp, "out")
    os.makedirs(out, exist_ok=True)
    keys.update({"maxstep": steps, "storage_spacing": 1, "datafile": os.path.join(out, "run")})
    path = os.path.join(tmp, "deck.input")
    deck.write_deck(path, keys)
    env = dict(os.environ, CITCOM_HIP_FORM=form)
    r = subprocess.run([binary, path], cwd=tmp, capture_output=True, text=True, timeout=600, env=env)
    # the reference ends a normal run through parallel_process_termination (exit 8)
    assert r.returncode in (0, 8), r.stderr[-2000:]
    return out


@pytest.mark.skipif(not (os.path.exists(PROGRAM_REF) and os.path.exists(PROGRAM_HIP)),
                    reason="oracle/_ref/citcom.mpi or citcom_hip.mpi not built")
@pytest.mark.parametrize("form", ["plain", "mixed16"])
def test_wrapped_program_outputs(form):
    """The reference program itself, pristine and with the GPU drop-in linked
    in (CITCOM_HIP_FORM = plain, mixed16), on the 16^3 deck at accuracy 1e-8
    for 20 steps: its own output files (.velo, .temp, .ave, written with 5-7
    significant digits) agree."""
    from citcom_amd import outputs
    with tempfile.TemporaryDirectory() as ta, tempfile.TemporaryDirectory() as tb:
        print("[gpu-test] pristine reference program, 20 steps", flush=True)
        da = run_program(PROGRAM_REF, "iso16@1e-8", 20, ta)
        print(f"[gpu-test] reference program with the GPU drop-in ({form}), 20 steps", flush=True)
        db = run_program(PROGRAM_HIP, "iso16@1e-8", 20, tb, form)
        va = outputs.read_velo(os.path.join(da, "run.velo.0.20"))
        vb = outputs.read_velo(os.path.join(db, "run.velo.0.20"))
        ta_ = outputs.read_temp(os.path.join(da, "run.temp.0.20"))
        tb_ = outputs.read_temp(os.path.join(db, "run.temp.0.20"))
        aa = [outputs.read_ave(os.path.join(da, f"run.ave.0.{k}")) for k in range(21)]
        ab = [outputs.read_ave(os.path.join(db, f"run.ave.0.{k}")) for k in range(21)]
    assert va.timesteps == vb.timesteps == 20
    ev, et = rel_l2(vb.values, va.values), rel_l2(tb_.values[:, 0], ta_.values[:, 0])
    dnu = max(abs(b.nu_top / a.nu_top - 1) for a, b in zip(aa, ab))
    dtime = abs(ab[-1].time / aa[-1].time - 1)
    dvrms = rel_l2(ab[-1].profiles[:, 1], aa[-1].profiles[:, 1])
    print(f"[gpu-test] {form}: program outputs at step 20: velo rel L2 {ev:.2e}, T {et:.2e}, max rel dNu_top {dnu:.2e}, "
          f"time {dtime:.2e}, vrms profile {dvrms:.2e}", flush=True)
    assert ev < 1e-5 and et < 1e-5 and dnu < 1e-5 and dtime < 1e-5 and dvrms < 1e-5
