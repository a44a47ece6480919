"""The 16^3 isoviscous Ra=1e4 workload (200 timesteps) on the GPU against the
reference's own run (tests/golden/cfg1_run.npz): per-step Nusselt number and
timestep, and the final T, U, P fields."""
import time

import numpy as np
import pytest

from conftest import golden, problem, rel_l2

pytestmark = pytest.mark.gpu


def test_cfg1_200_steps():
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    g = golden("cfg1_run.npz")
    p = problem("cfg1")
    c = Citcom(p)
    T = deck.perturbed_temperature(p)
    assert np.array_equal(T, g["step0_T"])
    c.initialize(T)
    c.general_stokes_solver()
    assert rel_l2(c.download("U"), g["step0_U"]) < 1e-6
    steps = g["steps"]
    t0 = time.time()
    worst_nu = worst_dt = 0.0
    for step in range(1, 201):
        r = c.timestep(with_nu=False)
        nu = c.heat_flux()  # after the Stokes solve, as the golden run records it
        worst_dt = max(worst_dt, abs(r["pg"]["timestep"] / steps[step, 2] - 1))
        worst_nu = max(worst_nu, abs(nu[0] / steps[step, 3] - 1))
        if step % 50 == 0:
            print(f"[gpu-test] cfg1 step {step} Nu {nu[0]:.6f} (ref {steps[step, 3]:.6f}) "
                  f"{time.time() - t0:.1f} s", flush=True)
    eT = rel_l2(c.download("T"), g["final_T"])
    eU = rel_l2(c.download("U"), g["final_U"])
    eP = rel_l2(c.download("P"), g["final_P"])
    print(f"[gpu-test] cfg1 200 steps: max rel dNu {worst_nu:.2e} max rel ddt {worst_dt:.2e} "
          f"T {eT:.2e} U {eU:.2e} P {eP:.2e}", flush=True)
    assert worst_nu < 1e-5
    assert eT < 1e-6 and eU < 1e-6 and eP < 1e-6
    c.close()
