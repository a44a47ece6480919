"""Run two precision forms side by side along their own trajectories.

    python -m citcom_amd.compare_forms --config tdep64 --steps 500 --forms plain,mixed --every 50

Both forms start from the same seeded temperature (the workload's restart
noise), run the step-0 Stokes solve and then --steps timesteps of the time
loop (Citcom.c:111-161: PG_timestep, heat_flux, general_stokes_solver), each
on its own trajectory.  Every step records dt, Nu_top, Uzawa iterations and
multigrid calls; every --every steps the temperature, velocity and pressure
of the two forms are compared (relative L2, the pressure up to its constant)
with Vrms.  Prints one line per checkpoint and a JSON summary at the end.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import deck
from .hip import Citcom

FORMS = {"plain": 0, "mixed": 1, "mixed16": 2}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="tdep64")
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--forms", default="plain,mixed")
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    names = a.forms.split(",")
    if len(names) != 2 or any(n not in FORMS for n in names):
        ap.error(f"--forms takes two of {sorted(FORMS)}")
    p = deck.problem_from_deck({k: str(v) for k, v in deck.workload(a.config).items()})
    T0 = deck.seeded_temperature(p, seed=a.seed)
    ctx = [Citcom(p, mixed=FORMS[n]) for n in names]
    rec = {n: {"dt": [], "nu": [], "its": [], "mg": [], "s": 0.0} for n in names}
    for n, c in zip(names, ctx):
        c.initialize(T0)
        c.general_stokes_solver()
    checkpoints = []
    for step in range(1, a.steps + 1):
        for n, c in zip(names, ctx):
            t0 = time.perf_counter()
            r = c.timestep()
            rec[n]["s"] += time.perf_counter() - t0
            rec[n]["dt"].append(r["pg"]["timestep"])
            rec[n]["nu"].append(r["Nu_top"])
            rec[n]["its"].append(r["stokes"]["iterations"])
            rec[n]["mg"].append(r["stokes"]["mg_cycles"])
        if step % a.every == 0 or step == a.steps:
            f = [{"T": c.download("T"), "U": c.download("U"), "P": c.download("P"), "vrms": c.vrms()} for c in ctx]
            b, m = f[0], f[1]
            cp = {"step": step, "time": float(np.sum(rec[names[0]]["dt"])),
                  "T": rel(m["T"], b["T"]), "U": rel(m["U"], b["U"]),
                  "P": rel(m["P"] - m["P"].mean(), b["P"] - b["P"].mean()),
                  "vrms": abs(m["vrms"] / b["vrms"] - 1), "vrms_" + names[0]: b["vrms"],
                  "Nu_top_" + names[0]: rec[names[0]]["nu"][-1],
                  "Nu_top": abs(rec[names[1]]["nu"][-1] / rec[names[0]]["nu"][-1] - 1)}
            checkpoints.append(cp)
            print(f"step {step}: {names[1]} against {names[0]}: T {cp['T']:.2e} U {cp['U']:.2e} P {cp['P']:.2e} "
                  f"vrms {cp['vrms']:.2e} Nu_top {cp['Nu_top']:.2e} (vrms {b['vrms']:.4g}, Nu_top "
                  f"{cp['Nu_top_' + names[0]]:.5g}); seconds so far {rec[names[0]]['s']:.1f} / "
                  f"{rec[names[1]]['s']:.1f}", flush=True)
    for c in ctx:
        c.close()
    dt0, dt1 = np.array(rec[names[0]]["dt"]), np.array(rec[names[1]]["dt"])
    nu0, nu1 = np.array(rec[names[0]]["nu"]), np.array(rec[names[1]]["nu"])
    out = {"config": a.config, "steps": a.steps, "forms": names, "seed": a.seed,
           "max_rel_dt": float(np.max(np.abs(dt1 / dt0 - 1))),
           "max_rel_nu_top": float(np.max(np.abs(nu1 / nu0 - 1))),
           "checkpoints": checkpoints,
           "per_form": {n: {"s_per_step": rec[n]["s"] / a.steps, "uzawa_iterations": int(np.sum(rec[n]["its"])),
                            "mg_calls": int(np.sum(rec[n]["mg"])),
                            "steps_at_piterations_cap": int(np.sum(np.array(rec[n]["its"]) >= p.p_iterations))}
                        for n in names}}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
