"""Time the reference program with the GPU drop-in linked in.

    python citcom_amd/wrap/time_dropin.py --config tdep128 --steps 4 --form plain --warm previous

Runs oracle/_ref/citcom_hip.mpi (the unmodified reference objects linked
with citcom_wrap.c; ``make -C oracle wrap``) on the workload's deck with the
seeded initial temperature read through the reference's restart reader, on
one MPI rank, and prints the per-step wall times the program itself logs
("CPU = ... for step n", Citcom.c:155-158) with its Uzawa iteration counts
("after (n) pressure loops", the shim's log line).  Step 1's time includes
the program's setup; steps 2.. are single passes of the time loop (host
work of the reference driver plus the device's PG step and Stokes solve).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from citcom_amd import deck  # noqa: E402

PROGRAM = os.path.join(ROOT, "oracle", "_ref", "citcom_hip.mpi")


def run(config: str, steps: int, form: str, warm: str, seed: int = 1) -> dict:
    keys = deck.workload(config)
    p = deck.problem_from_deck({k: str(v) for k, v in keys.items()})
    with tempfile.TemporaryDirectory() as tmp:
        deck.write_restart_temperature(os.path.join(tmp, "seed"), deck.seeded_temperature(p, seed=seed))
        keys.update({"datafile": os.path.join(tmp, "run"), "restart": 1, "restart_timesteps": 0,
                     "oldfile": os.path.join(tmp, "seed"), "maxstep": steps, "storage_spacing": 1000})
        path = os.path.join(tmp, "deck.input")
        deck.write_deck(path, keys)
        env = dict(os.environ, CITCOM_HIP_FORM=form, CITCOM_HIP_WARM=warm)
        r = subprocess.run([PROGRAM, path], cwd=tmp, capture_output=True, text=True, timeout=1800, env=env)
        log = open(os.path.join(tmp, "run.log0")).read()
    # a normal end of the reference program is exit code 8 (parallel_process_termination)
    if r.returncode not in (0, 8):
        raise RuntimeError(f"citcom_hip.mpi failed ({r.returncode}): {r.stderr[-1500:]}")
    times = {int(s): float(t) for t, s in re.findall(r"& CPU = ([0-9.eE+-]+) for step (\d+)", log)}
    its = [int(n) for n in re.findall(r"after \((\d+)\) pressure loops", r.stderr)]
    return {"config": config, "form": form, "warm": warm, "step_s": times, "uzawa_iterations": its}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="tdep128")
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--form", default="plain", choices=("plain", "mixed", "mixed16"))
    ap.add_argument("--warm", default="previous", choices=("previous", "extrap"))
    args = ap.parse_args()
    out = run(args.config, args.steps, args.form, args.warm)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
