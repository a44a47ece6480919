"""Timestep benchmark of the CitcomCU hot path on one MI355X.

One "step" is one pass of the reference's time loop (Citcom.c:111-161) with
every stage on the GPU: PG_timestep (SUPG energy step), heat_flux, then
general_stokes_solver (buoyancy forces, temperature-dependent viscosity,
stiffness rebuild on all multigrid levels, Uzawa CG around the full
multigrid velocity solve to accuracy 1e-6).

Default workload: 128x128x64 elements in a 2x2x1 box, 6 multigrid levels,
Ra = 1e7, eta = exp(-ln(1e4) T) clipped to [1e-4, 1], seeded initial
temperature.  The step-0 Stokes solve and the warmup steps are untimed.
The measured form is the mixed one (FP32 multigrid cycles inside FP64
iterative refinement, to the same 1e-6 criterion); the plain FP64 form is
timed beside it for the ratio.

Prints one JSON line: s/timestep (lower is better), the roofline of the
dominant kernel (the 8-colour Gauss-Seidel pass) from HIP events on the
library's stream, and a CPU baseline: the unmodified reference program under
MPI on up to 16 host cores (or, where MPI or that build is absent, the
reference's objects on one core, else the C restatement) timed on a bounded
sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from citcom_amd import deck  # noqa: E402

PEAK_HBM_GBS = 8000.0
COPY_PROBE_BYTES = 256 << 20
METRIC = "s/timestep at 128x128x64 elems: FP32+learned vs FP64 GPU, same 1e-6 residual"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_problem(name):
    return deck.problem_from_deck({k: str(v) for k, v in deck.workload(name).items()})


REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
REF_PROGRAM = os.path.join(ROOT, "oracle", "_ref", "citcom.mpi")


def cpu_baseline_reference(name, seed, mg_calls_per_step):
    """Time the unmodified reference (oracle/_ref/ref_harness, the
    reference's own objects) on one host thread on a bounded sample of the
    same workload: the operator build of one step, one multi_grid call and
    one PG_timestep on the seeded initial temperature (read through the
    reference's restart reader).  The per-step estimate scales the multi_grid
    time by the multigrid calls per step that the GPU run measured."""
    import subprocess
    import tempfile

    keys = deck.workload(name)
    p = make_problem(name)
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "out")
        keys["datafile"] = os.path.join(out, "run")
        keys.update({"restart": 1, "restart_timesteps": 0, "oldfile": os.path.join(out, "seed")})
        deck.write_restart_temperature(os.path.join(out, "seed"), deck.seeded_temperature(p, seed=seed))
        path = os.path.join(tmp, "deck.input")
        deck.write_deck(path, keys)
        r = subprocess.run([REF_HARNESS, path, os.path.join(tmp, "dump"), "time"], capture_output=True, text=True,
                           timeout=600, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError(f"reference timing sample failed: {r.stderr[-500:]}")
    f = r.stdout.split()
    t = {f[k]: float(f[k + 1]) for k in range(0, len(f) - 1, 2) if f[k] in ("build", "mg", "pg")}
    est = t["build"] + t["pg"] + t["mg"] * mg_calls_per_step
    return {"value": est, "unit": "s/timestep", "cores": 1, "kind": "reference",
            "sample": (f"unmodified reference (its own objects, gcc -O2, 1 MPI rank) at {name} size: operator build "
                       f"{t['build']:.2f} s + one PG_timestep {t['pg']:.2f} s + one multi_grid call {t['mg']:.2f} s x "
                       f"{mg_calls_per_step:.1f} multigrid calls per step (as counted on the GPU)")}


def cpu_baseline_port(name, seed, mg_calls_per_step):
    """The same sample on the C restatement (oracle/), where the reference
    build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    p = make_problem(name)
    o = pyoracle.Oracle(p, smoother=0)
    T = deck.seeded_temperature(p, seed=seed)
    nno = o.nno()
    t0 = time.perf_counter()
    o.set_viscosity(o.visc_from_T(T))
    t_rebuild = time.perf_counter() - t0
    F = o.assemble_forces(o.thermal_buoyancy(T))
    t0 = time.perf_counter()
    o.pg_timestep(T, np.zeros(nno, np.float32), np.zeros(3 * nno, np.float32))
    t_pg = time.perf_counter() - t0
    t0 = time.perf_counter()
    Fc = F.copy()
    d = np.zeros_like(F)
    pyoracle.lib().orc_multi_grid(o.h, d, Fc, None, 0.0, o.levmax)
    t_mg = time.perf_counter() - t0
    est = t_rebuild + t_pg + t_mg * mg_calls_per_step
    return {"value": est, "unit": "s/timestep", "cores": 1, "kind": "port",
            "sample": (f"C restatement on 1 host thread at {name} size: stiffness rebuild {t_rebuild:.2f} s + "
                       f"one PG step {t_pg:.2f} s + one multi_grid call {t_mg:.2f} s x "
                       f"{mg_calls_per_step:.1f} multigrid calls per step (as counted on the GPU)")}


def measured_traffic(kernel, config, field="hbm_bytes_per_launch"):
    """HBM bytes per launch of `kernel` on workload `config` (all its colour
    passes, or with field "finest_level_hbm_bytes_per_launch" the finest
    level's) from the committed counter summary
    (profiles/k_gs_colour_traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE
    runs, scaled by what the same counters read for a copy of known size), or
    None where it was not measured."""
    path = os.path.join(ROOT, "profiles", "k_gs_colour_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        rec = json.load(fh)
    per_kernel = rec.get(field)
    if rec.get("config") != config or not isinstance(per_kernel, dict):
        return None
    return per_kernel.get(kernel)


def mpi_decomposition(p, cap):
    """Largest rank count <= cap that the reference accepts for this grid
    (mgunit{x,y,z} % nproc{x,y,z} == 0), least subdomain surface first."""
    best = None
    mx, my, mz = p.mgunit
    for px in range(1, mx + 1):
        for py in range(1, my + 1):
            for pz in range(1, mz + 1):
                if mx % px or my % py or mz % pz or px * py * pz > cap:
                    continue
                nox, noy, noz = p.dims()
                ex, ey, ez = (nox - 1) // px, (noy - 1) // py, (noz - 1) // pz
                key = (px * py * pz, -(ex * ey + ey * ez + ex * ez))
                if best is None or key > best[0]:
                    best = (key, (px, py, pz))
    return best[1]


def cpu_baseline_mpi(name, seed, mg_calls_per_step, cap=16, iterations=(4, 14)):
    """Time the pristine reference program (oracle/_ref/citcom.mpi, built
    from the unmodified sources) under MPI on the host cores, on the same
    grid, deck and seeded temperature (one restart file per rank).

    Two short runs (maxstep 2, piterations = iterations[0], iterations[1])
    give step 2's wall time t and multigrid calls m from the reference's own
    log (Citcom.c:151-158; General_matrix_functions.c:494-498); the fit
    t = a + b m separates the per-step fixed work (PG_timestep, forces,
    viscosity, stiffness rebuild, heat flux) from the cost of one multigrid
    call, and the estimate is a + b x the multigrid calls per step counted on
    the GPU."""
    import re
    import shutil
    import subprocess
    import tempfile

    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not (os.path.exists(REF_PROGRAM) and os.path.exists(mpiexec)):
        raise RuntimeError("no reference program or mpiexec")
    p = make_problem(name)
    try:
        cap = min(cap, len(os.sched_getaffinity(0)))
    except AttributeError:
        cap = min(cap, os.cpu_count() or 1)
    nproc = mpi_decomposition(p, cap)
    ranks = nproc[0] * nproc[1] * nproc[2]
    T = deck.seeded_temperature(p, seed=seed)
    samples = []
    with tempfile.TemporaryDirectory() as tmp:
        deck.write_restart_temperature(os.path.join(tmp, "seed"), T, p.dims(), nproc)
        for its in iterations:
            keys = deck.workload(name)
            out = os.path.join(tmp, f"run{its}")
            os.makedirs(out)
            keys.update({"datafile": os.path.join(out, "run"), "restart": 1, "restart_timesteps": 0,
                         "oldfile": os.path.join(tmp, "seed"), "maxstep": 2, "piterations": its,
                         "nprocx": nproc[0], "nprocy": nproc[1], "nprocz": nproc[2]})
            path = os.path.join(out, "deck.input")
            deck.write_deck(path, keys)
            t0 = time.perf_counter()
            r = subprocess.run([mpiexec, "-launcher", "fork", "-n", str(ranks), REF_PROGRAM, path],
                               capture_output=True, text=True, timeout=900, cwd=out)
            text = open(os.path.join(out, "run.log0")).read()
            if "for step 2" not in text:
                raise RuntimeError(f"reference run failed (rc {r.returncode}): {r.stderr[-400:]}")
            step2 = text[text.index("for step 1"):text.index("for step 2")]
            t = float(re.findall(r"& CPU = ([0-9.eE+-]+) for step 2", text)[0])
            m = sum(int(a) for a in re.findall(r"residual \((\d+)\)\(\d+\)", step2))
            samples.append((its, t, m))
            log(f"reference under MPI ({ranks} ranks {nproc}): piterations {its}: step 2 {t:.2f} s, {m} multigrid "
                f"calls (run {time.perf_counter() - t0:.0f} s)")
    (_, t1, m1), (_, t2, m2) = samples
    b = (t2 - t1) / (m2 - m1)
    a = t1 - b * m1
    est = a + b * mg_calls_per_step
    return {"value": est, "unit": "s/timestep", "cores": ranks, "kind": "reference",
            "sample": (f"pristine reference program (oracle/_ref/citcom.mpi, gcc -O2), {ranks} MPI ranks "
                       f"(nprocx,y,z = {nproc[0]},{nproc[1]},{nproc[2]}) at {name} size on the seeded temperature: "
                       f"step 2 of two runs with piterations {iterations[0]} and {iterations[1]} "
                       f"({t1:.2f} s / {m1} multigrid calls, {t2:.2f} s / {m2}); fit: {a:.2f} s per step + "
                       f"{b:.3f} s per multigrid call x {mg_calls_per_step:.1f} multigrid calls per step "
                       f"(as counted on the GPU)")}


def cpu_baseline(name, seed, mg_calls_per_step):
    one_rank = None
    if os.path.exists(REF_HARNESS):
        try:
            one_rank = cpu_baseline_reference(name, seed, mg_calls_per_step)
        except Exception as e:  # noqa: BLE001 - report and fall back to the restatement
            log(f"reference timing sample unavailable ({e})")
    try:
        out = cpu_baseline_mpi(name, seed, mg_calls_per_step)
        if one_rank:
            out["sample"] += f"; one rank: {one_rank['value']:.1f} s/timestep ({one_rank['sample']})"
        return out
    except Exception as e:  # noqa: BLE001 - report and fall back
        log(f"reference MPI timing unavailable ({e})")
    if one_rank:
        return one_rank
    log("timing the C restatement")
    return cpu_baseline_port(name, seed, mg_calls_per_step)


def run_form(p, T, mixed, steps, warmup, graphs, profile, coarse_net=None):
    """Timestep loop of one form: step-0 solve and warmup untimed, then `steps`
    timed steps; with profile, one more step with kernel-stamped events on
    every colour pass and one on every matvec.  coarse_net: a trained
    learned coarse-level correction to use in every V-cycle."""
    import torch

    from citcom_amd.hip import Citcom

    form = {0: "plain", 1: "mixed", 2: "mixed16"}[int(mixed)] + ("+learned" if coarse_net is not None else "")
    t0 = time.perf_counter()
    c = Citcom(p, mixed=mixed)
    if not graphs:
        c.set_option("graphs", 0)
    c.initialize(T)
    if coarse_net is not None:
        from citcom_amd import learn
        learn.install_coarse_net(coarse_net, c)
    log(f"[{form}] setup {time.perf_counter() - t0:.1f} s, device memory {c.device_bytes() / 2**30:.2f} GiB")
    t0 = time.perf_counter()
    st0 = c.general_stokes_solver()
    log(f"[{form}] step-0 Stokes solve {time.perf_counter() - t0:.2f} s, {st0['iterations']} Uzawa iterations")
    for w in range(warmup):
        r = c.timestep()
        log(f"[{form}] warmup {w}: {r['stokes']['iterations']} Uzawa its, {r['stokes']['mg_cycles']} MG calls")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = [c.timestep() for _ in range(steps)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for k, r in enumerate(recs):
        log(f"[{form}] step {k}: dt {r['pg']['timestep']:.3e} Nu_top {r['Nu_top']:.5f} "
            f"Uzawa {r['stokes']['iterations']} MG {r['stokes']['mg_cycles']}")
    if coarse_net is not None:
        still_on, fallbacks = c.coarse_net_state()
        log(f"[{form}] learned coarse correction {'in use' if still_on else 'switched off'}; "
            f"{fallbacks} inner solve(s) fell back to the plain multigrid")
    out = {"s_per_step": elapsed / steps,
           "uzawa_iterations_per_step": float(np.mean([r["stokes"]["iterations"] for r in recs])),
           "mg_calls_per_step": float(np.mean([r["stokes"]["mg_cycles"] for r in recs])),
           "Nu_top": float(recs[-1]["Nu_top"])}
    if coarse_net is not None:
        out["coarse_net_in_use"], out["fallbacks"] = still_on, fallbacks
    if profile:
        c.profile("gs_colour")
        c.timestep()
        out["gs"] = c.profile_read()
        c.profile("matvec")
        c.timestep()
        out["mv"] = c.profile_read()
        # the finest level alone: its colour passes are the bandwidth-bound ones
        c.set_option("profile_level", p.levmax)
        c.profile("gs_colour")
        c.timestep()
        out["gs_finest"] = c.profile_read()
        c.set_option("profile_level", -1)
        c.profile(None)
    c.close()
    return out


def learned_form(p, T, args):
    """FP32 cycles with the learned coarse-level correction: train it on the
    workload family's coarsest-level problems (random temperature fields
    under a three-level mesh on the same coarse grid, exact coarse solves),
    then time the mixed form with it in every V-cycle"""
    from citcom_amd import learn
    t0 = time.perf_counter()
    levels = min(3, p.levels)
    train_set = learn.make_coarse_samples(args.config, levels, 48, 16, seed=args.seed)
    test_set = learn.make_coarse_samples(args.config, levels, 6, 8, seed=args.seed + 1)
    net, losses = learn.train_coarse(train_set, p.dims(0), epochs=400)
    err_net, err_smoother = learn.coarse_relative_errors(net, test_set, p.dims(0))
    t_train = time.perf_counter() - t0
    log(f"[mixed+learned] coarse correction trained on {len(train_set)} coarse problems in {t_train:.1f} s: "
        f"held-out error {err_net:.3f} (coarse smoother {err_smoother:.3f})")
    out = run_form(p, T, 1, args.compare_steps, 1, not args.no_graphs, profile=False, coarse_net=net)
    out.update(train_s=t_train, coarse_error=err_net, smoother_error=err_smoother, samples=len(train_set))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="tdep128")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--form", choices=("mixed", "plain", "mixed16"), default="mixed",
                    help="mixed: FP32 multigrid inside FP64 refinement; plain: the reference's FP64 arithmetic; "
                         "mixed16: mixed with the fine levels' stiffness stored in FP16, scaled by BI")
    ap.add_argument("--compare-steps", type=int, default=2,
                    help="timed steps of the plain FP64 form run beside the mixed form (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch multigrid kernels one by one")
    ap.add_argument("--piterations", type=int, default=None,
                    help="cap on Uzawa iterations (profiling runs only: changes the workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; the solver would shard by subdomain, "
                                   "but no multi-GPU figure is produced"}))
        sys.exit(2)

    import torch  # device sync for the timed region

    p = make_problem(args.config)
    if args.piterations is not None:
        p.p_iterations = args.piterations
    nox, noy, noz = p.dims()
    log(f"config {args.config}: {nox - 1}x{noy - 1}x{noz - 1} elements, {p.levels} levels")
    T = deck.seeded_temperature(p, seed=args.seed) if p.tdepv else deck.perturbed_temperature(p)
    mixed = {"plain": 0, "mixed": 1, "mixed16": 2}[args.form]
    main_run = run_form(p, T, mixed, args.steps, args.warmup, not args.no_graphs, profile=True)
    plain = fp16 = learned = None
    if mixed and args.compare_steps > 0:
        plain = run_form(p, T, 0, args.compare_steps, 1, not args.no_graphs, profile=False)
        if mixed == 1:
            fp16 = run_form(p, T, 2, args.compare_steps, 1, not args.no_graphs, profile=False)
            learned = learned_form(p, T, args)

    # known-size copy (256 MiB read + 256 MiB written): the reference point that
    # scales hardware byte counters (FETCH_SIZE / WRITE_SIZE) into HBM bytes
    src = torch.ones(COPY_PROBE_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    src.sum()
    torch.cuda.synchronize()
    # the practical bandwidth the same device reaches on plain streams (the
    # second denominator beside the 8 TB/s data-sheet peak): a copy (read +
    # write) and a read-only reduction over 256 MiB, timed on torch's stream
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    for _ in range(5):
        src.sum()
    e2.record()
    torch.cuda.synchronize()
    stream = {"copy_gbs": 5 * 2 * COPY_PROBE_BYTES / (e0.elapsed_time(e1) * 1e-3) / 1e9,
              "read_gbs": 5 * COPY_PROBE_BYTES / (e1.elapsed_time(e2) * 1e-3) / 1e9}
    log(f"device streams: copy {stream['copy_gbs']:.0f} GB/s, read {stream['read_gbs']:.0f} GB/s")
    del src, dst

    n_gs, ms_gs, bytes_gs = main_run["gs"]
    n_mv, ms_mv, bytes_mv = main_run["mv"]
    n_f, ms_f, bytes_f = main_run["gs_finest"]
    ach_f = (bytes_f / max(n_f, 1)) / ((ms_f / max(n_f, 1)) * 1e-3) / 1e9
    log(f"finest-level colour passes: {n_f} launches, {ms_f / max(n_f, 1) * 1e3:.1f} us each, {ach_f:.0f} GB/s")
    avg_ms = ms_gs / max(n_gs, 1)
    achieved = (bytes_gs / max(n_gs, 1)) / (avg_ms * 1e-3) / 1e9
    log(f"gs_colour: {n_gs} launches, {ms_gs:.1f} ms total, {achieved:.0f} GB/s; "
        f"matvec: {n_mv} launches, {ms_mv:.1f} ms, {(bytes_mv / max(ms_mv, 1e-9)) / 1e6:.0f} GB/s")
    s_per_step = main_run["s_per_step"]
    # <vector type, stiffness type>; mixed16 runs the FP16 variant on the two
    # finest levels and the FP32 one below, both averaged here
    kernel = {0: "k_gs_colour<double, float>", 1: "k_gs_colour<float, float>", 2: "k_gs_colour<float, _Float16>"}[mixed]
    visc = (f"viscosity exp(-{p.viscE[0]:.4g} T)" + (f", {p.num_mat} layers (visc0 {p.visc0[:p.num_mat]})"
                                                      if p.num_mat > 1 else "")) if p.tdepv else "isoviscous"
    config = {"workload": f"{args.config}: {nox - 1}x{noy - 1}x{noz - 1} elements, {p.levels} MG levels, "
                          f"Ra={p.rayleigh:g}, {visc}, accuracy={p.accuracy:g}",
              "form": {1: "FP32 multigrid cycles inside FP64 iterative refinement (FP64 residual, Uzawa CG, "
                          "coarsest level); FP32 stiffness as in the reference",
                       2: "FP32 multigrid cycles inside FP64 iterative refinement; the two finest levels' "
                          "stiffness rows scaled by BI and stored in FP16 for the cycles (the FP64 residual uses "
                          "the reference's FP32 stiffness)",
                       0: "plain: FP64 vectors, the reference's FP32 stiffness"}[mixed],
              "uzawa_iterations_per_step": main_run["uzawa_iterations_per_step"],
              "mg_calls_per_step": main_run["mg_calls_per_step"]}
    if plain is not None:
        config["plain_fp64_s_per_step"] = plain["s_per_step"]
        config["speedup_vs_plain_fp64"] = plain["s_per_step"] / s_per_step
        config["learned_warm_start"] = ("citcom_amd.learn (BF16 MFMA 3-D CNN) is the Uzawa initial guess where no "
                                        "previous solution exists; in this time loop the previous step's U, P are "
                                        "the initial guess, as in the reference, so it is not used here")
        log(f"plain FP64 form: {plain['s_per_step']:.3f} s/step; {args.form} form {s_per_step:.3f} s/step")
    if fp16 is not None:
        config["mixed16_s_per_step"] = fp16["s_per_step"]
        config["mixed16_speedup_vs_plain_fp64"] = plain["s_per_step"] / fp16["s_per_step"]
        config["mixed16_form"] = ("the same cycles with the two finest levels' stiffness rows scaled by BI and "
                                  "stored in FP16 (FP32 arithmetic; the FP64 residual keeps the reference's FP32 "
                                  "stiffness); same Uzawa iterations and multigrid calls")
        log(f"mixed16 form: {fp16['s_per_step']:.3f} s/step")
    if learned is not None:
        config["mixed_learned_s_per_step"] = learned["s_per_step"]
        config["mixed_learned_speedup_vs_plain_fp64"] = plain["s_per_step"] / learned["s_per_step"]
        config["mixed_learned_mg_calls_per_step"] = learned["mg_calls_per_step"]
        config["learned_coarse_correction"] = (
            f"FP32 cycles with a BF16-MFMA 3-D CNN in place of the coarsest level's smoother call in every "
            f"V-cycle, trained in {learned['train_s']:.0f} s (inside this run, untimed) on "
            f"{learned['samples']} exact coarse solves of the workload family; held-out error "
            f"{learned['coarse_error']:.3f} against {learned['smoother_error']:.3f} for the coarse smoother; "
            f"{'in use to the end' if learned['coarse_net_in_use'] else 'switched off by the plain-multigrid fallback'} "
            f"({learned['fallbacks']} inner solve(s) restarted without it); "
            f"{learned['s_per_step'] / s_per_step:.3f}x the mixed form's time per step")
        log(f"mixed+learned form: {learned['s_per_step']:.3f} s/step, {learned['mg_calls_per_step']:.0f} MG calls")
    finest = {"achieved": ach_f, "frac": ach_f / PEAK_HBM_GBS, "launches": n_f,
              "avg_launch_ms": ms_f / max(n_f, 1), "algorithmic_bytes_per_launch": bytes_f / max(n_f, 1),
              "traffic": measured_traffic(kernel, args.config, "finest_level_hbm_bytes_per_launch")}
    out = {
        "metric": METRIC,
        "value": s_per_step,
        "unit": "s/timestep",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * s_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": config,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic(kernel, args.config), "kernel": kernel,
                     "launches": n_gs, "avg_launch_ms": avg_ms, "finest_level": finest, "device_streams": stream},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(args.config, args.seed, main_run["mg_calls_per_step"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
