"""Timestep benchmark of the CitcomCU hot path on one MI355X.

One "step" is one pass of the reference's time loop (Citcom.c:111-161) with
every stage on the GPU: PG_timestep (SUPG energy step), heat_flux, then
general_stokes_solver (buoyancy forces, temperature-dependent viscosity,
stiffness rebuild on all multigrid levels, Uzawa CG around the full
multigrid velocity solve to accuracy 1e-6).

Default workload: 128x128x64 elements in a 2x2x1 box, 6 multigrid levels,
Ra = 1e7, eta = exp(-ln(1e4) T) clipped to [1e-4, 1], seeded initial
temperature.  Every form runs its own time loop from the same seeded state:
the step-0 Stokes solve and `warmup` steps untimed, then `steps` timed
steps, the same window for every form.  The value is the FP32+learned form
(mixed16+learned: FP32 multigrid cycles inside FP64 iterative refinement,
the fine levels' stiffness stored in FP16, and each Uzawa solve started
from the learned U-Net's correction of the last two solutions' extrapolation,
citcom_amd/weights/warm_unet.npz); the plain FP64 form (the reference
method, previous-solution start) and the precision-only form are timed
beside it on the same window, every solve to the same criterion (dp/p and
dv/v below 1e-6, or the piterations cap, counted per form).

Prints one JSON line: s/timestep (lower is better), the roofline of the
dominant kernel (the 8-colour Gauss-Seidel pass) from HIP events on the
library's stream, and a CPU baseline: one complete timestep of the
unmodified reference program under MPI on up to 16 host cores (or, where MPI
or that build is absent, an estimate from the C restatement on one core).
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


def cpu_baseline_port(name, seed, mg_calls_per_step):
    """Where the reference build or MPI is absent: the C restatement (oracle/)
    on one host thread, an estimate from a bounded sample (stiffness rebuild,
    one PG step and one multi_grid call, the latter scaled by the multigrid
    calls per step of the plain form)."""
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
            "sample": (f"estimate: C restatement on 1 host thread at {name} size, stiffness rebuild {t_rebuild:.2f} s + "
                       f"one PG step {t_pg:.2f} s + one multi_grid call {t_mg:.2f} s x "
                       f"{mg_calls_per_step:.1f} multigrid calls per step of the plain GPU form")}


def reference_step_times(stderr_text, log_text):
    """One complete timestep of the reference, from its own output.

    solve_Ahat_p_fhat prints "after (N) pressure loops and X sec for step s"
    on stderr (Stokes_flow_Incomp.c:482-488), X counted by MPI_Wtime from the
    start of the run's first Uzawa solve, so X(step 1) - X(step 0) is one
    whole pass of the time loop (Citcom.c:111-161: heat flux and velocity
    output of step 0, PG_timestep, forces, viscosity, stiffness rebuild and
    the Uzawa solve of step 1).  The multigrid calls of that step are the
    counts of its inner solves, " * residual (count)(cycles)" in the run log
    (General_matrix_functions.c:494-498) between the two Uzawa summaries.
    Returns (step seconds, multigrid calls, Uzawa iterations)."""
    import re

    after = {int(s): (int(n), float(x)) for n, x, s in
             re.findall(r"after \((\d+)\) pressure loops and ([0-9.eE+-]+) sec for step (\d+)", stderr_text)}
    if 0 not in after or 1 not in after:
        raise RuntimeError("the reference printed no Uzawa summary for steps 0 and 1")
    i0 = log_text.index("pressure loops and")
    i1 = log_text.index("pressure loops and", i0 + 1)
    mg = sum(int(a) for a in re.findall(r"residual \((\d+)\)\(\d+\)", log_text[i0:i1]))
    return after[1][1] - after[0][1], mg, after[1][0]


def cpu_baseline_mpi(name, seed, cap=16):
    """Time one complete timestep of the pristine reference program
    (oracle/_ref/citcom.mpi, built from the unmodified sources) under MPI on
    the host cores: the same grid, deck and seeded temperature (one restart
    file per rank), run to step 1 at the workload's own piterations and
    accuracy; the step's wall time and multigrid calls come from the
    reference's own output (reference_step_times)."""
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
    with tempfile.TemporaryDirectory() as tmp:
        deck.write_restart_temperature(os.path.join(tmp, "seed"), T, p.dims(), nproc)
        keys = deck.workload(name)
        keys.update({"datafile": os.path.join(tmp, "run"), "restart": 1, "restart_timesteps": 0,
                     "oldfile": os.path.join(tmp, "seed"), "maxstep": 1,
                     "nprocx": nproc[0], "nprocy": nproc[1], "nprocz": nproc[2]})
        path = os.path.join(tmp, "deck.input")
        deck.write_deck(path, keys)
        t0 = time.perf_counter()
        r = subprocess.run([mpiexec, "-launcher", "fork", "-n", str(ranks), REF_PROGRAM, path],
                           capture_output=True, text=True, timeout=900, cwd=tmp)
        wall = time.perf_counter() - t0
        log_text = open(os.path.join(tmp, "run.log0")).read()
    step_s, mg, its = reference_step_times(r.stderr, log_text)
    log(f"reference under MPI ({ranks} ranks {nproc}): step 1 {step_s:.2f} s, {its} Uzawa iterations, {mg} multigrid "
        f"calls (run {wall:.0f} s)")
    return {"value": step_s, "unit": "s/timestep", "cores": ranks, "kind": "reference",
            "uzawa_iterations": its, "mg_calls": mg,
            "sample": (f"measured: one complete timestep (step 1: PG_timestep, forces, viscosity, stiffness rebuild, "
                       f"Uzawa to accuracy {p.accuracy:g} or {p.p_iterations} iterations) of the pristine reference "
                       f"program (oracle/_ref/citcom.mpi, gcc -O2) on {ranks} MPI ranks (nprocx,y,z = "
                       f"{nproc[0]},{nproc[1]},{nproc[2]}) at {name} size on the seeded temperature: {its} Uzawa "
                       f"iterations, {mg} multigrid calls; wall time from the reference's own MPI_Wtime output "
                       f"(run to step 1: {wall:.0f} s including setup and the step-0 solve)")}


def cpu_baseline(name, seed, mg_calls_per_step):
    try:
        return cpu_baseline_mpi(name, seed)
    except Exception as e:  # noqa: BLE001 - report and fall back
        log(f"reference MPI timing unavailable ({e}); timing the C restatement")
    return cpu_baseline_port(name, seed, mg_calls_per_step)


FORMS = {
    # name: (mixed, warm_start, description)
    "plain": (0, 0, "the reference method in FP64 arithmetic (FP64 vectors, the reference's FP32 stiffness), each "
                    "Uzawa solve started from the previous solution as the reference does"),
    "mixed": (1, 0, "FP32 multigrid cycles inside FP64 iterative refinement (FP64 residual and Uzawa CG); the "
                    "cycles solve a small level exactly where there is one (17x17x9 nodes at 128x128x64 elements; "
                    "its inverse factored by rocSOLVER) and smooth the levels between it and the finest by 8-node "
                    "boxes, one launch per sweep; previous-solution start"),
    "mixed16": (2, 0, "as mixed, the two finest levels' stiffness rows scaled by BI and stored in FP16 for the cycles "
                      "(FP32 arithmetic; the FP64 residual keeps the reference's FP32 stiffness), previous-solution "
                      "start"),
    "mixed16+extrap": (2, 1, "as mixed16, each Uzawa solve started from the last solutions extrapolated along the "
                             "sequence of solves: 2 U_n - U_(n-1), P_n + 3/4 (P_n - P_(n-1)) + 1/4 (P_(n-1) - "
                             "P_(n-2)) (no network)"),
    "mixed16+learned": (2, 2, "as mixed16+extrap, the extrapolation corrected by the learned U-Net (BF16 MFMA, "
                              "citcom_amd/weights/warm_unet.npz) before each Uzawa solve"),
    "plain+extrap": (0, 1, "the plain FP64 form started from the extrapolation of mixed16+extrap (no network)"),
}


def run_form(p, T, form, steps, warmup, graphs, profile=False, same_state=False):
    """Timestep loop of one form on its own trajectory from the seeded state:
    the step-0 solve and `warmup` steps untimed, then `steps` timed steps.
    profile: one more step per level and smoother kernel (colour passes,
    box sweeps) with kernel-stamped events on its launches, then one on
    every matvec.
    same_state: one more step solved twice from the same state (same
    temperature, same previous solutions): with this form's initial guess and
    with the previous solution as it is (the reference's start)."""
    import torch

    from citcom_amd.hip import Citcom

    mixed, warm, _ = FORMS[form]
    t0 = time.perf_counter()
    c = Citcom(p, mixed=mixed)
    if not graphs:
        c.set_option("graphs", 0)
    if warm == 2:
        from citcom_amd import unet
        meta = unet.install(c)
        log(f"[{form}] U-Net of {meta['params']} parameters trained on {meta['samples']} samples "
            f"({meta['data_s']:.0f} s of solves, {meta['train_s']:.0f} s of training)")
    elif warm:
        c.set_option("warm_start", warm)
    c.initialize(T)
    log(f"[{form}] setup {time.perf_counter() - t0:.1f} s, device memory {c.device_bytes() / 2**30:.2f} GiB")
    t0 = time.perf_counter()
    st0 = c.general_stokes_solver()
    log(f"[{form}] step-0 Stokes solve {time.perf_counter() - t0:.2f} s, {st0['iterations']} Uzawa iterations")
    for w in range(warmup):
        r = c.timestep()
        log(f"[{form}] warmup {w + 1}: Uzawa {r['stokes']['iterations']} MG {r['stokes']['mg_cycles']}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = [c.timestep() for _ in range(steps)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cap = p.p_iterations
    for k, r in enumerate(recs):
        s = r["stokes"]
        log(f"[{form}] step {warmup + 1 + k}: dt {r['pg']['timestep']:.3e} Nu_top {r['Nu_top']:.5f} "
            f"Uzawa {s['iterations']} MG {s['mg_cycles']} dp/p {s['dpressure']:.2e} dv/v {s['dvelocity']:.2e}")
    its = [r["stokes"]["iterations"] for r in recs]
    out = {"s_per_step": elapsed / steps,
           "uzawa_iterations": its,
           "uzawa_iterations_per_step": float(np.mean(its)),
           "mg_calls_per_step": float(np.mean([r["stokes"]["mg_cycles"] for r in recs])),
           # the solve stops when dp/p and dv/v fall below accuracy, or at the
           # piterations cap (Stokes_flow_Incomp.c:393), or with alpha = 0
           # when an inner solve has nothing to do (valid = 0, :423-426)
           "steps_stopped_at_cap": int(sum(1 for i in its if i >= cap)),
           "Nu_top": float(recs[-1]["Nu_top"]),
           "elapsed_time": float(recs[-1]["pg"]["elapsed_time"]),
           "vrms": c.vrms(),
           "T_final": c.download("T")}
    if same_state:
        out["same_state"] = same_state_solves(c, form)
    if profile:
        # colour passes level by level (one more step each), per multigrid
        # call of that step: the FP16-stiffness form runs two template
        # instantiations of the colour pass on different levels.  The exact
        # coarse level's factorization runs on the work stream here, so that
        # the passes measured do not share the device with it
        c.set_option("direct_async", 0)
        out["gs_levels"] = {}
        for kind in ("gs_colour", "gs_box"):
            for lev in range(p.levels):
                c.set_option("profile_level", lev)
                c.profile(kind)
                r = c.timestep()
                n, ms, b = c.profile_read()
                mg = max(r["stokes"]["mg_cycles"], 1)
                if n:
                    out["gs_levels"][(kind, lev)] = {"launches": n / mg, "ms": ms / mg, "bytes": b / mg}
        c.set_option("profile_level", -1)
        c.profile("matvec")
        c.timestep()
        out["mv"] = c.profile_read()
        c.profile(None)
    c.close()
    return out


def same_state_solves(c, form):
    """Two more steps of this form's trajectory; the second one's solve is
    repeated from the same state (same temperature, same previous solutions)
    with other initial guesses: the form's own, the plain extrapolation
    (2 U_n - U_(n-1), P_n + 3/4 (P_n - P_(n-1)) + 1/4 (P_(n-1) - P_(n-2))),
    and the previous solution as it is (the reference's start), in the same
    arithmetic.  The form's own solve is the first, so
    the history it reads is intact."""
    import torch

    Ua, Pa = c.download("U"), c.download("P")
    c.timestep()
    Un, Pn = c.download("U"), c.download("P")
    c.PG_timestep()
    T = c.download("T")
    out = {}

    def timed(tag):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = c.general_stokes_solver()
        torch.cuda.synchronize()
        out[tag] = {"uzawa_iterations": st["iterations"], "mg_calls": st["mg_cycles"], "s": time.perf_counter() - t0,
                    "dpressure": st["dpressure"], "dvelocity": st["dvelocity"]}
        return c.download("U"), c.download("P")

    Uf, Pf = timed("form_guess")
    c.set_option("warm_start", 0)
    sols = {}
    for tag, (U0, P0) in (("extrapolation", (2 * Un - Ua, 2 * Pn - Pa)), ("previous_solution", (Un, Pn))):
        c.upload("T", T)
        c.thermal_buoyancy()
        c.upload("U", U0)
        c.upload("P", P0)
        sols[tag] = timed(tag)
    rel = lambda x, y: float(np.linalg.norm(x - y) / np.linalg.norm(y))  # noqa: E731
    Ub, Pb = sols["previous_solution"]
    out["U_rel_diff_form_vs_previous"] = rel(Uf, Ub)
    out["P_rel_diff_form_vs_previous"] = rel(Pf - Pf.mean(), Pb - Pb.mean())
    log(f"[{form}] same state: " + "; ".join(
        f"{k} Uzawa {v['uzawa_iterations']} in {v['s']:.2f} s" for k, v in out.items() if isinstance(v, dict)) +
        f"; U differs by {out['U_rel_diff_form_vs_previous']:.1e}, P by {out['P_rel_diff_form_vs_previous']:.1e}")
    return out


PEAK_BF16_TFLOPS = 2500.0   # MI355X dense BF16 MFMA


def unet_mfma(p, reps: int = 10):
    """MFMA throughput of the warm start's U-Net forward pass at the
    workload's finest grid (one volume, as the time loop runs it): the
    convolutions' multiply-adds (2 x voxels x 27 x ci x co per layer) over
    the pass's time, HIP events on the stream the kernels run on"""
    import torch

    from citcom_amd import unet
    from citcom_amd.learn import _p
    L = unet._lib()
    dims = p.dims()
    nv = [int(np.prod(unet.level_dims(dims, l))) for l in range(4)]
    level_of = [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0]
    flops = sum(2.0 * nv[level_of[k]] * 27 * ci * co for k, (ci, co) in enumerate(unet.LAYERS))
    w, b, _ = unet.load()
    wd = torch.from_numpy(w.view(np.int16)).cuda()
    bd = torch.from_numpy(b).cuda()
    x = torch.zeros((1, nv[0], 16), dtype=torch.bfloat16, device="cuda").normal_()
    out = torch.empty((1, nv[0], 16), dtype=torch.float32, device="cuda")
    ws = torch.empty(L.cl_unet_workspace(1, *dims), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(2):
        L.cl_unet_forward(_p(x), 1, *dims, _p(wd), _p(bd), _p(out), _p(ws), stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        if L.cl_unet_forward(_p(x), 1, *dims, _p(wd), _p(bd), _p(out), _p(ws), stream):
            raise RuntimeError("cl_unet_forward failed")
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    tf = flops / (ms * 1e-3) / 1e12
    return {"forward_ms": ms, "gflop": flops / 1e9, "tflops": tf, "frac_of_bf16_dense_peak": tf / PEAK_BF16_TFLOPS,
            "grid": list(dims)}


HEADLINE = "mixed16+learned"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="tdep128")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--form", choices=tuple(FORMS), default=HEADLINE, help="the form whose time is the value")
    ap.add_argument("--compare", default="plain,mixed16,mixed16+extrap",
                    help="forms timed beside it on the same window (comma-separated; '' for none)")
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
    graphs = not args.no_graphs
    main_run = run_form(p, T, args.form, args.steps, args.warmup, graphs, profile=True, same_state=True)
    others = {}
    for f in [f for f in args.compare.split(",") if f and f != args.form]:
        others[f] = run_form(p, T, f, args.steps, args.warmup, graphs)

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

    n_mv, ms_mv, bytes_mv = main_run["mv"]
    s_per_step = main_run["s_per_step"]
    mixed = FORMS[args.form][0]
    levels = main_run["gs_levels"]

    def passes(keys, name):
        """smoother launches (colour passes, box sweeps) of these (kind,
        level) keys, per multigrid call: launches, time, algorithmic bytes,
        the rate they reach, and their counter traffic"""
        n = sum(levels[k]["launches"] for k in keys)
        ms = sum(levels[k]["ms"] for k in keys)
        b = sum(levels[k]["bytes"] for k in keys)
        ach = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        return {"kernel": name, "levels": sorted(lv for _, lv in keys), "launches_per_mg_call": n,
                "ms_per_mg_call": ms, "avg_launch_ms": ms / max(n, 1e-9),
                "algorithmic_bytes_per_launch": b / max(n, 1e-9), "achieved": ach, "frac": ach / PEAK_HBM_GBS}

    # <vector type, stiffness type>: the FP16-stiffness form runs the FP16
    # instantiations on its two finest levels and the FP32 ones below; the
    # FP32 cycles smooth the levels between the exact level and the finest
    # by boxes (k_gs_box, one launch per sweep), the finest by colour passes
    fp16 = {lv for _, lv in levels if mixed == 2 and lv >= p.levels - 2 and lv >= 2}
    col = {k for k in levels if k[0] == "gs_colour"}
    box = {k for k in levels if k[0] == "gs_box"}
    groups = {"k_gs_colour<float, _Float16>": {k for k in col if k[1] in fp16},
              {0: "k_gs_colour<double, float>", 1: "k_gs_colour<float, float>",
               2: "k_gs_colour<float, float> (levels below the FP16 ones)"}[mixed]:
                  {k for k in col if k[1] not in fp16},
              "k_gs_box<_Float16>": {k for k in box if k[1] in fp16},
              "k_gs_box<float>": {k for k in box if k[1] not in fp16}}
    per_kernel = {k: passes(v, k) for k, v in groups.items() if v}
    if not col:
        sys.exit("bench: no colour-pass launches to measure on this workload (every level is smoothed in one "
                 "workgroup); use a larger --config")
    for k, r in per_kernel.items():
        # the counter summary's per-instantiation figures cover the level
        # sets of a run without boxes; with boxes the colour passes are the
        # finest level's alone
        r["traffic"] = (measured_traffic(k, args.config, "finest_level_hbm_bytes_per_launch")
                        if box and r["levels"] == [p.levels - 1] else measured_traffic(k, args.config))
        log(f"{k}: levels {r['levels']}, {r['launches_per_mg_call']:.0f} launches and {r['ms_per_mg_call']:.3f} ms "
            f"per multigrid call, {r['avg_launch_ms'] * 1e3:.2f} us each, {r['achieved']:.0f} GB/s")
    # the dominant kernel: the smoother, over all its launches (colour
    # passes and box sweeps, every instantiation and level; per
    # instantiation beside it)
    if box:
        allp = passes(col | box, "smoother: k_gs_colour and k_gs_box (every instantiation and level)")
        allp["traffic"] = measured_traffic("k_gs_colour (finest level) and k_gs_box (levels below)", args.config)
    else:
        allp = passes(col, "k_gs_colour (every instantiation and level)")
        allp["traffic"] = measured_traffic({0: "k_gs_colour<double, float>", 1: "k_gs_colour<float, float>",
                                            2: "k_gs_colour<float, _Float16> (two finest levels) and <float, float>"}
                                           [mixed], args.config)
    fin = passes({("gs_colour", p.levels - 1)}, "k_gs_colour, finest level")
    fin_kernel = [k for k, r in per_kernel.items() if p.levels - 1 in r["levels"]]
    fin["traffic"] = measured_traffic(fin_kernel[0], args.config, "finest_level_hbm_bytes_per_launch") \
        if fin_kernel else None
    log(f"all smoother launches: {allp['achieved']:.0f} GB/s; finest level {fin['avg_launch_ms'] * 1e3:.1f} us each, "
        f"{fin['achieved']:.0f} GB/s; matvec: {ms_mv:.1f} ms, {(bytes_mv / max(ms_mv, 1e-9)) / 1e6:.0f} GB/s")
    visc = (f"viscosity exp(-{p.viscE[0]:.4g} T)" + (f", {p.num_mat} layers (visc0 {p.visc0[:p.num_mat]})"
                                                      if p.num_mat > 1 else "")) if p.tdepv else "isoviscous"

    def summary(name, r):
        return {"form": FORMS[name][2], "s_per_step": r["s_per_step"],
                "uzawa_iterations_per_step": r["uzawa_iterations_per_step"],
                "uzawa_iterations": r["uzawa_iterations"], "mg_calls_per_step": r["mg_calls_per_step"],
                "steps_stopped_at_piterations_cap": r["steps_stopped_at_cap"], "Nu_top_last": r["Nu_top"],
                "vrms_last": r["vrms"], "model_time_last": r["elapsed_time"]}

    forms = {args.form: summary(args.form, main_run)}
    for f, r in others.items():
        forms[f] = summary(f, r)
        if f != "plain":
            continue
        # the forms run their own trajectories from the same seeded state; how
        # far apart they are after the same number of steps
        forms[args.form]["T_rel_diff_vs_plain_last"] = float(
            np.linalg.norm(main_run["T_final"] - r["T_final"]) / np.linalg.norm(r["T_final"]))
    learned = None
    if FORMS[args.form][1] == 2 or any(FORMS[f][1] == 2 for f in others):
        from citcom_amd import unet
        meta = unet.load()[2]
        learned = {"network": f"U-Net, {meta['params']} parameters, BF16 MFMA convolutions (FP32 accumulation)",
                   "trained_on": f"{meta['samples']} samples at {meta['dims']} nodes ({meta['base']} trajectories, "
                                 f"Ra and contrast drawn from [1e5, 1e7] x [1e2, 1e5])",
                   "data_s": meta["data_s"], "train_s": meta["train_s"], "epochs": meta["epochs"],
                   "holdout_error_extrapolation": meta["holdout_extrapolation_score"],
                   "holdout_error_network": meta["history"][-1].get("holdout"),
                   "cost": "trained once, offline (python -m citcom_amd.train_warm); not in the timed window",
                   "mfma": unet_mfma(p)}
        log(f"U-Net forward at {p.dims()}: {learned['mfma']['forward_ms']:.2f} ms, "
            f"{learned['mfma']['tflops']:.1f} TFLOP/s ({100 * learned['mfma']['frac_of_bf16_dense_peak']:.2f} % "
            f"of the BF16 dense peak)")
    config = {"workload": f"{args.config}: {nox - 1}x{noy - 1}x{noz - 1} elements, {p.levels} MG levels, "
                          f"Ra={p.rayleigh:g}, {visc}, accuracy={p.accuracy:g}, piterations={p.p_iterations}",
              "form": args.form, "window": f"steps {args.warmup + 1}..{args.warmup + args.steps} of each form's own "
                                           f"trajectory from the same seeded state (step-0 solve and "
                                           f"{args.warmup} warmup steps untimed), every form on the same window",
              "forms": forms, "same_state_check": main_run["same_state"]}
    if learned is not None:
        config["learned_warm_start"] = learned
    for f, r in others.items():
        config[f"speedup_vs_{f}"] = r["s_per_step"] / s_per_step
        log(f"{f}: {r['s_per_step']:.3f} s/step, {r['uzawa_iterations_per_step']:.1f} Uzawa iterations; "
            f"{args.form}: {s_per_step:.3f} s/step ({r['s_per_step'] / s_per_step:.2f}x)")
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
        "roofline": {"bound": "hbm", "achieved": allp["achieved"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": allp["frac"], "traffic": allp["traffic"], "kernel": allp["kernel"],
                     "levels": allp["levels"], "avg_launch_ms": allp["avg_launch_ms"],
                     "launches_per_mg_call": allp["launches_per_mg_call"],
                     "algorithmic_bytes_per_launch": allp["algorithmic_bytes_per_launch"],
                     "per_kernel": per_kernel, "finest_level": fin, "device_streams": stream},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        mg_plain = others["plain"]["mg_calls_per_step"] if "plain" in others else main_run["mg_calls_per_step"]
        out["cpu_baseline"] = cpu_baseline(args.config, args.seed, mg_plain)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
