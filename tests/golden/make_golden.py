"""Regenerate the golden fixtures in this directory from the reference.

Runs oracle/_ref/ref_harness (the unmodified reference objects linked with a
dump harness; build it with ``make -C oracle ref``) on the decks of
citcom_amd.deck and stores what the tests compare against:

* ``<deck>_setup.npz``  SHA-256 of every per-level array the reference
  builds (geometry, flags, viscosity, stiffness expanded to full rows, BI,
  BPI, elt_del) plus the finest-level fields; arrays are hashed because the
  restatement and the HIP path reproduce them bit for bit.
* ``<deck>_ops.npz``    outputs of n_assemble_del2_u, gauss_seidel,
  project_vector, interp_vector, assemble_div_u, assemble_grad_p for the
  harness's seeded inputs (regenerated in Python by ``xorshift``).
* ``<deck>_run.npz``    the first Stokes solve, the first PG step and the
  per-step (time, dt, Nu_top, Nu_bot) of a short run.
* ``iso16_run.npz``      the 16^3 isoviscous Ra=1e4 workload: 200 steps.
* ``<deck>_at_<acc>_run.npz``  the same runs with accuracy = acc: two solvers
  that stop on relative-increment tests at 1e-6 can differ by O(1e-6), so
  field parity at 1e-6 is shown where both are converged well below it.
* ``busse1a_steady.npz``  the reference program (oracle/_ref/citcom.mpi) on
  the case-1a deck under MPI (4 ranks), 20000 steps to its steady state:
  final T and V (assembled from the per-rank .temp/.velo files), Nu_top,
  Nu_bot and time from the top and bottom ranks' .ave files.
* ``<deck>_ave.npz``     the layer profiles of the .ave output (Have.T,
  Have.vrms, Have.Vi: heat_flux and averages) after 3 steps, with the final
  T and V they were computed from.
* ``tdep64_at_1e-9_run.npz``  the 64^3-element workload (5 levels, Ra 1e6,
  contrast 1e4, seeded temperature) at accuracy 1e-9 under MPI (8 ranks):
  the step-0 solve and the state after step 1, on every other node and
  element, and the per-step scalars.
* ``busse2_tic.npz``     the initial temperature the reference ships with its
  example deck (examples/Busse1993/case2-tic.temp.0.0), as read by the
  restart reader.
* ``busse*_setup.npz``   the setup hashes of the reference's example decks
  (no raw fields: 33^3 nodes); ``busse*_run.npz`` per-step (time, dt,
  Nu_top, Nu_bot) of 20 steps at the decks' own accuracy (1e-3);
  ``busse*_at_1e-8_run.npz`` the first Stokes solve and the state after 3
  steps (U, P, T in float32) at accuracy 1e-8.

Usage: python tests/golden/make_golden.py [deck ...]
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from citcom_amd import deck  # noqa: E402
import refdump  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
PROGRAM = os.path.join(ROOT, "oracle", "_ref", "citcom.mpi")
MPIEXEC = "/opt/conda/bin/mpiexec"
REF_EXAMPLE_TIC = "/root/reference/examples/Busse1993/case2-tic.temp.0.0"


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_harness(deck_name: str, workdir: str, mode: str, arg: str | None = None) -> str:
    keys = deck.workload(deck_name)
    keys["datafile"] = os.path.join(workdir, "out", deck_name)
    keys["oldfile"] = keys["datafile"]
    os.makedirs(os.path.join(workdir, "out"), exist_ok=True)
    if str(keys.get("restart", 0)) == "1":
        T = np.load(os.path.join(HERE, "busse2_tic.npz"))["T"]
        deck.write_restart_temperature(keys["oldfile"], T)
    path = os.path.join(workdir, f"{deck_name}.input")
    deck.write_deck(path, keys)
    out = os.path.join(workdir, f"{deck_name}_{mode}")
    cmd = [HARNESS, path, out, mode] + ([arg] if arg else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=workdir)
    return out


def setup_fixture(name: str, d: str, raw: bool = True) -> dict:
    p = deck.problem_from_deck({k: str(v) for k, v in deck.workload(name).items()})
    out = {}
    for lev in range(p.levels):
        nox, noy, noz = p.dims(lev)
        for key, fn in (("X", "X1.f32"), ("Y", "X2.f32"), ("Z", "X3.f32"), ("GNX", "GNX.f32"),
                        ("MASS", "MASS.f32"), ("EVI", "EVI.f32"), ("EG", "EG.f32"),
                        ("BI", "BI.f64"), ("BPI", "BPI.f64")):
            out[f"sha_{key}_{lev}"] = sha(refdump.load(d, fn, lev))
        out[f"sha_GDA_{lev}"] = sha(refdump.load(d, "GDA.f32", lev).reshape(-1, 11)[:, refdump.GDA_USED])
        out[f"sha_ECO_{lev}"] = sha(refdump.load(d, "ECO.f32", lev).reshape(-1, 13)[:, refdump.ECO_USED])
        out[f"sha_TWW_{lev}"] = sha(refdump.load(d, "TWW.f32", lev).reshape(-1, 9)[:, 1:])
        out[f"sha_NODEV_{lev}"] = sha(refdump.load(d, "NODE.u32", lev) & np.uint32(0xE))
        full = refdump.full_rows_from_half(*(refdump.load(d, f"K{i}.f32", lev) for i in (1, 2, 3)),
                                           refdump.load(d, "MAP.i32", lev), nox, noy, noz)
        out[f"sha_K_{lev}"] = sha(full)
    for key, fn in (("T", "T.f32"), ("buoyancy", "buoyancy.f32"), ("F", "F.f64"), ("TB3", "TB3.f32")):
        out[f"sha_{key}"] = sha(refdump.load(d, fn))
    # velocity and temperature boundary bits (VBX|VBZ|VBY|TBX|TBZ|TBY)
    out["sha_nodeflags"] = sha(refdump.load(d, "node.u32") & np.uint32(0x7E))
    if raw:
        out["T"] = refdump.load(d, "T.f32")
        out["F"] = refdump.load(d, "F.f64")
        out["EVI"] = refdump.load(d, "EVI.f32", p.levmax)
    return out


def ops_fixture(name: str, d: str) -> dict:
    p = deck.problem_from_deck({k: str(v) for k, v in deck.workload(name).items()})
    out = {}
    for lev in range(p.levels):
        for fn in ("mv_out.f64", "gs0_d.f64", "gs0_Ad.f64", "gs1_d.f64", "gs1_Ad.f64", "div_out.f64",
                   "grad_out.f64") + (("proj_out.f64", "interp_out.f64") if lev > 0 else ()):
            out[f"{fn.split('.')[0]}_{lev}"] = refdump.load(d, fn, lev)
    return out


def run_fixture(d: str, fields=("U.f64", "P.f64", "T.f32", "Tdot.f32", "V1.f32", "V2.f32", "V3.f32")) -> dict:
    out = {"steps": np.loadtxt(os.path.join(d, "steps.txt"), ndmin=2)}
    for tag in ("step0", "pg1", "final"):
        for fn in fields:
            path = os.path.join(d, f"{tag}_{fn}")
            if os.path.exists(path):
                out[f"{tag}_{fn.split('.')[0]}"] = refdump.load(d, f"{tag}_{fn}")
    return out


def steady_fixture(tmp: str, nsteps: int = 20000, nproc=(2, 1, 2)) -> dict:
    from citcom_amd import outputs
    keys = deck.workload("busse1a")
    out = os.path.join(tmp, "steady")
    os.makedirs(out, exist_ok=True)
    keys.update({"datafile": os.path.join(out, "run"), "maxstep": nsteps, "storage_spacing": nsteps // 10,
                 "nprocx": nproc[0], "nprocy": nproc[1], "nprocz": nproc[2]})
    path = os.path.join(out, "deck.input")
    deck.write_deck(path, keys)
    ranks = nproc[0] * nproc[1] * nproc[2]
    subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(ranks), PROGRAM, path], cwd=out,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return read_steady(out, nsteps, nproc)


def read_steady(out: str, nsteps: int, nproc) -> dict:
    from citcom_amd import outputs
    ranks = nproc[0] * nproc[1] * nproc[2]
    p = deck.problem_from_deck({k: str(v) for k, v in deck.workload("busse1a").items()})
    T = outputs.assemble([outputs.read_temp(os.path.join(out, f"run.temp.{r}.{nsteps}")).values[:, :1]
                          for r in range(ranks)], p.dims(), nproc)[:, 0]
    V = outputs.assemble([outputs.read_velo(os.path.join(out, f"run.velo.{r}.{nsteps}")).values
                          for r in range(ranks)], p.dims(), nproc)
    # ranks 0 .. nprocz-1 write .ave files, each with the Nusselt numbers of
    # its own subdomain's top and bottom (Process_buoyancy.c:166-193): the
    # surface values come from the top and the bottom rank
    top = outputs.read_ave(os.path.join(out, f"run.ave.{nproc[2] - 1}.{nsteps}"))
    bot = outputs.read_ave(os.path.join(out, f"run.ave.0.{nsteps}"))
    return {"T": T.astype(np.float32), "V": V.T.astype(np.float32).reshape(-1), "steps": np.int64(top.timesteps),
            "time": np.float64(top.time), "nu": np.array([top.nu_top, bot.nu_bot])}


def assemble_rank_dumps(d: str, tag: str, p, nproc) -> dict:
    """Global U (FP64 [3 nno]), P (FP64 [nel]) and T (FP32 [nno]) from the
    per-rank dumps of a harness run under MPI (<name>.r<rank>): nodes by
    rank_blocks (shared face nodes must agree), elements by the same blocks
    without their upper faces."""
    from citcom_amd import outputs
    nox, noy, noz = p.dims()
    ranks = nproc[0] * nproc[1] * nproc[2]
    T = outputs.assemble([refdump.load(d, f"{tag}_T.f32.r{r}").reshape(-1, 1) for r in range(ranks)],
                         p.dims(), nproc)[:, 0]
    U = outputs.assemble([refdump.load(d, f"{tag}_U.f64.r{r}").reshape(-1, 3) for r in range(ranks)],
                         p.dims(), nproc).reshape(-1)
    P = np.full((noy - 1, nox - 1, noz - 1), np.nan)
    for me, sy, sx, sz in deck.rank_blocks(p.dims(), nproc):
        blk = P[sy.start:sy.stop - 1, sx.start:sx.stop - 1, sz.start:sz.stop - 1]
        blk[...] = refdump.load(d, f"{tag}_P.f64.r{me}").reshape(blk.shape)
    assert not np.isnan(P).any()
    return {"T": T, "U": U, "P": P.reshape(-1)}


def tdep64_fixture(d: str, nproc=(2, 4, 1)) -> dict:
    """The 64^3-element workload at accuracy 1e-9: the step-0 solve and step 1
    of the reference (harness under MPI).  Stored on every other node and
    element (each axis) in FP32 to stay small: U [33^3][3], P [32^3], and T
    after step 1; the per-step (time, dt, Nu_top, Nu_bot)."""
    p = deck.problem_from_deck({k: str(v) for k, v in deck.workload("tdep64@1e-9").items()})
    nox, noy, noz = p.dims()
    out = {"steps": np.loadtxt(os.path.join(d, "steps.txt"), ndmin=2)}
    for tag in ("step0", "final"):
        f = assemble_rank_dumps(d, tag, p, nproc)
        out[f"{tag}_U2"] = f["U"].reshape(noy, nox, noz, 3)[::2, ::2, ::2].astype(np.float32).reshape(-1)
        out[f"{tag}_P2"] = f["P"].reshape(noy - 1, nox - 1, noz - 1)[::2, ::2, ::2].astype(np.float32).reshape(-1)
        if tag == "final":
            out["final_T2"] = f["T"].reshape(noy, nox, noz)[::2, ::2, ::2].astype(np.float32).reshape(-1)
    return out


def run_tdep64_mpi(tmp: str, nproc=(2, 4, 1), piterations: int = 3000) -> str:
    keys = deck.workload("tdep64@1e-9")
    p = deck.problem_from_deck({k: str(v) for k, v in keys.items()})
    deck.write_restart_temperature(os.path.join(tmp, "seed"), deck.seeded_temperature(p, seed=1), p.dims(), nproc)
    keys.update({"datafile": os.path.join(tmp, "run"), "restart": 1, "restart_timesteps": 0,
                 "oldfile": os.path.join(tmp, "seed"), "piterations": piterations,
                 "nprocx": nproc[0], "nprocy": nproc[1], "nprocz": nproc[2]})
    path = os.path.join(tmp, "tdep64.input")
    deck.write_deck(path, keys)
    out = os.path.join(tmp, "dump")
    ranks = nproc[0] * nproc[1] * nproc[2]
    subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(ranks), HARNESS, path, out, "run", "1"], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, cwd=tmp)
    return out


def main(names):
    if not os.path.exists(HARNESS):
        raise SystemExit(f"{HARNESS} missing: make -C oracle ref")
    with tempfile.TemporaryDirectory() as tmp:
        for name in names:
            if name == "tdep64_at_1e-9":
                # 8 MPI ranks, about 15 minutes on 8 cores (2258 + 2709 Uzawa
                # iterations); TDEP64_DUMP names an existing dump directory
                d = os.environ.get("TDEP64_DUMP") or run_tdep64_mpi(tmp)
                np.savez_compressed(os.path.join(HERE, "tdep64_at_1e-9_run.npz"), **tdep64_fixture(d))
                print("wrote", name)
                continue
            if name.endswith("_ave"):
                # the layer profiles of averages after 3 steps (and the
                # per-step scalars of the same run)
                base = name[:-4]
                d = run_harness(base, tmp, "run", "3")
                fx = {k: refdump.load(d, f"final_ave{k}.f32") for k in ("T", "vrms", "Vi")}
                for k in ("T", "V1", "V2", "V3"):
                    fx[f"final_{k}"] = refdump.load(d, f"final_{k}.f32")
                fx["steps"] = np.loadtxt(os.path.join(d, "steps.txt"), ndmin=2)
                np.savez_compressed(os.path.join(HERE, f"{base.replace('@', '_at_')}_ave.npz"), **fx)
                print("wrote", name)
                continue
            if name == "busse1a_steady":
                np.savez_compressed(os.path.join(HERE, "busse1a_steady.npz"), **steady_fixture(tmp))
                print("wrote", name)
                continue
            if name == "busse2_tic":
                T = deck.read_restart_temperature(REF_EXAMPLE_TIC)
                np.savez_compressed(os.path.join(HERE, "busse2_tic.npz"), T=T)
                print("wrote", name)
                continue
            if name.startswith("busse"):
                tag = name.replace("@", "_at_")
                if "@" in name:
                    d = run_harness(name, tmp, "run", "3")
                    fx = {k: (v.astype(np.float32) if k != "steps" else v)
                          for k, v in run_fixture(d, ("U.f64", "P.f64", "T.f32")).items()
                          if not k.startswith("pg1") and k != "step0_T"}
                    np.savez_compressed(os.path.join(HERE, f"{tag}_run.npz"), **fx)
                else:
                    d = run_harness(name, tmp, "setup")
                    np.savez_compressed(os.path.join(HERE, f"{name}_setup.npz"), **setup_fixture(name, d, raw=False))
                    d = run_harness(name, tmp, "run", "20")
                    np.savez_compressed(os.path.join(HERE, f"{name}_run.npz"), steps=run_fixture(d, ())["steps"])
                print("wrote", name)
                continue
            if name.startswith("iso16"):
                d = run_harness(name, tmp, "run", "200")
                tag = name.replace("@", "_at_")
                np.savez_compressed(os.path.join(HERE, f"{tag}_run.npz"), **run_fixture(d, ("U.f64", "P.f64", "T.f32")))
                continue
            if "@" in name:
                d = run_harness(name, tmp, "run", "3")
                np.savez_compressed(os.path.join(HERE, f"{name.replace('@', '_at_')}_run.npz"), **run_fixture(d))
                print("wrote", name)
                continue
            d = run_harness(name, tmp, "setup")
            np.savez_compressed(os.path.join(HERE, f"{name}_setup.npz"), **setup_fixture(name, d))
            d = run_harness(name, tmp, "ops", "7")
            np.savez_compressed(os.path.join(HERE, f"{name}_ops.npz"), **ops_fixture(name, d))
            d = run_harness(name, tmp, "run", "3")
            np.savez_compressed(os.path.join(HERE, f"{name}_run.npz"), **run_fixture(d))
            print("wrote", name)


if __name__ == "__main__":
    main(sys.argv[1:] or ["tiny", "tiny_tdep", "tiny@1e-9", "tiny_tdep@1e-9", "iso16", "iso16@1e-8"])
