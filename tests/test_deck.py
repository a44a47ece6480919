"""Deck handling: the reference's key=value format, its defaults, and the
node coordinates node_locations computes (checked against the reference's
own coordinates through the golden hashes)."""
import hashlib
import os
import shutil

import numpy as np
import pytest

from conftest import golden, problem

REF_PROGRAM = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                           "citcom.mpi")
MPIEXEC = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"


def test_parse_roundtrip(tmp_path):
    from citcom_amd import deck
    keys = deck.workload("tdep128")
    path = tmp_path / "d.input"
    deck.write_deck(str(path), keys)
    back = deck.read_deck(str(path))
    for k, v in keys.items():
        assert back[k] == str(v).strip('"'), k
    p = deck.problem_from_deck(back)
    assert p.dims() == (129, 129, 65) and p.levels == 6
    assert p.vmin_on and p.vmax_on and abs(p.viscE[0] - np.log(1e4)) < 1e-6


def test_comments_and_shared_lines():
    from citcom_amd import deck
    d = deck.parse_deck('a=1 b="x y"  # c=3\n  levels=4\n#z=9')
    assert d == {"a": "1", "b": "x y", "levels": "4"}


def test_coordinates_match_reference():
    for name in ("tiny", "tiny_tdep"):
        p = problem(name)
        g = golden(f"{name}_setup.npz")
        lev = p.levmax
        nox, noy, noz = p.dims(lev)
        X = np.broadcast_to(p.coords[0][None, :, None], (noy, nox, noz)).astype(np.float32).reshape(-1)
        Z = np.broadcast_to(p.coords[2][None, None, :], (noy, nox, noz)).astype(np.float32).reshape(-1)
        assert hashlib.sha256(np.ascontiguousarray(X).tobytes()).hexdigest() == str(g[f"sha_X_{lev}"])
        assert hashlib.sha256(np.ascontiguousarray(Z).tobytes()).hexdigest() == str(g[f"sha_Z_{lev}"])


def test_seeded_temperature():
    from citcom_amd import deck
    p = problem("tdep64")
    T = deck.seeded_temperature(p, seed=3)
    nox, noy, noz = p.dims()
    Tv = T.reshape(noy, nox, noz)
    assert np.all(Tv[:, :, 0] == 1.0) and np.all(Tv[:, :, -1] == 0.0)
    assert np.abs(Tv - (1 - p.coords[2])[None, None, :]).max() <= 0.0100001
    assert np.array_equal(T, deck.seeded_temperature(p, seed=3))


def test_restart_files_per_rank(tmp_path):
    """One restart file per rank over its local nodes, ranks in the order of
    parallel_domain_decomp0 (Parallel_related.c:108-121): reassembling the
    files gives the global field back, shared faces agreeing."""
    from citcom_amd import deck
    p = problem("tiny_tdep")
    T = deck.seeded_temperature(p, seed=5)
    nox, noy, noz = p.dims()
    nproc = (2, 2, 1)
    paths = deck.write_restart_temperature(str(tmp_path / "seed"), T, p.dims(), nproc)
    assert [os.path.basename(x) for x in paths] == [f"seed.temp.{r}.0" for r in range(4)]
    ex, ey, ez = (nox - 1) // 2, (noy - 1) // 2, noz - 1
    back = np.full((noy, nox, noz), np.nan)
    for me, path in enumerate(paths):
        with open(path) as fh:
            head = fh.readline().split()
            vals = np.array([float(x) for x in fh.read().split()])
        assert int(head[0]) == vals.size == (ex + 1) * (ey + 1) * (ez + 1)
        kz, jy = me % 1, me // 2
        ix = (me - jy * 2) // 1
        blk = back[jy * ey:jy * ey + ey + 1, ix * ex:ix * ex + ex + 1, kz * ez:kz * ez + ez + 1]
        v = vals.reshape(blk.shape)
        assert np.all(np.isnan(blk) | (blk == v))
        blk[...] = v
    assert np.array_equal(back.reshape(-1).astype(np.float32), T)


@pytest.mark.skipif(not (os.path.exists(REF_PROGRAM) and os.path.exists(MPIEXEC)),
                    reason="needs the reference program built in oracle/_ref and mpiexec")
def test_reference_under_mpi_reads_rank_files(tmp_path):
    """The pristine reference program run on 4 ranks from the per-rank
    restart files follows the 1-rank run: same dt, velocity and pressure
    norms within the solver tolerance."""
    import re
    import subprocess
    from citcom_amd import deck
    p = problem("tiny_tdep")
    T = deck.seeded_temperature(p, seed=1)
    logs = []
    for nproc in ((1, 1, 1), (2, 2, 1)):
        out = tmp_path / "".join(map(str, nproc))
        out.mkdir()
        keys = deck.workload("tiny_tdep")
        keys.update({"datafile": str(out / "run"), "restart": 1, "restart_timesteps": 0,
                     "oldfile": str(out / "seed"), "maxstep": 2,
                     "nprocx": nproc[0], "nprocy": nproc[1], "nprocz": nproc[2]})
        deck.write_restart_temperature(str(out / "seed"), T, p.dims(), nproc)
        deck.write_deck(str(out / "deck.input"), keys)
        subprocess.run([MPIEXEC, "-launcher", "fork", "-n", str(np.prod(nproc)), REF_PROGRAM,
                        str(out / "deck.input")], capture_output=True, timeout=120, cwd=str(out))
        logs.append((out / "run.log0").read_text())
    for key in (r"dt = ([0-9.e+-]+)", r"with v ([0-9.e+-]+)", r"p ([0-9.e+-]+) dp/p"):
        a, b = (float(re.findall(key, t)[-1]) for t in logs)
        assert abs(a - b) <= 1e-5 * abs(a), (key, a, b)
