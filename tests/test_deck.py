"""Deck handling: the reference's key=value format, its defaults, and the
node coordinates node_locations computes (checked against the reference's
own coordinates through the golden hashes)."""
import hashlib

import numpy as np

from conftest import golden, problem


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
