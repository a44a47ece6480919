"""The CPU restatement (oracle/) against the reference's own outputs.

Fixtures in tests/golden/ were written by tests/golden/make_golden.py from
the unmodified reference (oracle/_ref/ref_harness).  Geometry, flags,
viscosity, the assembled stiffness, BI, BPI, elt_del, the grid transfers,
div/grad, the SUPG step and the Nusselt number must agree bit for bit; the
matvec to summation-order rounding; the smoother to the float rounding of the
reference's incremental Ad bookkeeping; the Uzawa solution to the solver
tolerance.
"""
import hashlib

import numpy as np
import pytest

from conftest import golden, initial_temperature, problem, rel_l2
import refdump

DECKS = ["tiny", "tiny_tdep", "tiny_layer", "tiny2"]
# the reference's own example decks (examples/Busse1993): rigid lids, a
# non-cubic box, a graded z grid, the Arrhenius law, the shipped restart field
EXAMPLES = ["busse1a", "busse2"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make(pyoracle, name, smoother=0):
    p = problem(name)
    o = pyoracle.Oracle(p, smoother=smoother)
    T = initial_temperature(name, p)
    o.set_viscosity(o.visc_from_T(T))
    return p, o, T


@pytest.mark.parametrize("name", DECKS + EXAMPLES)
def test_setup_bitwise(pyoracle, name):
    p, o, T = make(pyoracle, name)
    g = golden(f"{name}_setup.npz")
    for lev in range(p.levels):
        for key in ("X", "Y", "Z", "GNX", "MASS", "EVI", "EG", "BI", "BPI", "K"):
            assert sha(o.get(key, lev)) == str(g[f"sha_{key}_{lev}"]), (key, lev)
        assert sha(o.get("GDA", lev).reshape(-1, 11)[:, refdump.GDA_USED]) == str(g[f"sha_GDA_{lev}"]), lev
        assert sha(o.get("ECO", lev).reshape(-1, 13)[:, refdump.ECO_USED]) == str(g[f"sha_ECO_{lev}"]), lev
        assert sha(o.get("TWW", lev).reshape(-1, 9)[:, 1:]) == str(g[f"sha_TWW_{lev}"]), lev
        assert sha(o.get("NODE", lev) & np.uint32(0xE)) == str(g[f"sha_NODEV_{lev}"]), lev
    assert sha(o.get("NODE") & np.uint32(0x7E)) == str(g["sha_nodeflags"])
    assert sha(T) == str(g["sha_T"])
    b = o.thermal_buoyancy(T)
    assert sha(b) == str(g["sha_buoyancy"])
    assert sha(o.assemble_forces(b)) == str(g["sha_F"])


@pytest.mark.parametrize("name", DECKS)
def test_operators_against_reference(pyoracle, name):
    p, o, T = make(pyoracle, name)
    g = golden(f"{name}_ops.npz")
    inp = refdump.harness_op_inputs(p, 7, o.strip)
    for lev in range(p.levels):
        x = inp[lev]
        mv = o.matvec(lev, x["u"])
        ref = g[f"mv_out_{lev}"]
        assert np.abs(mv - ref).max() <= 1e-14 * np.abs(ref).max()
        assert np.array_equal(o.div(lev, x["u"]), g[f"div_out_{lev}"])
        assert np.array_equal(o.grad(lev, x["p"]), g[f"grad_out_{lev}"])
        if lev > 0:
            assert np.array_equal(o.project(lev, x["u"]), g[f"proj_out_{lev}"])
            assert np.array_equal(o.interp(lev, x["c"]), g[f"interp_out_{lev}"])
        # lexicographic node-block Gauss-Seidel: the reference scatters float
        # increments into Ad, the restatement recomputes each row in double
        d, Ad = o.gauss_seidel(lev, np.zeros_like(x["F"]), x["F"], 3, 0)
        for mine, key in ((d, "gs0_d"), (Ad, "gs0_Ad")):
            ref = g[f"{key}_{lev}"]
            assert np.abs(mine - ref).max() <= 1e-6 * np.abs(ref).max(), key
        d, Ad = o.gauss_seidel(lev, x["d0"], x["F"], 2, 1)
        for mine, key in ((d, "gs1_d"), (Ad, "gs1_Ad")):
            ref = g[f"{key}_{lev}"]
            assert np.abs(mine - ref).max() <= 1e-6 * np.abs(ref).max(), key


@pytest.mark.parametrize("name", DECKS)
def test_stokes_pg_nu_against_reference(pyoracle, name):
    p, o, T = make(pyoracle, name)
    g = golden(f"{name}_run.npz")
    nno = o.nno()
    F = o.assemble_forces(o.thermal_buoyancy(T))
    U, P, st = o.solve_stokes(np.zeros(3 * nno), np.zeros(o.nel()), F)
    assert rel_l2(U, g["step0_U"]) < 1e-6
    assert rel_l2(P, g["step0_P"]) < 1e-6
    V = np.concatenate([g["step0_V1"], g["step0_V2"], g["step0_V3"]])
    nu = o.heat_flux(T, V)
    assert nu[0] == pytest.approx(g["steps"][0, 3], rel=1e-8)
    assert nu[1] == pytest.approx(g["steps"][0, 4], rel=1e-8)
    T1, Td1, dt, _ = o.pg_timestep(T, np.zeros(nno, np.float32), V)
    assert np.float32(dt) == np.float32(g["steps"][1, 2])
    assert np.array_equal(T1, g["pg1_T"])
    assert np.array_equal(Td1, g["pg1_Tdot"])


def test_iso16_first_steps(pyoracle):
    """16^3 isoviscous Ra=1e4: the first five steps of the time loop."""
    from citcom_amd import deck
    p = problem("iso16")
    o = pyoracle.Oracle(p)
    o.set_viscosity(o.visc_from_T(deck.perturbed_temperature(p)))
    g = golden("iso16_run.npz")
    T = deck.perturbed_temperature(p)
    assert np.array_equal(T, g["step0_T"])
    nno, nel = o.nno(), o.nel()
    U, P, _ = o.solve_stokes(np.zeros(3 * nno), np.zeros(nel), o.assemble_forces(o.thermal_buoyancy(T)))
    assert rel_l2(U, g["step0_U"]) < 1e-6 and rel_l2(P, g["step0_P"]) < 1e-6
    Tdot = np.zeros(nno, np.float32)
    for step in range(1, 6):
        V = pyoracle.v_from_vector(U, nno).reshape(-1)
        T, Tdot, dt, _ = o.pg_timestep(T, Tdot, V)
        assert dt == pytest.approx(g["steps"][step, 2], rel=1e-5)
        U, P, _ = o.solve_stokes(U, P, o.assemble_forces(o.thermal_buoyancy(T)))
        # the harness records Nu after the step's Stokes solve
        nu = o.heat_flux(T, pyoracle.v_from_vector(U, nno).reshape(-1))
        assert nu[0] == pytest.approx(g["steps"][step, 3], rel=1e-5)


@pytest.mark.parametrize("name", EXAMPLES)
def test_example_decks_tight_solve(pyoracle, name):
    """The first Stokes solve of the reference's example decks at accuracy
    1e-8 (lexicographic smoother, as the reference)."""
    p, o, T = make(pyoracle, name + "@1e-8")
    g = golden(f"{name}_at_1e-8_run.npz")
    U, P, st = o.solve_stokes(np.zeros(3 * o.nno()), np.zeros(o.nel()), o.assemble_forces(o.thermal_buoyancy(T)))
    assert rel_l2(U, g["step0_U"]) < 1e-6 and rel_l2(P, g["step0_P"]) < 1e-6
