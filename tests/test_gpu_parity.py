"""HIP path (libcitcom_hip.so through its C ABI) against the CPU restatement
and the reference's golden outputs, on the test decks.

Operators that the reference computes deterministically are compared bit for
bit: geometry, viscosity on every level, stiffness rows, BI, BPI, elt_del,
forces, grid transfers, div/grad and the 8-colour smoother (against the
restatement's 8-colour variant).  Solutions are compared to the solver
tolerance (relative L2 < 1e-6 on the fields).
"""
import numpy as np
import pytest

from conftest import golden, initial_temperature, problem, rel_l2
import refdump

pytestmark = pytest.mark.gpu

DECKS = ["tiny", "tiny_tdep", "tiny_layer", "tiny2"]


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module", params=DECKS)
def pair(request, pyoracle):
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    name = request.param
    say("deck", name)
    p = problem(name)
    c = Citcom(p)
    o = pyoracle.Oracle(p, smoother=1)
    T = deck.perturbed_temperature(p)
    c.initialize(T)
    c.get_system_viscosity()
    c.construct_stiffness_B_matrix()
    o.set_viscosity(o.visc_from_T(T))
    yield name, p, c, o, T
    c.close()


def test_geometry_bitwise(pair):
    name, p, c, o, T = pair
    for lev in range(p.levels):
        for key in ("GNX", "GDA", "ECO", "TWW", "MASS"):
            assert np.array_equal(c.download(key, lev), o.get(key, lev)), (key, lev)
        assert np.array_equal(c.download("NODE", lev) & np.uint32(0x7E), o.get("NODE", lev) & np.uint32(0x7E))


def test_operator_build_bitwise(pair):
    name, p, c, o, T = pair
    for lev in range(p.levels):
        assert np.array_equal(c.download("EVI", lev), o.get("EVI", lev)), ("EVI", lev)
        assert np.array_equal(c.download("EG", lev), o.get("EG", lev)), ("EG", lev)
        assert np.array_equal(c.download("BI", lev), o.get("BI", lev)), ("BI", lev)
        assert np.array_equal(c.download("BPI", lev), o.get("BPI", lev)), ("BPI", lev)
        assert np.array_equal(c.download("K", lev), o.get("K", lev)), ("K", lev)


def test_buoyancy_and_forces(pair):
    name, p, c, o, T = pair
    b = c.download("buoyancy")
    bo = o.thermal_buoyancy(T)
    # the horizontal average is a parallel double sum on the GPU
    assert np.abs(b - bo).max() <= 4 * np.finfo(np.float32).eps * np.abs(bo).max()
    c.assemble_forces()
    assert np.array_equal(c.download("F"), o.assemble_forces(b))


def test_operators_bitwise(pair):
    name, p, c, o, T = pair
    inp = refdump.harness_op_inputs(p, 7, o.strip)
    for lev in range(p.levels):
        x = inp[lev]
        assert np.array_equal(c.apply("matvec", lev, x["u"]), o.matvec(lev, x["u"])), lev
        assert np.array_equal(c.apply("div", lev, x["u"]), o.div(lev, x["u"])), lev
        assert np.array_equal(c.apply("grad", lev, x["p"]), o.grad(lev, x["p"])), lev
        if lev > 0:
            assert np.array_equal(c.apply("project", lev, x["u"]), o.project(lev, x["u"])), lev
            assert np.array_equal(c.apply("interp", lev, x["c"]), o.interp(lev, x["c"])), lev
        d, Ad = c.apply("gs0", lev, None, x["F"], sweeps=3)
        do, Ado = o.gauss_seidel(lev, np.zeros_like(x["F"]), x["F"], 3, 0)
        assert np.array_equal(d, do) and np.array_equal(Ad, Ado), lev
        d, Ad = c.apply("gs1", lev, x["d0"], x["F"], sweeps=2)
        do, Ado = o.gauss_seidel(lev, x["d0"], x["F"], 2, 1)
        assert np.array_equal(d, do) and np.array_equal(Ad, Ado), lev


def test_matches_reference_operators(pair):
    """The same inputs through the reference (golden fixture): transfers and
    div/grad exact; the colour smoother converges to the same fixed point, so
    only its residual reduction is compared."""
    name, p, c, o, T = pair
    g = golden(f"{name}_ops.npz")
    inp = refdump.harness_op_inputs(p, 7, o.strip)
    for lev in range(p.levels):
        x = inp[lev]
        ref = g[f"mv_out_{lev}"]
        assert np.abs(c.apply("matvec", lev, x["u"]) - ref).max() <= 1e-14 * np.abs(ref).max()
        assert np.array_equal(c.apply("div", lev, x["u"]), g[f"div_out_{lev}"])
        assert np.array_equal(c.apply("grad", lev, x["p"]), g[f"grad_out_{lev}"])
        if lev > 0:
            assert np.array_equal(c.apply("project", lev, x["u"]), g[f"proj_out_{lev}"])
            assert np.array_equal(c.apply("interp", lev, x["c"]), g[f"interp_out_{lev}"])


def test_stokes_against_reference(pair):
    """accuracy 1e-6 on both sides: the 8-colour and the lexicographic
    smoother stop at different iterates, each within its relative-increment
    tolerance of the discrete solution, so they agree to O(1e-6)."""
    name, p, c, o, T = pair
    g = golden(f"{name}_run.npz")
    c.upload("U", np.zeros(3 * c.nno()))
    c.upload("P", np.zeros(c.nel()))
    c.thermal_buoyancy()
    c.assemble_forces()
    st = c.solve_constrained_flow_iterative()
    U, P = c.download("U"), c.download("P")
    eu, ep = rel_l2(U, g["step0_U"]), rel_l2(P, g["step0_P"])
    say(name, "uzawa", st["iterations"], "its; rel L2 vs reference at accuracy 1e-6: U", eu, "P", ep)
    assert eu < 1e-5 and ep < 1e-5
    V = c.download("V")
    assert np.array_equal(V, U.reshape(-1, 3).T.astype(np.float32).reshape(-1))


@pytest.mark.parametrize("name", DECKS)
def test_stokes_converged_parity(name):
    """Both solvers converged well below the bar (accuracy 1e-9): velocity and
    pressure agree with the reference's to 1e-6 relative L2."""
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    p = problem(f"{name}@1e-9")
    g = golden(f"{name}_at_1e-9_run.npz")
    c = Citcom(p)
    c.initialize(deck.perturbed_temperature(p))
    st = c.general_stokes_solver()
    U, P = c.download("U"), c.download("P")
    eu, ep = rel_l2(U, g["step0_U"]), rel_l2(P, g["step0_P"])
    say(name, "accuracy 1e-9:", st["iterations"], "Uzawa its; rel L2 U", eu, "P", ep)
    assert eu < 1e-6 and ep < 1e-6
    c.close()


def test_vrms(pair):
    """volume-RMS velocity of the reference's step-0 velocity: the GPU
    reduction against the restatement's sequential sum"""
    name, p, c, o, T = pair
    g = golden(f"{name}_run.npz")
    V = np.concatenate([g["step0_V1"], g["step0_V2"], g["step0_V3"]])
    c.upload("V", V)
    assert c.vrms() == pytest.approx(o.vrms(V), rel=1e-12)


def test_pg_and_nu_against_reference(pair):
    """PG_timestep and heat_flux from the reference's own step-0 state: the
    energy step reproduces the reference bit for bit."""
    name, p, c, o, T = pair
    g = golden(f"{name}_run.npz")
    V = np.concatenate([g["step0_V1"], g["step0_V2"], g["step0_V3"]])
    c.upload("T", T)
    c.upload("Tdot", np.zeros(c.nno(), np.float32))
    c.upload("V", V)
    nu = c.heat_flux()
    assert nu[0] == pytest.approx(g["steps"][0, 3], rel=1e-6)
    assert nu[1] == pytest.approx(g["steps"][0, 4], rel=1e-6)
    st = c.PG_timestep()
    assert np.float32(st["timestep"]) == np.float32(g["steps"][1, 2])
    assert np.array_equal(c.download("T"), g["pg1_T"])
    assert np.array_equal(c.download("Tdot"), g["pg1_Tdot"])


def test_tight_solve_fixed_point(pair, pyoracle):
    """At accuracy 1e-10 the GPU (8-colour smoother) and the restatement with
    the reference's lexicographic smoother reach the same discrete solution."""
    from citcom_amd.hip import Citcom
    name, p, c, o, T = pair
    import dataclasses
    q = dataclasses.replace(p, accuracy=1e-10)
    c2 = Citcom(q)
    c2.initialize(T)
    c2.general_stokes_solver()
    o2 = pyoracle.Oracle(q, smoother=0)
    o2.set_viscosity(o2.visc_from_T(T))
    F = o2.assemble_forces(o2.thermal_buoyancy(T))
    U, P, _ = o2.solve_stokes(np.zeros(3 * o2.nno()), np.zeros(o2.nel()), F)
    e_u, e_p = rel_l2(c2.download("U"), U), rel_l2(c2.download("P"), P)
    say(name, "tight solve rel L2 U", e_u, "P", e_p)
    assert e_u < 1e-8 and e_p < 1e-8
    c2.close()


def test_graph_replay_matches_eager():
    """Each multigrid call replays one captured HIP graph; a full Stokes solve
    (hundreds of replays, with host reads between them) must give the same
    bits as launching the kernels one by one."""
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    p = problem("tiny_tdep")
    T = deck.perturbed_temperature(p)
    out = []
    for graphs in (1, 0, 1):
        c = Citcom(p)
        c.set_option("graphs", graphs)
        c.initialize(T)
        st = c.general_stokes_solver()
        r = c.timestep()
        out.append((c.download("U"), c.download("P"), c.download("T"), st["mg_cycles"] + r["stokes"]["mg_cycles"]))
        c.close()
    say("graph replays per run:", out[0][3])
    for k in range(3):
        assert np.array_equal(out[0][k], out[1][k]) and np.array_equal(out[2][k], out[1][k])


def test_print_convergence(capfd):
    """print_convergence writes one AhatP line per Uzawa iteration and one
    residual line per inner solve (Stokes_flow_Incomp.c:501-526,
    General_matrix_functions.c:494-498), and changes no result."""
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    p = problem("tiny_tdep")
    T = deck.perturbed_temperature(p)
    res = []
    for flag in (0, 1):
        c = Citcom(p)
        c.set_option("print_convergence", flag)
        c.initialize(T)
        capfd.readouterr()
        st = c.general_stokes_solver()
        err = capfd.readouterr().err
        res.append((c.download("U"), st))
        c.close()
    lines = err.splitlines()
    st = res[1][1]
    assert sum(ln.startswith("AhatP (") for ln in lines) == st["iterations"]
    assert sum("residual (" in ln for ln in lines) == st["inner_solves"]
    last = [ln for ln in lines if ln.startswith("AhatP (")][-1]
    assert f"dp/p {st['dpressure']:9.3e}" in last
    assert np.array_equal(res[0][0], res[1][0])


@pytest.mark.parametrize("name", ["tiny_tdep", "iso16"])
@pytest.mark.parametrize("small", [1, 0])
def test_smoother_kernels_bitwise(name, small, pyoracle):
    """The three smoother implementations (whole level in LDS, one workgroup
    with the stiffness in L2, one launch per colour pass; iso16's levels use
    one each) give the restatement's 8-colour iterate bit for bit, and so do
    per-colour launches on every level (small_levels = 0)."""
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    p = problem(name)
    c = Citcom(p)
    c.set_option("small_levels", small)
    o = pyoracle.Oracle(p, smoother=1)
    T = deck.perturbed_temperature(p)
    c.initialize(T)
    c.get_system_viscosity()
    c.construct_stiffness_B_matrix()
    o.set_viscosity(o.visc_from_T(T))
    inp = refdump.harness_op_inputs(p, 11, o.strip)
    for lev in range(p.levels):
        x = inp[lev]
        d, Ad = c.apply("gs0", lev, None, x["F"], sweeps=3)
        do, Ado = o.gauss_seidel(lev, np.zeros_like(x["F"]), x["F"], 3, 0)
        assert np.array_equal(d, do) and np.array_equal(Ad, Ado), (lev, "gs0")
        d, Ad = c.apply("gs1", lev, x["d0"], x["F"], sweeps=2)
        do, Ado = o.gauss_seidel(lev, x["d0"], x["F"], 2, 1)
        assert np.array_equal(d, do) and np.array_equal(Ad, Ado), (lev, "gs1")
    if small:
        # the coarsest level's smoother call of multi_grid, as a matrix
        F = inp[0]["F"]
        assert np.array_equal(c.apply("coarse", 0, None, F), o.coarse_solve(F))
        d, _ = o.gauss_seidel(0, np.zeros_like(F), F, p.v_steps_low, 0)
        assert rel_l2(c.apply("coarse", 0, None, F), d) < 1e-6
        if 3 * c.nno(1) <= 1500 and p.levels > 2:
            # the level above, as an affine map (with and without a guess)
            x = inp[1]
            assert np.array_equal(c.apply("affine0", 1, None, x["F"]), o.affine_solve(None, x["F"]))
            assert np.array_equal(c.apply("affine1", 1, x["d0"], x["F"]), o.affine_solve(x["d0"], x["F"]))
            d, _ = o.gauss_seidel(1, x["d0"], x["F"], p.down_heavy, 1)
            assert rel_l2(c.apply("affine1", 1, x["d0"], x["F"]), d) < 1e-6
    say(name, "small_levels", small, "smoother bitwise on", p.levels, "levels")
    c.close()


@pytest.mark.parametrize("mixed", [0, 1, 2])
@pytest.mark.parametrize("name", ["busse1a", "tdep64"])
def test_one_workgroup_smoother_same_bits(name, mixed):
    """A whole gauss_seidel call in one workgroup (unknowns in LDS, stiffness
    read through L2: busse1a's 405-node level, tdep64's 729-node level) gives
    the per-colour launches' iterate bit for bit, in the FP64,
    FP32 and FP16-stiffness forms: the step-0 solve and two timesteps agree
    exactly with the kernel on and off."""
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    p = problem(name)
    runs = []
    for wg in (1, 0):
        c = Citcom(p, mixed=mixed)
        c.set_option("gs_wg", wg)
        c.initialize(deck.seeded_temperature(p, seed=3))
        st = c.general_stokes_solver()
        for _ in range(2):
            r = c.timestep(with_nu=False)
        runs.append((c.download("U"), c.download("P"), c.download("T"), st["mg_cycles"], r["stokes"]["mg_cycles"]))
        c.close()
    say(name, "mixed", mixed, "one-workgroup smoother on/off: MG calls", runs[0][3:], runs[1][3:])
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mixed", [0, 1, 2])
@pytest.mark.parametrize("name", ["tdep128", "tdep64"])
def test_one_xcd_smoother_same_bits(name, mixed):
    """A whole gauss_seidel call as one launch of a few workgroups on one XCD
    (k_gs_xcd: tdep128's 2601-node level, tdep64's 4913-node level; colour
    passes separated by barriers through that XCD's L2) gives the per-colour
    launches' iterate bit for bit, in the FP64, FP32 and FP16-stiffness
    forms: the step-0 solve (Uzawa capped at 20 iterations) and two
    timesteps agree exactly with the kernel on and off.  In the FP32 forms
    the box smoother, which would otherwise take tdep64's level, and the
    exact coarse level, which would otherwise take tdep128's, are off: the
    one-XCD kernel smooths those levels, and the runs do not depend on
    whether rocSOLVER's factorization of the exact level gives the same
    last bits twice (once, at tdep128 with the FP16 stiffness, two runs
    differed in U's last bits with the same multigrid calls)."""
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    runs = []
    for on in (1, 0):
        p = problem(name)
        p.p_iterations = 20
        c = Citcom(p, mixed=mixed)
        if mixed:
            c.set_option("box_gs", 0)
            c.set_option("coarse_direct", 0)
        c.set_option("xcd_gs", on)
        c.initialize(deck.seeded_temperature(p, seed=5))
        st = c.general_stokes_solver()
        for _ in range(2):
            r = c.timestep(with_nu=False)
        runs.append((c.download("U"), c.download("P"), c.download("T"), st["mg_cycles"], r["stokes"]["mg_cycles"]))
        c.close()
    say(name, "mixed", mixed, "one-XCD smoother on/off: MG calls", runs[0][3:], runs[1][3:])
    for a, b in zip(runs[0], runs[1]):
        assert np.array_equal(a, b)


def test_one_xcd_smoother_calls():
    """Single smoother calls on tdep128's 2601-node level through the
    one-XCD kernel, with and without an initial guess and with 3 and 2
    sweeps in turn (the launches' barrier counter carries over from call to
    call): d and K d equal the per-colour launches' bit for bit."""
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    p = problem("tdep128")
    c = Citcom(p)
    c.initialize(deck.seeded_temperature(p, seed=2))
    c.get_system_viscosity()
    c.construct_stiffness_B_matrix()
    lev = 2
    assert c.nno(lev) == 2601
    rng = np.random.default_rng(4)
    n3 = 3 * c.nno(lev)
    for call in range(6):
        F = c.apply("strip", lev, rng.standard_normal(n3))
        d0 = c.apply("strip", lev, rng.standard_normal(n3))
        op, sweeps = ("gs1", 3) if call % 2 else ("gs0", 2 + call % 3)
        out = []
        for on in (1, 0):
            c.set_option("xcd_gs", on)
            out.append(c.apply(op, lev, d0 if op == "gs1" else None, F, sweeps=sweeps))
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1]), (call, op, sweeps)
    say("one-XCD smoother calls on the 2601-node level: 6 calls bit for bit")
    c.close()


def test_unsupported_problem_is_refused():
    """Viscosity laws other than rheol 0, 3 and 11, and unknown precision
    forms, are refused at setup, not computed as something else."""
    import dataclasses
    from citcom_amd.hip import ChipError, Citcom
    p = dataclasses.replace(problem("tiny_tdep"), rheol=1)
    with pytest.raises(ChipError, match="rheol 1 is not supported"):
        Citcom(p)
    with pytest.raises(ChipError, match="mixed must be 0, 1 or 2"):
        Citcom(problem("tiny_tdep"), mixed=3)


@pytest.mark.parametrize("name", ["tiny_tdep", "iso16@1e-8"])
def test_averages_against_reference(name):
    """the .ave layer profiles (Have.T from heat_flux, Have.vrms and Have.Vi
    from averages, Process_velocity.c:179-224) on the device: from the
    reference's own final state (its T and V uploaded, the viscosity formed
    from T), against the profiles the reference computed from it after the
    same 3-step run"""
    from citcom_amd.hip import Citcom
    p = problem(name)
    tag = name.replace("@", "_at_")
    ref = golden(f"{tag}_ave.npz")
    c = Citcom(p)
    c.upload("T", ref["final_T"])
    c.upload("V", np.concatenate([ref["final_V1"], ref["final_V2"], ref["final_V3"]]))
    c.get_system_viscosity()
    got = c.averages()
    for k in ("T", "vrms", "Vi"):
        e = rel_l2(got[k], ref[k])
        say(f"{name}: Have.{k} against the reference: {e:.1e}")
        assert e < 1e-6, (k, e)
    # the same profiles at the end of the GPU's own 3-step run
    c.close()
    c = Citcom(p)
    c.initialize(initial_temperature(name, p))
    c.general_stokes_solver()
    for _ in range(3):
        c.timestep()
    got = c.averages()
    for k in ("T", "vrms", "Vi"):
        e = rel_l2(got[k], ref[k])
        say(f"{name}: Have.{k} after the GPU's own 3 steps: {e:.1e}")
        assert e < (1e-6 if "@" in name else 1e-5), (k, e)
    c.close()


def test_apply_refuses_short_inputs():
    """chip_apply copies a fixed number of doubles from its host inputs: the
    Python mirror refuses missing or short arrays before the call"""
    from citcom_amd.hip import ChipError, Citcom
    p = problem("tiny")
    c = Citcom(p)
    n3 = 3 * c.nno()
    for op, kw in (("gs0", {}), ("gs1", {"x": np.zeros(n3)}), ("matvec", {"x": np.zeros(n3 - 1)}),
                   ("grad", {"x": np.zeros(c.nel() + 1)}), ("coarse", {"x2": np.zeros(5)})):
        with pytest.raises(ChipError, match="needs"):
            c.apply(op, c.levmax if op != "coarse" else 0, **kw)
    c.close()


def test_large_level_operators_bitwise(pyoracle):
    """the kernels of levels above 20000 nodes (the three-wavefront matvec,
    the colour passes) against the restatement, bit for bit, on the 64^3
    workload's two finest levels (274625 and 35937 nodes)"""
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    p = problem("tdep64")
    T = deck.seeded_temperature(p, seed=1)
    o = pyoracle.Oracle(p, smoother=1)
    o.set_viscosity(o.visc_from_T(T))
    c = Citcom(p)
    c.upload("T", T)
    c.get_system_viscosity()
    c.construct_stiffness_B_matrix()
    rng = np.random.default_rng(5)
    for lev in (p.levmax, p.levmax - 1):
        n3 = 3 * c.nno(lev)
        u = o.strip(lev, rng.standard_normal(n3))
        assert np.array_equal(c.apply("matvec", lev, u), o.matvec(lev, u)), lev
        F = o.strip(lev, rng.standard_normal(n3))
        d, Ad = c.apply("gs0", lev, None, F, sweeps=1)
        do, Ado = o.gauss_seidel(lev, np.zeros(n3), F, 1, 0)
        assert np.array_equal(d, do) and np.array_equal(Ad, Ado), lev
    c.close()
