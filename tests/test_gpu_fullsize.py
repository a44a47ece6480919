"""The measured workloads at full size, where no CPU run of the reference or
the restatement finishes inside a test: tdep64 (64^3 elements, 5 levels,
275 k nodes), tdep128 (128x128x64 elements, 6 levels, 1.08 M nodes) and
layer256 (256x256x128 elements, 7 levels, 8.5 M nodes).  Checked through
properties that hold at any size:

- the stiffness of every level is symmetric (u.Kv = v.Ku on stripped random
  vectors, to the rounding of its FP32 coefficients) and positive (u.Ku > 0);
- grad is the transpose of div (p.div u = grad p.u);
- after the step-0 Stokes solve the momentum residual F - K U - grad P is at
  the inner solves' tolerance relative to F, and the Uzawa monitors
  (incompressibility, dv/v) are what the loop stopped on;
- the plain, mixed and FP16-stiffness forms reach the same velocity and
  pressure, to the rounding of their cycles (FP32) or to the spread of a
  solve stopped at the Uzawa cap before its pressure has converged (FP16).
"""
import numpy as np
import pytest

from conftest import golden, problem, rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def start(name, mixed=0, piterations=None):
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    p = problem(name)
    if piterations is not None:
        p.p_iterations = piterations
    c = Citcom(p, mixed=mixed)
    c.initialize(deck.seeded_temperature(p, seed=1))
    return p, c


@pytest.mark.parametrize("name", ["tdep64", "tdep128", "layer256"])
def test_fullsize_operator_properties(name):
    say("deck", name, "operators")
    p, c = start(name)
    c.get_system_viscosity()
    c.construct_stiffness_B_matrix()
    rng = np.random.default_rng(11)
    for lev in range(p.levels):
        n3 = 3 * c.nno(lev)
        u = c.apply("strip", lev, rng.standard_normal(n3))
        v = c.apply("strip", lev, rng.standard_normal(n3))
        Ku, Kv = c.apply("matvec", lev, u), c.apply("matvec", lev, v)
        a, b = float(u @ Kv), float(v @ Ku)
        assert abs(a - b) <= 1e-6 * (abs(u) @ np.abs(Kv)), (lev, a, b)
        assert float(u @ Ku) > 0.0, lev
    lev = p.levmax
    u = c.apply("strip", lev, rng.standard_normal(3 * c.nno(lev)))
    pr = rng.standard_normal(c.nel(lev))
    lhs, rhs = float(pr @ c.apply("div", lev, u)), float(c.apply("grad", lev, pr) @ u)
    assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + abs(rhs))
    c.close()


def momentum_residual(c):
    lev = c.levmax
    F = c.download("F")
    r = F - c.apply("matvec", lev, c.download("U")) - c.apply("grad", lev, c.download("P"))
    r = c.apply("strip", lev, r)
    return float(np.linalg.norm(r) / np.linalg.norm(F))


@pytest.mark.parametrize("name,piterations", [("tdep64", None), ("tdep128", None), ("layer256", 10)])
def test_fullsize_forms_agree(name, piterations):
    out = {}
    for mixed in (0, 1, 2):
        p, c = start(name, mixed, piterations)
        st = c.general_stokes_solver()
        U, P = c.download("U"), c.download("P")
        rm = momentum_residual(c)
        say(f"{name} mixed={mixed}: {st['iterations']} Uzawa its, {st['mg_cycles']} MG calls, dv/v "
            f"{st['dvelocity']:.2e} dp/p {st['dpressure']:.2e} div/v {st['incompressibility']:.2e}, "
            f"momentum residual {rm:.2e}")
        out[mixed] = (U, P - P.mean(), st)
        c.close()
        assert rm < 1e-4   # measured 4.6e-6 (tdep128), 1.1e-6 (layer256)
        assert st["iterations"] == p.p_iterations or (st["dvelocity"] < p.accuracy and st["dpressure"] < p.accuracy)
    # The FP32 cycles' preconditioner differs from the plain form's (the
    # levels between the exact one and the finest smoothed by boxes; where
    # there is one, the 17x17x9-node level solved exactly; FP16 stiffness
    # with mixed = 2), so at the 375-iteration cap, where no form has
    # converged the pressure (dp/p 4-5e-4), the solve stops at another
    # iterate: U 3.3e-6, P 6.3e-4 (FP32, tdep64); U 5.6e-6, P 1.5e-3 (FP32)
    # and U 6.6e-6, P 7.0e-3 (FP16) at tdep128.  With the colour passes on
    # every level (option box_gs 0) and no exact level, the FP32 cycles give
    # the plain form's iterates to rounding (tdep64: U 1.0e-6, P 3.6e-5).
    bounds = {1: (1e-4, 3e-2), 2: (1e-4, 3e-2)}
    for mixed in (1, 2):
        eu, ep = rel_l2(out[mixed][0], out[0][0]), rel_l2(out[mixed][1], out[0][1])
        say(f"{name}: mixed={mixed} against plain: U {eu:.2e}, P {ep:.2e}")
        assert eu < bounds[mixed][0] and ep < bounds[mixed][1]


def test_tdep64_trajectory_forms():
    """The 64^3-element workload (Ra 1e6, contrast 1e4, seeded noise): 20
    steps of the plain FP64 form and of the FP32 refinement form, each on its
    own trajectory from the same state; dt and Nu_top per step and the final
    fields against each other (the per-field error of the FP32 cycles: the
    FP32 form with the colour passes on every level, option box_gs 0, so
    that the arithmetic is the only difference)"""
    runs = {}
    for mixed in (0, 1):
        p, c = start("tdep64", mixed)
        if mixed:
            c.set_option("box_gs", 0)
        c.general_stokes_solver()
        rec = [c.timestep() for _ in range(20)]
        runs[mixed] = {"dt": np.array([r["pg"]["timestep"] for r in rec]),
                       "nu": np.array([r["Nu_top"] for r in rec]),
                       "its": [r["stokes"]["iterations"] for r in rec],
                       "U": c.download("U"), "P": c.download("P"), "T": c.download("T"), "vrms": c.vrms()}
        c.close()
    a, b = runs[0], runs[1]
    e_dt = float(np.abs(b["dt"] / a["dt"] - 1).max())
    e_nu = float(np.abs(b["nu"] / a["nu"] - 1).max())
    eu, et = rel_l2(b["U"], a["U"]), rel_l2(b["T"], a["T"])
    ep = rel_l2(b["P"] - b["P"].mean(), a["P"] - a["P"].mean())
    say(f"tdep64 20 steps, FP32 refinement against plain FP64: dt {e_dt:.1e}, Nu_top {e_nu:.1e}, U {eu:.1e}, "
        f"P {ep:.1e}, T {et:.1e}, vrms {abs(b['vrms'] / a['vrms'] - 1):.1e}; Uzawa its plain {a['its']}, "
        f"mixed {b['its']}")
    assert e_dt < 1e-5 and e_nu < 1e-5
    assert et < 1e-6 and eu < 1e-5 and ep < 1e-3


def test_tdep64_against_reference():
    """The 64^3-element workload against the reference run at accuracy 1e-9
    (tests/golden/tdep64_at_1e-9_run.npz: the reference under MPI, 8 ranks;
    2258 and 2709 Uzawa iterations): the step-0 solve and step 1 (dt, Nu,
    the fields after it).  Fields on every other node and element, as the
    fixture keeps them; the pressure up to its free-slip constant.

    On 8 ranks the reference's Gauss-Seidel sweeps each subdomain and
    exchanges the boundary sums (General_matrix_functions.c:1273, :1336), so
    its iterates differ from the one-domain sweep, and the relative-increment
    criterion (Stokes_flow_Incomp.c:442-443) stops the two loops at different
    points of the slowly converging pressure modes.  Measured: this build's
    step-0 pressure at 1e-9 is 2.2e-6 from its own solve run to 1e-11, and
    the reference's is 7.9e-6 from that solve (8.7e-6 from ours at 1e-9);
    the velocity agrees to 4.4e-8.  Hence the pressure bound of 2e-5 at
    step 0 against 1e-6 for the velocity."""
    from citcom_amd import deck
    from citcom_amd.hip import Citcom
    g = golden("tdep64_at_1e-9_run.npz")
    p = problem("tdep64@1e-9")
    p.p_iterations = 3000
    nox, noy, noz = p.dims()
    c = Citcom(p)
    c.initialize(deck.seeded_temperature(p, seed=1))

    def sub(U, P):
        u2 = U.reshape(noy, nox, noz, 3)[::2, ::2, ::2].reshape(-1)
        p2 = P.reshape(noy - 1, nox - 1, noz - 1)[::2, ::2, ::2].reshape(-1)
        return u2, p2

    def pdiff(a, b):
        return rel_l2(a - a.mean(), b - b.mean())
    st = c.general_stokes_solver()
    u2, p2 = sub(c.download("U"), c.download("P"))
    eu, ep = rel_l2(u2, g["step0_U2"]), pdiff(p2, g["step0_P2"])
    say(f"tdep64 step 0 at 1e-9: {st['iterations']} Uzawa iterations (reference 2258), U {eu:.1e}, P {ep:.1e} "
        f"(P with its constant {rel_l2(p2, g['step0_P2']):.1e})")
    assert eu < 1e-6 and ep < 2e-5
    r = c.timestep()
    # the harness records heat_flux after the step's Stokes solve
    # (ref_harness.c, run_steps); timestep() returns the one the time loop
    # computes before it (process_temp_field, Citcom.c:129), from the
    # previous velocity
    nu = c.heat_flux()
    ref = g["steps"][1]
    u2, p2 = sub(c.download("U"), c.download("P"))
    t2 = c.download("T").reshape(noy, nox, noz)[::2, ::2, ::2].reshape(-1)
    e = {"dt": abs(r["pg"]["timestep"] / ref[2] - 1), "Nu_top": abs(nu[0] / ref[3] - 1),
         "Nu_bot": abs(nu[1] / ref[4] - 1), "U": rel_l2(u2, g["final_U2"]), "P": pdiff(p2, g["final_P2"]),
         "T": rel_l2(t2, g["final_T2"])}
    say(f"tdep64 step 1 at 1e-9: {r['stokes']['iterations']} Uzawa iterations (reference 2709); " +
        ", ".join(f"{k} {v:.1e}" for k, v in e.items()))
    c.close()
    # measured: dt 5.5e-11, T 7.7e-10, U 9.6e-8, P 1.7e-6 (the step-0
    # pressure spread carried into step 1's start), Nu_top 4e-9, Nu_bot 3.3e-6
    # (its advective part is large: |Vz| reaches 1e4 in the bottom layers)
    assert e["dt"] < 1e-6 and e["T"] < 1e-6 and e["U"] < 1e-6 and e["P"] < 1e-5
    assert e["Nu_top"] < 1e-5 and e["Nu_bot"] < 1e-5


def test_direct_coarse_level():
    """The FP32 cycles' exact solve on the 17x17x9-node level of tdep128
    (the level's stiffness factored and inverted in FP64 by rocSOLVER, the
    inverse applied in FP32): K x = F for random right-hand sides to the
    FP32 inverse's rounding; and the mixed forms' step-0 Stokes solve with
    it on and off reaches the same velocity and pressure"""
    p, c = start("tdep128", mixed=1)
    c.get_system_viscosity()
    c.construct_stiffness_B_matrix()
    lev = 2
    assert c.nno(lev) == 17 * 17 * 9
    rng = np.random.default_rng(7)
    worst = 0.0
    for _ in range(3):
        F = c.apply("strip", lev, rng.standard_normal(3 * c.nno(lev)))
        x = c.apply("direct", lev, None, F)
        r = c.apply("strip", lev, c.apply("matvec", lev, x) - F)
        worst = max(worst, float(np.linalg.norm(r) / np.linalg.norm(F)))
    say(f"direct solve on level {lev}: |K x - F| / |F| = {worst:.2e}")
    c.close()
    assert worst < 2e-3   # measured 4.6e-4 (the inverse applied in FP16)
    # measured: 505 against 510 multigrid calls (FP32), 515 against 523
    # (FP16 stiffness); U 5.6e-6 and 6.8e-6, P 1.5e-3 and 6.5e-3 apart at
    # the 375-iteration cap, where no form has converged the pressure
    out = {}
    for mixed in (1, 2):
        for on in (1, 0):
            p, c = start("tdep128", mixed)
            c.set_option("coarse_direct", on)
            st = c.general_stokes_solver()
            out[(mixed, on)] = (c.download("U"), c.download("P"), st)
            say(f"tdep128 mixed={mixed} coarse_direct={on}: {st['iterations']} Uzawa its, {st['mg_cycles']} MG calls")
            c.close()
        (U1, P1, s1), (U0, P0, s0) = out[(mixed, 1)], out[(mixed, 0)]
        eu, ep = rel_l2(U1, U0), rel_l2(P1 - P1.mean(), P0 - P0.mean())
        say(f"tdep128 mixed={mixed}: direct against the full hierarchy: U {eu:.2e}, P {ep:.2e}")
        assert eu < 1e-4 and ep < 3e-2
