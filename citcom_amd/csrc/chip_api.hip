// C ABI and host orchestration of libcitcom_hip: the Uzawa loop
// (solve_Ahat_p_fhat), the full multigrid (solve_del2_u / multi_grid), the
// SUPG step (PG_timestep) and the per-step stiffness rebuild, all on the
// device.  The host only keeps loop control: one 8-byte read per multigrid
// call (the residual test) and one small read per Uzawa iteration (the
// convergence monitors), as the reference's loops need them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "chip_kernels.h"
#include "../../include/citcom_learn.h"

namespace chip {

static thread_local std::string g_err;
void set_error(const std::string &s) { g_err = s; }

namespace {

// slots of chip_ctx::dots
enum : int {
    D_TMP = 0,     // [0..3] scratch reductions
    D_MG = 4,      // [4] AU.AU, [5] AU.res
    D_R1Z1 = 6,    // [6], [7] alternate between Uzawa iterations
    D_S2AH = 8,
    D_VV = 9,      // [9] V.V, [10] u1.u1
    D_PP = 11,     // [11] P.P, [12] Ah.Ah, [13] s2.s2
    D_INIT = 14,   // [14] r1.r1, [15] V.V
    D_WS = 16,     // [16..19] warm-start feature scales
    D_HF = 24,     // [24..27] heat-flux sums
    D_XST = 28,    // [28] status of the one-XCD smoother launches (nonzero: one stopped early)
};

inline hipStream_t S(const chip_ctx *c) { return c->stream; }

int sync_dots(chip_ctx *c) {
    CHIP_CHECK(hipMemcpyAsync(c->h_dots, c->dots, 32 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CHIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

inline float unford(unsigned u) {
    const unsigned b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

int sync_minmax(chip_ctx *c, float *mn, float *mx) {
    CHIP_CHECK(hipMemcpyAsync(c->h_minmax, c->minmax, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    CHIP_CHECK(hipStreamSynchronize(c->stream));
    if (mn) *mn = unford(c->h_minmax[0]);
    if (mx) *mx = unford(c->h_minmax[1]);
    return 0;
}

// ---- profiling of one kernel family -------------------------------------
// When the profiled family matches, a and b are handed to the launch, which
// stamps them with the kernel's own start and end (hipExtLaunchKernelGGL).
struct ProfGuard {
    hipEvent_t a = nullptr, b = nullptr;
    ProfGuard(chip_ctx *c, int kernel, double bytes, int lev) {
        if (c->prof.kernel != kernel || (c->prof.level >= 0 && c->prof.level != lev)) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            a = b = nullptr;
            return;
        }
        c->prof.ev.emplace_back(a, b);
        c->prof.bytes.push_back(bytes);
    }
};

int prof_flush(chip_ctx *c) {
    Prof &pr = c->prof;
    if (pr.ev.empty()) return 0;
    CHIP_CHECK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < pr.ev.size(); i++) {
        float ms = 0.0f;
        CHIP_CHECK(hipEventElapsedTime(&ms, pr.ev[i].first, pr.ev[i].second));
        pr.total_ms += ms;
        pr.total_bytes += pr.bytes[i];
        pr.launches++;
        (void)hipEventDestroy(pr.ev[i].first);
        (void)hipEventDestroy(pr.ev[i].second);
    }
    pr.ev.clear();
    pr.bytes.clear();
    return 0;
}

// algorithmic bytes: 243 coefficients per node (FP32, or FP16 where the
// level has the scaled FP16 stiffness and the cycle runs in FP32) plus the
// vectors (matvec: read u, write Au; colour pass: read/write d, read F and BI)
template <typename V>
double coef_bytes(const Level &L) { return (sizeof(V) == 4 && L.Kh) ? 243.0 * 2 : 243.0 * 4; }
template <typename V>
double mv_bytes(const Level &L) { return coef_bytes<V>(L) + 6 * sizeof(V); }
template <typename V>
double gs_bytes(const Level &L) { return coef_bytes<V>(L) + 12 * sizeof(V); }

// the multigrid work vectors of a level in one precision
template <typename V>
struct LV {
    V *vel, *res, *fl, *del, *AU;
    const V *BI;
};
template <typename V>
LV<V> vecs(const Level &L);
template <>
LV<double> vecs<double>(const Level &L) { return {L.vel, L.res, L.fl, L.del, L.AU, L.BI}; }
template <>
LV<float> vecs<float>(const Level &L) { return {L.velf, L.resf, L.flf, L.delf, L.AUf, L.BIf}; }

template <typename V>
void do_matvec(chip_ctx *c, int lev, const V *u, V *Au, const V *rhs = nullptr, V *res = nullptr,
               V *ucopy = nullptr) {
    const Level &L = c->L[lev];
    ProfGuard pg(c, 2, mv_bytes<V>(L) * L.g.nno, lev);
    matvec(S(c), L, u, Au, rhs, res, pg.a, pg.b, ucopy);
}

// gauss_seidel (General_matrix_functions.c:1160-1361), 8-colour; returns Ad = K d
// (with res: also res = rhs - Ad, formed by the final matvec).  guess: 1
// start from d, 0 from zero, 2 from zero with d already zero (no fill)
template <typename V>
void do_gs(chip_ctx *c, int lev, V *d, const V *F, V *Ad, int sweeps, int guess, const V *rhs = nullptr,
           V *res = nullptr);

// a smoother call of multi_grid: the affine form where the level has one
template <typename V>
void mg_gs(chip_ctx *c, int lev, V *d, const V *F, V *Ad, int sweeps, int guess, const V *rhs = nullptr,
           V *res = nullptr) {
    const Level &L = c->L[lev];
    if (L.Ma && c->small_levels && sweeps == L.affine_sweeps) {
        V *out;
        if constexpr (sizeof(V) == sizeof(double)) out = L.tmp;
        else out = L.tmpf;
        affine_apply(S(c), L, guess == 1 ? (const V *)d : nullptr, F, out);
        do_matvec(c, lev, (const V *)out, Ad, rhs, res, d);   // and d = out
        return;
    }
    do_gs(c, lev, d, F, Ad, sweeps, guess, rhs, res);
}

template <typename V>
void do_gs(chip_ctx *c, int lev, V *d, const V *F, V *Ad, int sweeps, int guess, const V *rhs, V *res) {
    const Level &L = c->L[lev];
    if constexpr (sizeof(V) == sizeof(double)) {
        if (c->small_levels && gs_small_kind(L)) {
            gs_small(S(c), L, d, F, Ad, sweeps, guess);
            if (res) sub(S(c), res, rhs, Ad, 3L * L.g.nno);
            return;
        }
    }
    const V *BI = vecs<V>(L).BI;
    if (c->small_levels && c->xcd_gs && sweeps > 0 && gs_xcd_ok(L)) {
        ProfGuard pg(c, 4, gs_bytes<V>(L) * L.g.nno * sweeps + mv_bytes<V>(L) * L.g.nno, lev);
        gs_xcd(S(c), L, BI, d, F, Ad, rhs, res, sweeps, guess, c->dots + D_XST, pg.a, pg.b);
        return;
    }
    if (c->small_levels && c->gs_wg && gs_wg_ok<V>(L)) {
        ProfGuard pg(c, 3, gs_bytes<V>(L) * L.g.nno * sweeps + mv_bytes<V>(L) * L.g.nno, lev);
        gs_wg(S(c), L, BI, d, F, Ad, rhs, res, sweeps, guess, pg.a, pg.b);
        return;
    }
    if constexpr (sizeof(V) == 4) {
        if (c->box_gs && L.Kbox && sweeps > 0) {
            // sweeps alternate between d and the level's scratch; a sweep
            // from zero reads nothing, so it can write d itself
            const float *src = guess == 1 ? d : nullptr;
            float *dst = guess == 1 ? L.tmpf : d;
            for (int cyc = 0; cyc < sweeps; cyc++) {
                ProfGuard pg(c, 5, gs_bytes<V>(L) * L.g.nno, lev);
                gs_box(S(c), L, BI, src, dst, F, pg.a, pg.b);
                src = dst;
                dst = dst == d ? L.tmpf : d;
            }
            do_matvec(c, lev, (const V *)src, Ad, rhs, res, src == d ? nullptr : d);
            return;
        }
    }
    if (guess == 0) fill(S(c), d, V(0), 3L * L.g.nno);
    for (int cyc = 0; cyc < sweeps; cyc++)
        for (int col = 0; col < 8; col++) {
            ProfGuard pg(c, 1, gs_bytes<V>(L) * L.g.cn[col], lev);
            gs_colour(S(c), L, BI, col, d, F, pg.a, pg.b);
        }
    do_matvec(c, lev, (const V *)d, Ad, rhs, res);
}

// multi_grid (General_matrix_functions.c:525-653).  F holds the right-hand
// side on entry and the residual on exit; the correction is left in
// L[levmax].vel.  Coarse right-hand sides share storage with fl[] (rhs[lev]
// is a copy of fl[lev] in the reference, and fl[lev] is not read again).
// This enqueues the whole fixed kernel sequence; |F|^2 lands in dots[D_TMP].
// the coarsest-level smoother call of multi_grid (AU[levmin] is not read
// afterwards): the learned correction when it is switched on (a failed
// launch is reported by multi_grid), else the coarse matrix where there is
// one
template <typename V>
void coarse_smooth(chip_ctx *c, V *vel, const V *rhs) {
    const Level &L = c->L[c->levmin];
    if (c->cnet_on) {
        if (cnet_apply(S(c), c, L, rhs, vel)) c->cnet_failed = true;
        return;
    }
    if (L.Mc && c->small_levels)
        coarse_apply(S(c), L, rhs, vel);
    else
        do_gs(c, c->levmin, vel, rhs, vecs<V>(L).AU, c->p.v_steps_low, 0);
}

// the V-cycles of multi_grid in one precision; F is the finest right-hand
// side, the correction is left in the finest vel and the finest residual in
// its res
template <typename V>
void enqueue_cycles(chip_ctx *c, V *F) {
    const int levmax = c->levmax;
    // the FP32 cycles stop at the level solved exactly, where there is one
    const bool direct = sizeof(V) == 4 && c->direct_ok && !c->cnet_on;
    const int levmin = direct ? c->dlev : c->levmin;
    const chip_problem &p = c->p;
    Level *L = c->L;
    auto fl = [&](int lev) { return lev == levmax ? F : vecs<V>(L[lev]).fl; };
    auto bottom = [&](V *vel, const V *rhs) {
        if (direct)
            dense_apply(S(c), c->dAh, c->dAs, rhs, vel, 3 * L[levmin].g.nno);
        else
            coarse_smooth(c, vel, rhs);
    };
    for (int lev = levmax; lev > levmin; lev--) project(S(c), L[lev], L[lev - 1], (const V *)fl(lev), fl(lev - 1));
    bottom(vecs<V>(L[levmin]).vel, (const V *)fl(levmin));
    for (int lev = levmin + 1; lev <= levmax; lev++) {
        interp(S(c), L[lev - 1], L[lev], (const V *)vecs<V>(L[lev - 1]).vel, vecs<V>(L[lev]).vel);
        for (int Vn = 1; Vn <= p.mg_cycle; Vn++) {
            for (int dlev = lev; dlev >= levmin + 1; dlev--) {
                const LV<V> v = vecs<V>(L[dlev]);
                const int cycles = (dlev == levmax) ? p.v_steps_high : p.down_heavy;
                V *rhs = (dlev == lev) ? fl(dlev) : v.fl;
                // res = rhs - AU formed by the smoother's final matvec; below
                // the top of the V-cycle the smoother starts from zero, which
                // the restriction above it has written
                mg_gs(c, dlev, v.vel, (const V *)rhs, v.AU, cycles, dlev == lev ? 1 : 2, (const V *)rhs, v.res);
                project(S(c), L[dlev], L[dlev - 1], (const V *)v.res, vecs<V>(L[dlev - 1]).fl, true,
                        vecs<V>(L[dlev - 1]).vel);
            }
            bottom(vecs<V>(L[levmin]).vel, (const V *)vecs<V>(L[levmin]).fl);
            for (int ulev = levmin + 1; ulev <= lev; ulev++) {
                const LV<V> v = vecs<V>(L[ulev]);
                const int cycles = (ulev == levmax) ? p.v_steps_high : p.up_heavy;
                const long n = 3L * L[ulev].g.nno;
                interp(S(c), L[ulev - 1], L[ulev], (const V *)vecs<V>(L[ulev - 1]).vel, v.del);
                mg_gs(c, ulev, v.del, (const V *)v.res, v.AU, cycles, 1);
                V *resout = ulev == levmax ? v.res : nullptr;
                if (line_search_ok(n)) {
                    line_search(S(c), n, v.vel, (const V *)v.del, resout, (const V *)v.AU, (const V *)v.res,
                                c->dots + D_MG);
                    continue;
                }
                dots(S(c), c, (int)n, 2, (const V *)v.AU, (const V *)v.AU, v.AU, v.res, nullptr, nullptr, nullptr,
                     nullptr, c->dots + D_MG);
                mg_update(S(c), v.vel, v.del, resout, v.AU, c->dots + D_MG, n);
            }
        }
    }
}

// one multi_grid call on F (= c->sr, FP64): F holds the right-hand side on
// entry and the residual on exit; the correction is left in L[levmax].vel
// (FP64).  Plain form: the reference's FP64 cycles, residual from the line
// search.  Mixed form (FP32 multigrid inside FP64 refinement): the cycles run
// on an FP32 copy of F, the FP32 correction is added to the FP64 solution
// (u1, so vel is not used) and the residual F = Ah - K u1 is formed in FP64;
// solve_del2_u always solves K u1 = Ah.  |F|^2 lands in dots[D_TMP].
void enqueue_multi_grid(chip_ctx *c, double *F) {
    const Level &L = c->L[c->levmax];
    const long neq = 3L * L.g.nno;
    if (!c->p.mixed) {
        enqueue_cycles<double>(c, F);
        copy(S(c), F, L.res, neq);
    } else {
        to_f32(S(c), L.flf, F, neq);
        enqueue_cycles<float>(c, L.flf);
        add_f32(S(c), c->u1, L.velf, neq);
        do_matvec<double>(c, c->levmax, (const double *)c->u1, c->sd, (const double *)c->Ah, F);
    }
    dots(S(c), c, (int)neq, 1, (const double *)F, (const double *)F, nullptr, nullptr, nullptr, nullptr, nullptr,
         nullptr, c->dots + D_TMP);
}

// a k_gs_xcd launch that stopped early leaves a nonzero status in
// dots[D_XST] (read by the last sync_dots): reported as an error, after the
// status and the levels' barrier counters are cleared for later launches
int xcd_status(chip_ctx *c) {
    const double st = c->h_dots[D_XST];
    if (st == 0.0) return 0;
    (void)hipMemsetAsync(c->dots + D_XST, 0, sizeof(double), c->stream);
    for (int lev = c->levmin; lev <= c->levmax; lev++)
        if (c->L[lev].xbar) (void)hipMemsetAsync(c->L[lev].xbar, 0, sizeof(XcdBar), c->stream);
    (void)hipStreamSynchronize(c->stream);
    set_error(st == 2.0 ? "the one-XCD smoother's workgroups were not on one XCD (set option xcd_gs = 0)"
                        : "a barrier of the one-XCD smoother passed its poll limit (set option xcd_gs = 0)");
    return -1;
}

// F is always c->sr here, so one captured graph serves every call
int multi_grid(chip_ctx *c, double *F, double *residual) {
    if (c->use_graphs && !c->prof.kernel && F == c->sr) {
        if (!c->mg_graph) {
            hipGraph_t graph;
            CHIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            enqueue_multi_grid(c, F);
            CHIP_CHECK(hipStreamEndCapture(c->stream, &graph));
            if (c->cnet_failed) {
                (void)hipGraphDestroy(graph);
                c->cnet_failed = false;
                set_error("multi_grid: the learned coarse correction failed to launch");
                return -1;
            }
            CHIP_CHECK(hipGraphInstantiate(&c->mg_graph, graph, nullptr, nullptr, 0));
            CHIP_CHECK(hipGraphDestroy(graph));
        }
        CHIP_CHECK(hipGraphLaunch(c->mg_graph, c->stream));
    } else {
        enqueue_multi_grid(c, F);
        if (c->cnet_failed) {
            c->cnet_failed = false;
            set_error("multi_grid: the learned coarse correction failed to launch");
            return -1;
        }
    }
    if (sync_dots(c) || xcd_status(c)) return -1;
    *residual = std::sqrt(c->h_dots[D_TMP] / (3.0 * c->L[c->levmax].g.nno));
    return 0;
}

// the captured multigrid graph holds the coarsest level's smoother choice
int drop_mg_graph(chip_ctx *c) {
    if (c->mg_graph) {
        CHIP_CHECK(hipStreamSynchronize(S(c)));
        CHIP_CHECK(hipGraphExecDestroy(c->mg_graph));
        c->mg_graph = nullptr;
    }
    return 0;
}

// the multigrid loop of solve_del2_u: d0 += corrections until the residual
// reaches acc; *residual in: r0, out: the last residual
int mg_loop(chip_ctx *c, double *d0, const double *F, double acc, double r0, int *count, double *residual) {
    const long neq = 3L * c->L[c->levmax].g.nno;
    double res = r0, prev = r0;
    int cnt = 0;
    while (res > acc) {
        if (multi_grid(c, c->sr, &res)) return -1;
        if (!c->p.mixed) axpy_scalar(S(c), d0, c->L[c->levmax].vel, 1.0, neq);
        cnt++;
        // the plain multigrid is the fallback of the learned coarse
        // correction: a call that does not reduce the residual (or gives a
        // non-finite one) switches the correction off for the rest of the
        // run and restarts this solve with the coarsest level's smoother
        if (c->cnet_on && !(res < prev)) {
            c->cnet_on = false;
            c->cnet_fallbacks++;
            if (drop_mg_graph(c)) return -1;
            if (c->print_convergence)
                std::fprintf(stderr, " *  learned coarse correction off: residual %.3e after %.3e\n", res, prev);
            copy(S(c), c->sr, F, neq);
            fill(S(c), d0, 0.0, neq);
            res = prev = r0;
            cnt = 0;
            continue;
        }
        prev = res;
        if (cnt > 10000) {
            set_error("solve_del2_u: multigrid does not converge");
            return -1;
        }
    }
    *count = cnt;
    *residual = res;
    return 0;
}

// solve_del2_u (General_matrix_functions.c:368-520, multigrid branch):
// d0 = K^-1 F; F is not modified.  The first solve after the learned coarse
// correction is switched on is a calibration: it runs with the coarsest
// level's smoother, then again with the correction, and the correction
// stays on only if its mean residual reduction per multigrid call is the
// stronger; otherwise the plain solve's result is kept and the correction
// switched off (so the run continues exactly as the plain one would).
int solve_del2_u(chip_ctx *c, double *d0, const double *F, double acc, int *valid, int *count) {
    if (c->p.mixed && (d0 != c->u1 || F != c->Ah)) {
        set_error("solve_del2_u: the mixed form solves K u1 = Ah only");
        return -1;
    }
    const Level &L = c->L[c->levmax];
    const long neq = 3L * L.g.nno;
    copy(S(c), c->sr, F, neq);
    fill(S(c), d0, 0.0, neq);
    dots(S(c), c, (int)neq, 1, c->sr, c->sr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c->dots + D_TMP);
    if (sync_dots(c)) return -1;
    const double r0 = std::sqrt(c->h_dots[D_TMP] / neq);
    acc = std::max(acc, r0 * c->p.accuracy);
    *valid = (r0 < acc) ? 0 : 1;
    double residual = r0;
    int cnt = 0;
    if (c->cnet_on && !c->cnet_calibrated && r0 > acc) {
        c->cnet_calibrated = true;
        // plain first, its result kept in the finest level's scratch
        c->cnet_on = false;
        if (drop_mg_graph(c) || mg_loop(c, d0, F, acc, r0, &cnt, &residual)) return -1;
        const double f_plain = std::pow(residual / r0, 1.0 / std::max(cnt, 1));
        const int cnt_plain = cnt;
        double *keep = c->cnet_keep;
        copy(S(c), keep, d0, neq);
        // then with the correction
        c->cnet_on = true;
        copy(S(c), c->sr, F, neq);
        fill(S(c), d0, 0.0, neq);
        if (drop_mg_graph(c) || mg_loop(c, d0, F, acc, r0, &cnt, &residual)) return -1;
        const double f_net = std::pow(residual / r0, 1.0 / std::max(cnt, 1));
        if (c->cnet_on && !(f_net < f_plain)) {   // not the stronger: back to the plain run's result
            c->cnet_on = false;
            c->cnet_fallbacks++;
            if (drop_mg_graph(c)) return -1;
            copy(S(c), d0, keep, neq);
            cnt = cnt_plain;
        }
        if (c->print_convergence)
            std::fprintf(stderr, " *  learned coarse correction calibration: reduction per multigrid call %.3e "
                                 "(plain %.3e): %s\n", f_net, f_plain, c->cnet_on ? "kept" : "switched off");
    } else if (mg_loop(c, d0, F, acc, r0, &cnt, &residual)) {
        return -1;
    }
    if (c->print_convergence)
        std::fprintf(stderr, " *  residual (%03d) = %.3e from %.3e to %.3e\n", cnt, residual, r0, acc);
    *count = cnt;
    return 0;
}

// solve_Ahat_p_fhat (Stokes_flow_Incomp.c:295-497) + v_from_vector (:530-552)
int uzawa(chip_ctx *c, chip_stokes_stats *st) {
    const Level &L = c->L[c->levmax];
    const long neq = 3L * L.g.nno, npno = L.g.nel;
    const int gneq = (int)neq, gnpno = (int)npno;
    const double imp = c->p.accuracy;
    chip_stokes_stats s{};
    hipStream_t q = S(c);

    dots(q, c, gneq, 1, c->F, c->F, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c->dots + D_TMP);
    if (sync_dots(c)) return -1;
    const double v_res = std::sqrt(c->h_dots[D_TMP] / gneq);

    grad(q, L, c->P, c->Ah);
    do_matvec(c, c->levmax, c->U, c->u1);
    init_residual(q, c->Ah, c->F, c->Ah, c->u1, neq);
    strip(q, L, c->Ah);
    int valid = 0, cnt = 0;
    if (solve_del2_u(c, c->u1, c->Ah, imp * v_res, &valid, &cnt)) return -1;
    s.mg_cycles += cnt;
    s.inner_solves++;
    strip(q, L, c->u1);
    axpy_scalar(q, c->U, c->u1, 1.0, neq);
    div(q, L, c->U, c->r1);
    dots(q, c, gnpno, 1, c->r1, c->r1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c->dots + D_INIT);
    dots(q, c, gneq, 1, c->U, c->U, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c->dots + D_INIT + 1);
    if (sync_dots(c)) return -1;
    s.pdotp = std::sqrt(c->h_dots[D_INIT] / gnpno);
    s.vdotv = std::sqrt(c->h_dots[D_INIT + 1] / gneq);
    s.incompressibility = s.pdotp / s.vdotv;

    int count = 0;
    float dpressure = 1.0f, dvelocity = 1.0f;
    double *r0 = c->r0, *r1 = c->r1, *r2 = c->r2, *z0 = c->z0, *z1 = c->z1, *s1 = c->s1, *s2 = c->s2, *sh;
    while ((count < c->p.p_iterations) && (dpressure >= imp || dvelocity >= imp)) {
        double *r1z1 = c->dots + D_R1Z1 + (count & 1), *r0z0 = c->dots + D_R1Z1 + ((count + 1) & 1);
        uz_z(q, z1, L.BPI, r1, npno);
        dots(q, c, gnpno, 1, r1, z1, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, r1z1);
        uz_search(q, s2, z1, s1, r1z1, r0z0, count == 0, npno);
        grad(q, L, s2, c->Ah);
        if (solve_del2_u(c, c->u1, c->Ah, imp * v_res, &valid, &cnt)) return -1;
        s.mg_cycles += cnt;
        s.inner_solves++;
        strip(q, L, c->u1);
        div(q, L, c->u1, c->Ah);
        dots(q, c, gnpno, 1, s2, c->Ah, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c->dots + D_S2AH);
        uz_pupdate(q, r2, r1, c->Ah, c->P, s2, r1z1, c->dots + D_S2AH, valid, npno);
        uz_vupdate(q, c->U, c->u1, r1z1, c->dots + D_S2AH, valid, neq);
        div(q, L, c->U, c->Ah);
        dots(q, c, gneq, 2, c->U, c->U, c->u1, c->u1, nullptr, nullptr, nullptr, nullptr, c->dots + D_VV);
        dots(q, c, gnpno, 3, c->P, c->P, c->Ah, c->Ah, s2, s2, nullptr, nullptr, c->dots + D_PP);
        if (sync_dots(c)) return -1;
        const double *h = c->h_dots;
        const double alpha = valid ? h[D_R1Z1 + (count & 1)] / h[D_S2AH] : 0.0;
        double vdotv = h[D_VV], pdotp = h[D_PP];
        s.incompressibility = std::sqrt((gneq / gnpno) * (1.0e-32 + h[D_PP + 1] / (1.0e-32 + vdotv)));
        dpressure = (float)(alpha * std::sqrt(h[D_PP + 2] / (1.0e-32 + pdotp)));
        dvelocity = (float)(alpha * std::sqrt(h[D_VV + 1] / (1.0e-32 + vdotv)));
        s.vdotv = std::sqrt(vdotv / gneq);
        s.pdotp = std::sqrt(pdotp / gnpno);
        count++;
        if (c->print_convergence)   // generate_log_message, Stokes_flow_Incomp.c:501-526
            std::fprintf(stderr, "AhatP (%03d) with v %12.6e dv/v %9.3e div/v %9.3e  p %12.6e dp/p %9.3e\n", count,
                         s.vdotv, dvelocity, s.incompressibility, s.pdotp, dpressure);
        sh = s1; s1 = s2; s2 = sh;
        sh = r0; r0 = r1; r1 = r2; r2 = sh;
        sh = z0; z0 = z1; z1 = sh;
    }
    // keep the rotated buffers for the next call, as the reference's statics
    c->r0 = r0; c->r1 = r1; c->r2 = r2; c->z0 = z0; c->z1 = z1; c->s1 = s1; c->s2 = s2;
    v_from_vector(q, L, c->U, c->V);
    s.iterations = count;
    s.dpressure = dpressure;
    s.dvelocity = dvelocity;
    if (st) *st = s;
    return prof_flush(c);
}

// the FP32 cycles' exact coarsest-level solve (mixed forms): after every
// rebuild, the level's stiffness assembled dense, factored (potrf) and
// inverted (potri) in FP64, the inverse kept as FP16 rows scaled by their
// largest entry (to about 5e-4 of it).  The first
// factorization runs on the work stream and is used at once; later ones
// (assembly, factorization, inverse, conversion) run on a side stream
// while the step's solve goes on with the previous inverse (an exact solve
// of the previous step's stiffness, a slightly weaker coarse correction),
// and are taken up at the next rebuild by one device copy.  A factorization
// that fails (the matrix not positive definite in FP64) leaves the cycles
// on the full hierarchy until one succeeds; the captured multigrid graph
// follows the change.  The learned coarse-level correction, when on, keeps
// the full hierarchy (it works on its coarsest level).

// the pending side-stream factorization (already waited for by rebuild):
// when both its steps succeeded, its FP32 inverse replaces the one the
// cycles read.  *state: 0 none was pending, 1 taken, 2 it failed
int direct_take(chip_ctx *c, int *state) {
    *state = 0;
    if (!c->fact_pending) return 0;
    c->fact_pending = false;
    int info[2] = {1, 1};
    CHIP_CHECK(hipMemcpyAsync(info, c->dinfo, sizeof info, hipMemcpyDeviceToHost, c->side));
    CHIP_CHECK(hipStreamSynchronize(c->side));
    *state = 2;
    if (info[0] != 0 || info[1] != 0) return 0;
    const size_t n = 3 * (size_t)c->L[c->dlev].g.nno;
    CHIP_CHECK(hipMemcpyAsync(c->dAh, c->dAh_stage, n * n * sizeof(uint16_t), hipMemcpyDeviceToDevice, S(c)));
    CHIP_CHECK(hipMemcpyAsync(c->dAs, c->dAs_stage, n * sizeof(float), hipMemcpyDeviceToDevice, S(c)));
    *state = 1;
    return 0;
}

int direct_factor(chip_ctx *c) {
    bool was = c->direct_ok;
    int state = 0;
    if (direct_take(c, &state)) return -1;
    const bool taken = state == 1;
    if (state == 2) {   // the last one failed: factor this step's on the work stream
        c->direct_ok = false;
        if (drop_mg_graph(c)) return -1;
        was = false;
    }
    if (c->dlev < 0 || !c->direct_on) {
        c->direct_ok = false;
        return (c->direct_ok != was) ? drop_mg_graph(c) : 0;
    }
    const Level &L = c->L[c->dlev];
    const int n = 3 * L.g.nno;
    if (!c->rb) {
        rocblas_handle h;
        // no atomics in the library's algorithms: they would make the
        // factorization's last bits depend on timing
        if (rocblas_create_handle(&h) != rocblas_status_success ||
            rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed) != rocblas_status_success) {
            set_error("coarse direct solve: rocblas_create_handle failed");
            return -1;
        }
        c->rb = h;
        CHIP_CHECK(hipStreamCreate(&c->side));
        CHIP_CHECK(hipEventCreateWithFlags(&c->ev_k, hipEventDisableTiming));
        CHIP_CHECK(hipEventCreateWithFlags(&c->ev_fact, hipEventDisableTiming));
    }
    rocblas_handle h = (rocblas_handle)c->rb;
    const bool async = c->direct_async && (was || taken);   // an inverse to go on with
    if (async) {
        // the side stream reads this step's stiffness (rebuild waits for it
        // before the next one overwrites it)
        CHIP_CHECK(hipEventRecord(c->ev_k, S(c)));
        CHIP_CHECK(hipStreamWaitEvent(c->side, c->ev_k, 0));
        dense_assemble(c->side, L, c->dA);
        if (rocblas_set_stream(h, c->side) != rocblas_status_success ||
            rocsolver_dpotrf(h, rocblas_fill_lower, n, c->dA, n, c->dinfo) != rocblas_status_success ||
            rocsolver_dpotri(h, rocblas_fill_lower, n, c->dA, n, c->dinfo + 1) != rocblas_status_success) {
            set_error("coarse direct solve: rocSOLVER failed to launch");
            return -1;
        }
        sym_to_f32(c->side, c->dA, c->dAinv, n);
        rowscale_f16(c->side, c->dAinv, c->dAh_stage, c->dAs_stage, n);
        CHIP_CHECK(hipGetLastError());
        CHIP_CHECK(hipEventRecord(c->ev_fact, c->side));
        c->fact_pending = true;
        c->direct_ok = true;
        return 0;
    }
    dense_assemble(S(c), L, c->dA);
    int info = 0;
    if (rocblas_set_stream(h, S(c)) != rocblas_status_success ||
        rocsolver_dpotrf(h, rocblas_fill_lower, n, c->dA, n, c->dinfo) != rocblas_status_success) {
        set_error("coarse direct solve: rocsolver_dpotrf failed to launch");
        return -1;
    }
    CHIP_CHECK(hipMemcpyAsync(&info, c->dinfo, sizeof(int), hipMemcpyDeviceToHost, S(c)));
    CHIP_CHECK(hipStreamSynchronize(S(c)));
    if (info == 0) {
        if (rocsolver_dpotri(h, rocblas_fill_lower, n, c->dA, n, c->dinfo + 1) != rocblas_status_success) {
            set_error("coarse direct solve: rocsolver_dpotri failed to launch");
            return -1;
        }
        CHIP_CHECK(hipMemcpyAsync(&info, c->dinfo + 1, sizeof(int), hipMemcpyDeviceToHost, S(c)));
        CHIP_CHECK(hipStreamSynchronize(S(c)));
    }
    c->direct_ok = info == 0;
    if (c->direct_ok) {
        sym_to_f32(S(c), c->dA, c->dAinv, n);
        rowscale_f16(S(c), c->dAinv, c->dAh, c->dAs, n);
        CHIP_CHECK(hipGetLastError());
    }
    return (c->direct_ok != was) ? drop_mg_graph(c) : 0;
}

int rebuild(chip_ctx *c) {
    hipStream_t q = S(c);
    // a side-stream factorization still reading the stiffness finishes
    // before the stiffness is rebuilt
    if (c->fact_pending) CHIP_CHECK(hipStreamWaitEvent(q, c->ev_fact, 0));
    for (int lev = c->levmax; lev > c->levmin; lev--) project_viscosity(q, c->L[lev], c->L[lev - 1]);
    for (int lev = c->levmin; lev <= c->levmax; lev++) elt_g(q, c->L[lev]);
    for (int lev = c->levmax; lev >= c->levmin; lev--) node_ks(q, c->L[lev], c->p);
    for (int lev = c->levmax; lev >= c->levmin; lev--) build_bpi(q, c->L[lev], c->p.precond);
    if (c->cnet_w) cnet_logeta(q, c->L[c->levmin], c->cnet_logeta);
    if (c->L[c->levmin].Mc) coarse_build(q, c->L[c->levmin], c->p.v_steps_low);
    for (int lev = c->levmin; lev <= c->levmax; lev++)
        if (c->L[lev].Ma) affine_build(q, c->L[lev], c->L[lev].affine_sweeps);
    if (c->p.mixed)
        for (int lev = c->levmin; lev <= c->levmax; lev++) to_f32(q, c->L[lev].BIf, c->L[lev].BI, 3L * c->L[lev].g.nno);
    for (int lev = c->levmin; lev <= c->levmax; lev++)
        if (c->L[lev].Kh) scale_half(q, c->L[lev]);
    for (int lev = c->levmin; lev <= c->levmax; lev++) box_tiles(q, c->L[lev]);
    CHIP_CHECK(hipGetLastError());
    return direct_factor(c);
}

// PG_timestep (Advection_diffusion.c:251-345)
int pg_timestep(chip_ctx *c, chip_pg_stats *st) {
    hipStream_t q = S(c);
    const Level &L = c->L[c->levmax];
    const long nno = L.g.nno;
    float adv, Tint1, Tint = 0.0f;
    timestep_tmax(q, c, c->V, c->T);
    if (sync_minmax(c, &adv, &Tint1)) return -1;
    float timestep;
    if (c->p.fixed_timestep != 0.0f) timestep = c->p.fixed_timestep;
    else {
        const float a = c->p.finetunedt * adv;
        timestep = (float)(1.0e-32 + (a < c->diff_timestep ? a : c->diff_timestep));
    }
    fcopy(q, c->T1, c->T, nno);
    fcopy(q, c->Tdot1, c->Tdot, nno);
    float dt_reduced = 1.0f;
    int last_sub = 1, iredo;
    do {
        timestep *= dt_reduced;
        iredo = 0;
        predictor(q, L, c->T, c->Tdot, (float)((1.0 - c->p.gamma) * timestep));
        for (int pass = 0; pass < c->p.temp_iterations; pass++) {
            pg_residual(q, c, c->T, c->Tdot, c->V, c->Eres);
            pg_gather_correct(q, c, c->Eres, c->T, c->Tdot, c->p.gamma * timestep);
        }
        tmax(q, c, c->T);
        if (sync_minmax(c, nullptr, &Tint)) return -1;
        if (Tint / Tint1 > 1.01f) {
            fcopy(q, c->T, c->T1, nno);
            fcopy(q, c->Tdot, c->Tdot1, nno);
            iredo = 1;
            dt_reduced *= 0.5f;
            last_sub++;
        }
    } while (iredo == 1 && last_sub <= 5);
    c->elapsed_time += timestep;
    conform_temperature(q, c, c->T);
    thermal_buoyancy(q, c, c->T, c->buoy);
    CHIP_CHECK(hipGetLastError());
    if (st) {
        st->timestep = timestep;
        st->T_interior = Tint;
        st->sub_iterations = last_sub;
        st->elapsed_time = c->elapsed_time;
    }
    return 0;
}

int heat_flux_nu(chip_ctx *c, float *nu) {
    heat_flux(S(c), c, c->T, c->V);
    if (sync_dots(c)) return -1;
    const double *h = c->h_dots + D_HF;
    const float inp[4] = {(float)h[0], (float)h[1], (float)h[2], (float)h[3]};
    nu[1] = inp[0] / (inp[1] * 4);
    nu[0] = inp[2] / (inp[3] * 4);
    return 0;
}

// the initial guess of the Uzawa solve: the previous solution (the
// reference, Stokes_flow_Incomp.c:354-366 start from E->U, E->P as the last
// solve left them) or, with the warm start, that solution extrapolated
// along the sequence of solves (the timestep follows the flow speed, so the
// change per step is more regular than the change per unit time): the
// velocity linearly, 2 U_n - U_(n-1); the pressure with its increment
// averaged over the last two solves 3:1, P_n + 3/4 (P_n - P_(n-1)) +
// 1/4 (P_(n-1) - P_(n-2)) (linearly while only two are kept).  The
// solves stop on relative increments, so they leave remainders in the slow
// pressure modes that alternate from solve to solve; the average halves an
// alternating part of the increment and keeps a steady one (over 60 steps
// of the 64^3 and 32 of the 128x128x64 workload: 16 % and 13 % fewer Uzawa
// iterations than the linear rule).  With warm_start = 2 the U-Net's
// correction of that extrapolation is added, once three solves are kept.
// When training samples are recorded (or the network runs) the network's
// input is formed here and the target after the solve (warm_finish).
int warm_start_guess(chip_ctx *c) {
    c->feat_pending = false;
    if (!c->warm_start || c->hist < 1) return 0;
    hipStream_t q = S(c);
    const Level &L = c->L[c->levmax];
    const long nno = L.g.nno, neq = 3L * nno;
    const bool net = c->warm_start == 2 && c->wn_set && c->hist >= 3;
    const bool rec = c->rec_x && c->rec_y && c->hist >= 3;
    const double r = c->hist >= 2 ? 1.0 : 0.0;
    const double p1 = c->hist >= 3 ? 0.75 : r, p2 = c->hist >= 3 ? 0.25 : 0.0;
    extrapolate(q, c->U, c->Uprev, c->Uprev2, r, 0.0, (net || rec) ? c->D2u : nullptr, neq);
    extrapolate(q, c->P, c->Pprev, c->Pprev2, p1, p2, (net || rec) ? c->D2p : nullptr, L.g.nel);
    c->wscale[6] = c->hist;
    if (!net && !rec) return 0;
    // features: the temperature's departure from its own extrapolation, the
    // nodal pressures, log10 viscosity (EVI is this step's), and their scales
    warm_dT(q, c->T, c->Th0, c->Th1, nno, c->wn_dT);
    e2n(q, L, c->P, nullptr, c->wn_pn);
    e2n(q, L, c->P, c->Pprev, c->wn_dpn);
    e2n(q, L, c->D2p, nullptr, c->wn_d2pn);
    cnet_logeta(q, L, c->wn_logeta);
    dots(q, c, (int)neq, 2, (const double *)c->U, (const double *)c->U, (const double *)c->D2u,
         (const double *)c->D2u, nullptr, nullptr, nullptr, nullptr, c->dots + D_WS);
    dots(q, c, (int)nno, 3, (const float *)c->wn_dT, (const float *)c->wn_dT, (const float *)c->wn_pn,
         (const float *)c->wn_pn, (const float *)c->wn_d2pn, (const float *)c->wn_d2pn, nullptr, nullptr,
         c->dots + D_WS + 2);
    if (sync_dots(c)) return -1;
    const double *h = c->h_dots + D_WS;
    const double s_u = std::sqrt(h[0] / neq), s_2u = std::sqrt(h[1] / neq), s_dT = std::sqrt(h[2] / nno),
                 s_p = std::sqrt(h[3] / nno), s_2p = std::sqrt(h[4] / nno);
    double *w = c->wscale;
    w[0] = s_u, w[1] = s_p, w[2] = s_dT, w[3] = s_2u, w[4] = s_2p, w[5] = 0.0, w[7] = 1.0;
    auto inv = [](double v) { return v > 0.0 ? (float)(1.0 / v) : 0.0f; };
    warm_features(q, L, c->T, c->wn_dT, c->wn_logeta, c->U, c->Uprev, c->D2u, c->wn_pn, c->wn_dpn, c->wn_d2pn,
                  inv(s_dT), inv(s_u), inv(s_p), inv(s_2u), inv(s_2p), c->wn_x);
    copy(q, c->Ue, c->U, neq);
    copy(q, c->Pe, c->P, L.g.nel);
    // the extrapolation's error is of the order of the second difference:
    // the network's outputs are in those units
    const bool scaled = s_2u > 0.0 && s_2p > 0.0;
    if (rec) {
        CHIP_CHECK(hipMemcpyAsync(c->rec_x, c->wn_x, 16 * (size_t)nno * sizeof(uint16_t), hipMemcpyDeviceToDevice, q));
        c->feat_pending = scaled;
    }
    if (net && scaled) {
        if (cl_unet_forward(c->wn_x, 1, L.g.nox, L.g.noy, L.g.noz, c->wn_w, c->wn_b, c->wn_out, c->wn_ws, q)) {
            set_error("warm start: the U-Net's forward pass failed to launch");
            return -1;
        }
        warm_apply(q, L, c->wn_out, s_2u, s_2p, c->U, c->P);
    }
    CHIP_CHECK(hipGetLastError());
    return 0;
}

// after the solve: the training target of a recorded sample, then the
// solve joins the history
int warm_finish(chip_ctx *c) {
    if (!c->warm_start) return 0;
    hipStream_t q = S(c);
    const Level &L = c->L[c->levmax];
    if (c->feat_pending) {
        e2n(q, L, c->P, c->Pe, c->wn_dpn);
        warm_target(q, L, c->U, c->Ue, c->wn_dpn, 1.0 / c->wscale[3], 1.0 / c->wscale[4], c->rec_y);
        c->feat_pending = false;
    }
    fcopy(q, c->Th1, c->Th0, L.g.nno);
    fcopy(q, c->Th0, c->T, L.g.nno);
    c->hist++;
    CHIP_CHECK(hipGetLastError());
    return 0;
}

int stokes_step(chip_ctx *c, chip_stokes_stats *st) {
    hipStream_t q = S(c);
    const Level &L = c->L[c->levmax];
    // get_system_viscosity + construct_stiffness_B_matrix (Drive_solvers.c:123-126);
    // the viscosity first, so that the warm start reads this step's
    const bool update = !c->built || c->p.visc_update;
    if (update) visc_from_T(q, c, c->T, c->L[c->levmax].EVI);
    if (warm_start_guess(c)) return -1;
    conform_velocity(q, L, c->U);
    assemble_forces(q, c, c->buoy, c->F);
    if (update) {
        c->built = true;
        if (rebuild(c)) return -1;
    }
    if (uzawa(c, st)) return -1;
    return warm_finish(c);
}

}  // namespace
}  // namespace chip

using namespace chip;

extern "C" {

const char *chip_last_error(void) { return chip::g_err.c_str(); }

// every entry point taking a context refuses a null one
#define CHIP_NEED_CTX(c)                                 \
    do {                                                 \
        if (!(c)) {                                      \
            chip::set_error(std::string(__func__) + ": null context"); \
            return -1;                                   \
        }                                                \
    } while (0)

int chip_create(const chip_problem *p, chip_ctx **out) {
    *out = nullptr;
    if (!p || !p->xc || !p->yc || !p->zc) {
        set_error("chip_create: missing problem or coordinates");
        return -1;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error("chip_create: no HIP device");
        return -2;
    }
    chip_ctx *c = new chip_ctx();
    c->p = *p;
    const int f = 1 << (p->levels - 1);
    c->xc.assign(p->xc, p->xc + p->mgunitx * f + 1);
    c->yc.assign(p->yc, p->yc + p->mgunity * f + 1);
    c->zc.assign(p->zc, p->zc + p->mgunitz * f + 1);
    c->p.xc = c->xc.data();
    c->p.yc = c->yc.data();
    c->p.zc = c->zc.data();
    if (upload_constants()) {
        set_error("chip_create: constant upload failed");
        delete c;
        return -1;
    }
    (void)lds_limit();
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        set_error("chip_create: stream");
        delete c;
        return -1;
    }
    c->own_stream = true;
    if (build_mesh(c)) {
        chip_destroy(c);
        return -1;
    }
    *out = c;
    return 0;
}

void chip_destroy(chip_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    if (c->rb) (void)rocblas_destroy_handle((rocblas_handle)c->rb);
    if (c->ev_k) (void)hipEventDestroy(c->ev_k);
    if (c->ev_fact) (void)hipEventDestroy(c->ev_fact);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->mg_graph) (void)hipGraphExecDestroy(c->mg_graph);
    for (void *ptr : c->allocs) (void)hipFree(ptr);
    if (c->h_dots) (void)hipHostFree(c->h_dots);
    if (c->h_minmax) (void)hipHostFree(c->h_minmax);
    for (auto &e : c->prof.ev) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int chip_set_stream(chip_ctx *c, void *hip_stream) {
    CHIP_NEED_CTX(c);
    if (!hip_stream) return 0;
    if (c->own_stream) {
        CHIP_CHECK(hipStreamSynchronize(c->stream));
        CHIP_CHECK(hipStreamDestroy(c->stream));
        c->own_stream = false;
    }
    c->stream = (hipStream_t)hip_stream;
    return 0;
}

void *chip_get_stream(chip_ctx *c) { return (void *)c->stream; }

static int resolve(chip_ctx *c, const char *which, int lev, void **ptr, size_t *elem, long *count) {
    if (!c || !which) {
        set_error("null context or field name");
        return -1;
    }
    if (lev < 0) lev = c->levmax;
    if (lev < c->levmin || lev > c->levmax) {
        set_error("level out of range");
        return -1;
    }
    const Level &L = c->L[lev];
    const Geo &gf = c->L[c->levmax].g;
    const std::string w(which);
    if (w == "T") { *ptr = c->T; *elem = 4; *count = gf.nno; }
    else if (w == "Tdot") { *ptr = c->Tdot; *elem = 4; *count = gf.nno; }
    else if (w == "U") { *ptr = c->U; *elem = 8; *count = 3L * gf.nno; }
    else if (w == "P") { *ptr = c->P; *elem = 8; *count = gf.nel; }
    else if (w == "F") { *ptr = c->F; *elem = 8; *count = 3L * gf.nno; }
    else if (w == "V") { *ptr = c->V; *elem = 4; *count = 3L * gf.nno; }
    else if (w == "buoyancy") { *ptr = c->buoy; *elem = 4; *count = gf.nno; }
    else if (w == "EVI") { *ptr = L.EVI; *elem = 4; *count = 8L * L.g.nel; }
    else if (w == "BI") { *ptr = L.BI; *elem = 8; *count = 3L * L.g.nno; }
    else if (w == "BPI") { *ptr = L.BPI; *elem = 8; *count = L.g.nel; }
    else if (w == "EG") { *ptr = L.EG; *elem = 4; *count = 24L * L.g.nel; }
    else if (w == "MASS") { *ptr = L.MASS; *elem = 4; *count = L.g.nno; }
    else if (w == "NODE") { *ptr = L.NODE; *elem = 4; *count = L.g.nno; }
    else if (w == "GNX") { *ptr = L.GNX; *elem = 4; *count = (long)GNX_SZ * L.g.nel; }
    else if (w == "GDA") { *ptr = L.GDA; *elem = 4; *count = (long)GDA_SZ * L.g.nel; }
    else if (w == "ECO") { *ptr = L.ECO; *elem = 4; *count = (long)ECO_SZ * L.g.nel; }
    else if (w == "TWW") { *ptr = L.TWW; *elem = 4; *count = (long)TWW_SZ * L.g.nel; }
    else {
        set_error("unknown array " + w);
        return -1;
    }
    return 0;
}

int chip_upload(chip_ctx *c, const char *which, int lev, const void *host, long count) {
    if (!host) {
        set_error("chip_upload: null host pointer");
        return -1;
    }
    void *ptr;
    size_t elem;
    long n;
    if (resolve(c, which, lev, &ptr, &elem, &n)) return -1;
    if (count != n) {
        set_error(std::string("chip_upload ") + which + ": size mismatch");
        return -1;
    }
    CHIP_CHECK(hipMemcpyAsync(ptr, host, n * elem, hipMemcpyHostToDevice, c->stream));
    CHIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int chip_download(chip_ctx *c, const char *which, int lev, void *host, long count) {
    if (!c || !which || !host) {
        set_error("chip_download: null argument");
        return -1;
    }
    const std::string w(which);
    if (w == "K") {
        // device tiles [q/64][slot*9 + r*3 + c][q%64], colour order -> host [n][r*81 + slot*3 + c]
        if (lev < 0) lev = c->levmax;
        if (lev < c->levmin || lev > c->levmax) {
            set_error("level out of range");
            return -1;
        }
        const Level &L = c->L[lev];
        const Geo &g = L.g;
        if (count != (long)g.nno * KROW) {
            set_error("chip_download K: size mismatch");
            return -1;
        }
        std::vector<float> dev((size_t)KROW * g.pitch);
        CHIP_CHECK(hipMemcpyAsync(dev.data(), L.K, dev.size() * 4, hipMemcpyDeviceToHost, c->stream));
        CHIP_CHECK(hipStreamSynchronize(c->stream));
        float *out = (float *)host;
        for (int col = 0; col < 8; col++)
            for (int q = g.cs[col]; q < g.cs[col] + g.cn[col]; q++) {
                const int r0 = q - g.cs[col], kk = r0 % g.cz[col], t = r0 / g.cz[col];
                const int k = 2 * kk + (col & 1), i = 2 * (t % g.cx[col]) + ((col >> 1) & 1),
                          j = 2 * (t / g.cx[col]) + ((col >> 2) & 1);
                const int n = k + g.noz * (i + g.nox * j);
                for (int s = 0; s < 27; s++)
                    for (int r = 0; r < 3; r++)
                        for (int cc = 0; cc < 3; cc++)
                            out[(size_t)n * KROW + r * 81 + s * 3 + cc] =
                                dev[(size_t)(q >> 6) * KROW * 64 + (size_t)(s * 9 + r * 3 + cc) * 64 + (q & 63)];
            }
        return 0;
    }
    void *ptr;
    size_t elem;
    long n;
    if (resolve(c, which, lev, &ptr, &elem, &n)) return -1;
    if (count != n) {
        set_error(std::string("chip_download ") + which + ": size mismatch");
        return -1;
    }
    CHIP_CHECK(hipMemcpyAsync(host, ptr, n * elem, hipMemcpyDeviceToHost, c->stream));
    CHIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int chip_get_system_viscosity(chip_ctx *c) {
    CHIP_NEED_CTX(c);
    visc_from_T(c->stream, c, c->T, c->L[c->levmax].EVI);
    CHIP_CHECK(hipGetLastError());
    return 0;
}

int chip_construct_stiffness_B_matrix(chip_ctx *c) {
    CHIP_NEED_CTX(c);
    c->built = true;
    if (rebuild(c)) return -1;
    CHIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int chip_apply(chip_ctx *c, const char *op, int lev, const double *in, const double *in2, double *out, double *out2,
               int arg) {
    CHIP_NEED_CTX(c);
    if (!op) {
        set_error("chip_apply: null op");
        return -1;
    }
    if (lev < c->levmin || lev > c->levmax) {
        set_error("chip_apply: level out of range");
        return -1;
    }
    const std::string o(op);
    const Level &L = c->L[lev];
    const long nv = 3L * L.g.nno, np = L.g.nel;
    hipStream_t q = c->stream;
    // scratch on the level's MG vectors (and the coarser level's for transfers)
    double *a = L.del, *b = L.res, *d = L.AU;
    auto up = [&](double *dst, const double *src, long n) {
        return hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, q);
    };
    auto down = [&](double *dst, const double *src, long n) {
        return hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, q);
    };
    if (o == "matvec") {
        CHIP_CHECK(up(a, in, nv));
        do_matvec(c, lev, a, b);
        CHIP_CHECK(down(out, b, nv));
    } else if (o == "gs0" || o == "gs1") {
        const int guess = (o == "gs1");
        if (guess) CHIP_CHECK(up(a, in, nv));
        CHIP_CHECK(up(b, in2, nv));
        do_gs(c, lev, a, b, d, arg, guess);
        CHIP_CHECK(down(out, a, nv));
        CHIP_CHECK(down(out2, d, nv));
        if (sync_dots(c) || xcd_status(c)) return -1;
    } else if (o == "box0" || o == "box1") {   // FP32 box smoother call: arg sweeps, then Ad = K d
        if (!L.Kbox || !L.velf) {
            set_error("box: the level has no box smoother (FP32 cycles of a mixed form, levels between the "
                      "exact or affine level and the finest)");
            return -1;
        }
        const int guess = (o == "box1");
        float *af = L.delf, *bf = L.resf, *df = L.AUf;
        if (guess) {
            CHIP_CHECK(up(a, in, nv));
            to_f32(q, af, a, nv);
        }
        CHIP_CHECK(up(b, in2, nv));
        to_f32(q, bf, b, nv);
        const bool keep = c->box_gs;
        c->box_gs = true;
        do_gs(c, lev, af, (const float *)bf, df, arg, guess);
        c->box_gs = keep;
        fill(q, a, 0.0, nv);
        add_f32(q, a, af, nv);
        fill(q, d, 0.0, nv);
        add_f32(q, d, df, nv);
        CHIP_CHECK(down(out, a, nv));
        CHIP_CHECK(down(out2, d, nv));
        CHIP_CHECK(hipStreamSynchronize(q));
    } else if (o == "affine0" || o == "affine1") {
        if (!L.Ma) {
            set_error("affine: the level has no affine smoother form");
            return -1;
        }
        const int guess = (o == "affine1");
        if (guess) CHIP_CHECK(up(a, in, nv));
        CHIP_CHECK(up(b, in2, nv));
        affine_apply(q, L, guess ? (const double *)a : nullptr, (const double *)b, L.tmp);
        CHIP_CHECK(down(out, L.tmp, nv));
    } else if (o == "direct") {
        if (lev != c->dlev || !c->direct_ok) {
            set_error("direct: the level solved exactly by the FP32 cycles, after a rebuild");
            return -1;
        }
        CHIP_CHECK(up(b, in2, nv));
        dense_apply(q, c->dAh, c->dAs, (const double *)b, a, (int)nv);
        CHIP_CHECK(down(out, a, nv));
    } else if (o == "coarse") {
        if (lev != c->levmin || !L.Mc) {
            set_error("coarse: the coarsest level has no coarse matrix");
            return -1;
        }
        CHIP_CHECK(up(b, in2, nv));
        coarse_apply(q, L, b, a);
        CHIP_CHECK(down(out, a, nv));
    } else if (o == "cnet") {
        if (lev != c->levmin || !c->cnet_layers) {
            set_error("cnet: the coarsest level, after chip_set_coarse_net");
            return -1;
        }
        CHIP_CHECK(up(b, in2, nv));
        if (cnet_apply(q, c, L, (const double *)b, a)) {
            set_error("cnet: launch failed");
            return -1;
        }
        CHIP_CHECK(down(out, a, nv));
    } else if (o == "project") {
        if (lev == c->levmin) {
            set_error("project: no coarser level");
            return -1;
        }
        const Level &C = c->L[lev - 1];
        CHIP_CHECK(up(a, in, nv));
        project(q, L, C, a, C.AU, false);
        CHIP_CHECK(down(out, C.AU, 3L * C.g.nno));
    } else if (o == "interp") {
        if (lev == c->levmin) {
            set_error("interp: no coarser level");
            return -1;
        }
        const Level &C = c->L[lev - 1];
        CHIP_CHECK(up(C.AU, in, 3L * C.g.nno));
        interp(q, C, L, C.AU, a, false);
        CHIP_CHECK(down(out, a, nv));
    } else if (o == "div") {
        CHIP_CHECK(up(a, in, nv));
        div(q, L, a, b);
        CHIP_CHECK(down(out, b, np));
    } else if (o == "grad") {
        CHIP_CHECK(up(a, in, np));
        grad(q, L, a, b);
        CHIP_CHECK(down(out, b, nv));
    } else if (o == "strip") {
        CHIP_CHECK(up(a, in, nv));
        strip(q, L, a);
        CHIP_CHECK(down(out, a, nv));
    } else {
        set_error("chip_apply: unknown op " + o);
        return -1;
    }
    CHIP_CHECK(hipGetLastError());
    CHIP_CHECK(hipStreamSynchronize(q));
    return prof_flush(c);
}

int chip_solve_del2_u(chip_ctx *c, const double *F, double *d0, double acc, int *valid, int *cycles) {
    CHIP_NEED_CTX(c);
    const long neq = 3L * c->L[c->levmax].g.nno;
    CHIP_CHECK(hipMemcpyAsync(c->Ah, F, neq * 8, hipMemcpyHostToDevice, c->stream));
    int v = 0, n = 0;
    if (solve_del2_u(c, c->u1, c->Ah, acc, &v, &n)) return -1;
    CHIP_CHECK(hipMemcpyAsync(d0, c->u1, neq * 8, hipMemcpyDeviceToHost, c->stream));
    CHIP_CHECK(hipStreamSynchronize(c->stream));
    if (valid) *valid = v;
    if (cycles) *cycles = n;
    return prof_flush(c);
}

int chip_solve_constrained_flow_iterative(chip_ctx *c, chip_stokes_stats *st) {
    CHIP_NEED_CTX(c);
    if (c->warm_start) {   // the warm start's guess replaces the uploaded U, P
        if (warm_start_guess(c)) return -1;
        conform_velocity(S(c), c->L[c->levmax], c->U);
    }
    if (uzawa(c, st)) return -1;
    return warm_finish(c);
}

int chip_PG_timestep(chip_ctx *c, chip_pg_stats *st) {
    CHIP_NEED_CTX(c);
    return pg_timestep(c, st);
}

int chip_thermal_buoyancy(chip_ctx *c) {
    CHIP_NEED_CTX(c);
    thermal_buoyancy(c->stream, c, c->T, c->buoy);
    CHIP_CHECK(hipGetLastError());
    return 0;
}

int chip_assemble_forces(chip_ctx *c) {
    CHIP_NEED_CTX(c);
    assemble_forces(c->stream, c, c->buoy, c->F);
    CHIP_CHECK(hipGetLastError());
    return 0;
}

int chip_heat_flux(chip_ctx *c, float *nu) {
    CHIP_NEED_CTX(c);
    return heat_flux_nu(c, nu);
}

int chip_vrms(chip_ctx *c, double *vrms_out) {
    CHIP_NEED_CTX(c);
    vrms(c->stream, c, c->V, c->dots + D_TMP);
    if (sync_dots(c)) return -1;
    *vrms_out = std::sqrt(c->h_dots[D_TMP] / c->h_dots[D_TMP + 1]);
    return 0;
}

int chip_averages(chip_ctx *c, float *T_ave, float *vrms_ave, float *visc_ave) {
    CHIP_NEED_CTX(c);
    const Geo &g = c->L[c->levmax].g;
    if (!c->Have) {
        CHIP_CHECK(hipMalloc(&c->Have, 3 * (size_t)g.noz * sizeof(float)));
        c->allocs.push_back(c->Have);
        CHIP_CHECK(hipMalloc(&c->ave_scratch, (size_t)g.nno * sizeof(float)));
        c->allocs.push_back(c->ave_scratch);
        c->dev_bytes += (3L * g.noz + g.nno) * (long)sizeof(float);
    }
    averages(c->stream, c, c->Have, c->ave_scratch);
    CHIP_CHECK(hipGetLastError());
    std::vector<float> h(3 * (size_t)g.noz);
    CHIP_CHECK(hipMemcpyAsync(h.data(), c->Have, h.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    CHIP_CHECK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < g.noz; k++) {
        if (T_ave) T_ave[k] = h[k];
        // Have.vrms[i] = sqrt(Have.vrms[i]) on a float (Process_velocity.c:219-220)
        if (vrms_ave) vrms_ave[k] = (float)std::sqrt((double)h[g.noz + k]);
        if (visc_ave) visc_ave[k] = h[2 * g.noz + k];
    }
    return 0;
}

int chip_general_stokes_solver(chip_ctx *c, chip_stokes_stats *st) {
    CHIP_NEED_CTX(c);
    const int r = stokes_step(c, st);
    return r;
}

int chip_timestep(chip_ctx *c, chip_pg_stats *pg, chip_stokes_stats *st, float *nu) {
    CHIP_NEED_CTX(c);
    c->solution_cycles++;
    if (pg_timestep(c, pg)) return -1;
    if (nu && heat_flux_nu(c, nu)) return -1;
    return stokes_step(c, st);
}

int chip_profile_enable(chip_ctx *c, const char *kernel, int on) {
    CHIP_NEED_CTX(c);
    const std::string k(kernel ? kernel : "");
    if (prof_flush(c)) return -1;
    if (!on) {
        c->prof.kernel = 0;
        return 0;
    }
    c->prof.kernel = (k == "gs_colour") ? 1 : (k == "matvec") ? 2 : (k == "gs_wg") ? 3 : (k == "gs_xcd") ? 4 : (k == "gs_box") ? 5 : 0;
    c->prof.launches = 0;
    c->prof.total_ms = c->prof.total_bytes = 0.0;
    if (!c->prof.kernel) {
        set_error("chip_profile_enable: unknown kernel " + k);
        return -1;
    }
    return 0;
}

int chip_profile_read(chip_ctx *c, long *launches, double *total_ms, double *bytes) {
    CHIP_NEED_CTX(c);
    if (prof_flush(c)) return -1;
    if (launches) *launches = c->prof.launches;
    if (total_ms) *total_ms = c->prof.total_ms;
    if (bytes) *bytes = c->prof.total_bytes;
    return 0;
}

long chip_device_bytes(chip_ctx *c) { return c ? c->dev_bytes : -1; }

int chip_set_coarse_net(chip_ctx *c, int layers, const uint16_t *weights, const float *bias, const float *out_scale) {
    CHIP_NEED_CTX(c);
    if (layers < 1 || layers > CNET_MAX_LAYERS || !weights || !bias || !out_scale) {
        set_error("chip_set_coarse_net: 1.." + std::to_string(CNET_MAX_LAYERS) +
                  " layers and non-null weights, biases and output scales");
        return -1;
    }
    if (!cnet_fits(c->L[c->levmin].g)) {
        set_error("chip_set_coarse_net: the coarsest level has more than 1024 nodes");
        return -1;
    }
    if (alloc_coarse_net(c)) return -1;
    hipStream_t q = c->stream;
    CHIP_CHECK(hipMemcpyAsync(c->cnet_w, weights, (size_t)layers * 16 * 27 * 16 * sizeof(uint16_t),
                              hipMemcpyHostToDevice, q));
    CHIP_CHECK(hipMemcpyAsync(c->cnet_b, bias, (size_t)layers * 16 * sizeof(float), hipMemcpyHostToDevice, q));
    CHIP_CHECK(hipMemcpyAsync(c->cnet_ys, out_scale, 3 * sizeof(float), hipMemcpyHostToDevice, q));
    // EVI of the coarsest level holds the last rebuild's viscosity (or the
    // isoviscous start); every later rebuild forms the feature again
    cnet_logeta(q, c->L[c->levmin], c->cnet_logeta);
    CHIP_CHECK(hipGetLastError());
    CHIP_CHECK(hipStreamSynchronize(q));
    c->cnet_layers = layers;
    if (c->mg_graph) {   // the layer count is part of the captured sequence
        CHIP_CHECK(hipGraphExecDestroy(c->mg_graph));
        c->mg_graph = nullptr;
    }
    return 0;
}

int chip_coarse_net_state(chip_ctx *c, int *on, int *fallbacks) {
    CHIP_NEED_CTX(c);
    if (on) *on = c->cnet_on ? 1 : 0;
    if (fallbacks) *fallbacks = c->cnet_fallbacks;
    return 0;
}

int chip_set_warm_net(chip_ctx *c, const uint16_t *weights, long n_weights, const float *bias, long n_bias) {
    CHIP_NEED_CTX(c);
    long nw = 0, nb = 0;
    cl_unet_params(&nw, &nb);
    if (!weights || !bias || n_weights != nw || n_bias != nb) {
        set_error("chip_set_warm_net: the U-Net takes " + std::to_string(nw) + " weights and " + std::to_string(nb) +
                  " biases");
        return -1;
    }
    const Geo &g = c->L[c->levmax].g;
    for (int n : {g.nox, g.noy, g.noz})
        if (n < 17 || (n - 1) % 8) {
            set_error("chip_set_warm_net: the finest grid needs n - 1 divisible by 8 and n >= 17 nodes per axis");
            return -1;
        }
    if (alloc_warm_start(c) || alloc_warm_net(c)) return -1;
    CHIP_CHECK(hipMemcpyAsync(c->wn_w, weights, (size_t)nw * sizeof(uint16_t), hipMemcpyHostToDevice, c->stream));
    CHIP_CHECK(hipMemcpyAsync(c->wn_b, bias, (size_t)nb * sizeof(float), hipMemcpyHostToDevice, c->stream));
    CHIP_CHECK(hipStreamSynchronize(c->stream));
    c->wn_set = true;
    return 0;
}

int chip_warm_record(chip_ctx *c, void *features, void *targets) {
    CHIP_NEED_CTX(c);
    if ((features == nullptr) != (targets == nullptr)) {
        set_error("chip_warm_record: give both buffers or neither");
        return -1;
    }
    if (features && alloc_warm_start(c)) return -1;
    c->rec_x = static_cast<uint16_t *>(features);
    c->rec_y = static_cast<float *>(targets);
    return 0;
}

int chip_warm_scales(chip_ctx *c, double *out) {
    CHIP_NEED_CTX(c);
    for (int i = 0; i < 8; i++) out[i] = c->wscale[i];
    return 0;
}

int chip_set_option(chip_ctx *c, const char *name, int value) {
    CHIP_NEED_CTX(c);
    const std::string n(name ? name : "");
    if (n == "graphs") {
        c->use_graphs = value != 0;
        return 0;
    }
    if (n == "profile_level") {  // restrict chip_profile_enable to one level (-1: all)
        c->prof.level = value;
        return 0;
    }
    if (n == "print_convergence") {
        c->print_convergence = value != 0;
        return 0;
    }
    if (n == "warm_start") {   // a new history starts with the next solve
        if (value < 0 || value > 2) {
            set_error("chip_set_option: warm_start is 0 (previous solution), 1 (extrapolated) or 2 (extrapolated "
                      "and corrected by the learned network)");
            return -1;
        }
        if (value == 2 && !c->wn_set) {
            set_error("chip_set_option: warm_start 2 needs chip_set_warm_net first");
            return -1;
        }
        if (value && alloc_warm_start(c)) return -1;
        c->warm_start = value;
        c->hist = 0;
        return 0;
    }
    if (n == "coarse_net" && value && !c->cnet_layers) {
        set_error("chip_set_option: coarse_net needs chip_set_coarse_net first");
        return -1;
    }
    if (n == "coarse_net" && value) c->cnet_calibrated = false;   // the next solve calibrates it
    if (n == "direct_async") {
        c->direct_async = value != 0;
        return 0;
    }
    if (n == "coarse_direct") {   // the FP32 cycles' exact coarsest-level solve
        c->direct_on = value != 0;
        if (c->built && direct_factor(c)) return -1;
        if (!c->built && !c->direct_on && drop_mg_graph(c)) return -1;
        return 0;
    }
    if (n == "xcd_gs" && value) {   // relies on round-robin dispatch to the XCDs: checked here
        for (int lev = c->levmin; lev <= c->levmax; lev++) {
            if (!gs_xcd_ok(c->L[lev])) continue;
            bool same = false;
            if (xcd_colocated(c->stream, xcd_groups(c->L[lev]), c->xcd_ids, &same)) {
                set_error("chip_set_option: XCD dispatch check failed");
                return -1;
            }
            if (!same) {
                set_error("chip_set_option: xcd_gs needs the workgroups of a launch dispatched round-robin to the "
                          "XCDs, which this device does not do");
                return -1;
            }
        }
    }
    if (n == "small_levels" || n == "gs_wg" || n == "coarse_net" || n == "xcd_gs" || n == "box_gs") {
        (n == "small_levels" ? c->small_levels
         : n == "gs_wg"      ? c->gs_wg
         : n == "xcd_gs"     ? c->xcd_gs
         : n == "box_gs"     ? c->box_gs
                             : c->cnet_on) = value != 0;
        if (c->mg_graph) {   // the captured kernel sequence changes
            CHIP_CHECK(hipStreamSynchronize(c->stream));
            CHIP_CHECK(hipGraphExecDestroy(c->mg_graph));
            c->mg_graph = nullptr;
        }
        return 0;
    }
    set_error("chip_set_option: unknown option " + n);
    return -1;
}

}  // extern "C"
