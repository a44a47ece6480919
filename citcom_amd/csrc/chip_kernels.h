// Launch wrappers of the gfx950 kernels (chip_kernels.hip).  All take the
// work stream and device pointers; none allocates.
#pragma once

#include "chip_common.h"

namespace chip {

// master-element constants into __constant__ memory
int upload_constants();
// largest LDS allocation per workgroup; also raises the kernels' dynamic LDS
// limit once, outside any stream capture
size_t lds_limit();

// ---- stiffness rebuild --------------------------------------------------
void visc_from_T(hipStream_t s, const chip_ctx *c, const float *T, float *evi);
void project_viscosity(hipStream_t s, const Level &fine, const Level &coarse);
void elt_g(hipStream_t s, const Level &L);
void node_ks(hipStream_t s, const Level &L, const chip_problem &p);
void build_bpi(hipStream_t s, const Level &L, int precond);

// ---- vector operators ---------------------------------------------------
// The multigrid operators take FP64 (plain form) or FP32 (mixed form)
// vectors; the stiffness is FP32 in both.
// Au = K u (strip BC rows); optionally res = rhs - Au as well.  With events
// e0/e1 the launch records the kernel's own start and end time in them.
template <typename V>
void matvec(hipStream_t s, const Level &L, const V *u, V *Au, const V *rhs, V *res, hipEvent_t e0 = nullptr,
            hipEvent_t e1 = nullptr, V *ucopy = nullptr);
// one colour pass of the node-block Gauss-Seidel (BI in the vector type)
template <typename V>
void gs_colour(hipStream_t s, const Level &L, const V *BI, int colour, V *d, const V *F, hipEvent_t e0 = nullptr,
               hipEvent_t e1 = nullptr);
// every sweep of a small level in one workgroup, then Ad = K d
void gs_small(hipStream_t s, const Level &L, double *d, const double *F, double *Ad, int sweeps, int guess);
// 1: the level's stiffness and vectors fit in LDS (gs_small), 0: one launch
// per colour pass
int gs_small_kind(const Level &L);
// a whole gauss_seidel call in one workgroup with d in LDS and the FP32
// stiffness read through L2, then Ad = K d (and res = rhs - Ad if res):
// for levels gs_wg_ok accepts (small enough, no FP16 stiffness)
template <typename V>
bool gs_wg_ok(const Level &L);
template <typename V>
void gs_wg(hipStream_t s, const Level &L, const V *BI, V *d, const V *F, V *Ad, const V *rhs, V *res, int sweeps,
           int guess, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// a whole gauss_seidel call as one launch of a few workgroups on one XCD,
// colour passes separated by barriers through that XCD's L2 (k_gs_xcd), then
// Ad = K d (and res = rhs - Ad); a nonzero *status after the launch means it
// stopped early (2: the workgroups were not on one XCD, 1: a barrier's poll
// limit) and its results are not valid
bool gs_xcd_ok(const Level &L);
int xcd_groups(const Level &L);
template <typename V>
void gs_xcd(hipStream_t s, const Level &L, const V *BI, V *d, const V *F, V *Ad, const V *rhs, V *res, int sweeps,
            int guess, double *status, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// box smoother of the FP32 cycles: one sweep as one launch, a workgroup per
// box of the level's partition; each box runs the 8 colours in order on its
// own nodes (in LDS) and reads the other boxes' nodes from din, as they were
// before the sweep, so the result does not depend on the order the boxes
// run in.  din == nullptr: the sweep starts from zero.  Writes every node of
// dout (dout != din).
bool box_part(int nno_z, int nno_x, int nno_y, BoxPart &bp);
size_t box_tile_elems(const Level &L);           // size of L.Kbox (L.nbox, L.Kbox_half set)
void box_tiles(hipStream_t s, const Level &L);   // L.Kbox from L.Kh or L.K
void gs_box(hipStream_t s, const Level &L, const float *BI, const float *din, float *dout, const float *F,
            hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// set-up check: do the active blocks of an 8 nwg-block launch share one XCD
// (four launches)?  ids_dev holds 16 unsigned
int xcd_colocated(hipStream_t s, int nwg, unsigned *ids_dev, bool *same);
// level above the coarsest: its smoother call as d = Ma din + Mb F
bool affine_ok(const Level &L);
void affine_build(hipStream_t s, const Level &L, int sweeps);
template <typename V>
void affine_apply(hipStream_t s, const Level &L, const V *din, const V *F, V *out);
// exact coarsest-level solve of the FP32 cycles: A (column-major, n = 3 nno)
// <- the level's stiffness with identity rows and columns at fixed
// components (zeroed first); M (row-major FP32) <- the symmetric matrix
// whose lower triangle A holds; H, sc <- M's rows scaled by their largest
// entry, in FP16 (bits), and the scales; out = diag(sc) H F
void dense_assemble(hipStream_t s, const Level &L, double *A);
void sym_to_f32(hipStream_t s, const double *A, float *M, int n);
void rowscale_f16(hipStream_t s, const float *M, uint16_t *H, float *sc, int n);
template <typename V>
void dense_apply(hipStream_t s, const uint16_t *H, const float *sc, const V *F, V *out, int n);
// coarsest level: Mc = the v_steps_low-sweep smoother as a matrix, d = Mc F
void coarse_build(hipStream_t s, const Level &L, int sweeps);
template <typename V>
void coarse_apply(hipStream_t s, const Level &L, const V *F, V *d);
// (zero, if given, is a coarse vector set to zero in the same launch)
template <typename V>
void project(hipStream_t s, const Level &fine, const Level &coarse, const V *AU, V *AD, bool do_strip = true,
             V *zero = nullptr);
template <typename V>
void interp(hipStream_t s, const Level &coarse, const Level &fine, const V *AD, V *AU, bool do_strip = true);
void div(hipStream_t s, const Level &L, const double *U, double *divU);
void grad(hipStream_t s, const Level &L, const double *P, double *gradP);
void strip(hipStream_t s, const Level &L, double *v);

// ---- learned coarse-level correction (no reference counterpart) ----------
// per node of level L: log10 viscosity, the mean over each adjacent
// element's Gauss points averaged over those elements
void cnet_logeta(hipStream_t s, const Level &L, float *out);
// d = the network's coarse correction for right-hand side F on the coarsest
// level: inputs (F / rms F, log10 eta), output scaled back by rms F, BC rows
// 0; one workgroup, for levels cnet_fits accepts (at most 1024 nodes)
bool cnet_fits(const Geo &g);
template <typename V>
int cnet_apply(hipStream_t s, const chip_ctx *c, const Level &L, const V *F, V *d);

// ---- reductions: out[m] = sum a_m . b_m over n entries (m < nd <= 4), FP64
// (only a0, b0 fix the vector type; the others may be nullptr)
template <typename T>
struct same_t {
    using type = T;
};
template <typename V>
void dots(hipStream_t s, chip_ctx *c, int n, int nd, const V *a0, const V *b0, typename same_t<const V *>::type a1,
          typename same_t<const V *>::type b1, typename same_t<const V *>::type a2, typename same_t<const V *>::type b2,
          typename same_t<const V *>::type a3, typename same_t<const V *>::type b3, double *out);

// ---- elementwise ------------------------------------------------------------
void fill(hipStream_t s, double *v, double val, long n);
void fill(hipStream_t s, float *v, float val, long n);
void to_f32(hipStream_t s, float *dst, const double *src, long n);
void scale_half(hipStream_t s, const Level &L);
void add_f32(hipStream_t s, double *y, const float *x, long n);   // y += x
void copy(hipStream_t s, double *dst, const double *src, long n);
void axpy_scalar(hipStream_t s, double *y, const double *x, double a, long n);   // y += a x
// x += r1 (x - h1) + r2 (h1 - h2); d2 (optional) = x - 2 h1 + h2; h2 = h1;
// h1 = old x
void extrapolate(hipStream_t s, double *x, double *h1, double *h2, double r1, double r2, double *d2, long n);
void sub(hipStream_t s, double *r, const double *a, const double *b, long n);    // r = a - b
void sub(hipStream_t s, float *r, const float *a, const float *b, long n);
// vel += alpha del, alpha = sc[1]/sc[0]; if res: res -= alpha AU
// multi_grid's line search on a short vector in one launch: sc = (AU.AU,
// AU.res), vel += (sc[1] / sc[0]) del, resout -= the same times AU (if
// resout); for n where line_search_ok
bool line_search_ok(long n);
template <typename V>
void line_search(hipStream_t s, long n, V *vel, const V *del, V *resout, const V *AU, const V *res, double *sc);
void mg_update(hipStream_t s, double *vel, const double *del, double *res, const double *AU, const double *sc, long n);
void mg_update(hipStream_t s, float *vel, const float *del, float *res, const float *AU, const double *sc, long n);
// Uzawa pieces (Stokes_flow_Incomp.c:393-464)
void uz_z(hipStream_t s, double *z1, const double *BPI, const double *r1, long n);
void uz_search(hipStream_t s, double *s2, const double *z1, const double *s1, const double *num, const double *den,
               int first, long n);
void uz_pupdate(hipStream_t s, double *r2, const double *r1, const double *Ah, double *P, const double *s2,
                const double *num, const double *den, int valid, long n);
void uz_vupdate(hipStream_t s, double *V, const double *u1, const double *num, const double *den, int valid, long n);
void init_residual(hipStream_t s, double *Ah, const double *F, const double *gradP, const double *KV, long n);
void v_from_vector(hipStream_t s, const Level &L, const double *U, float *V);
void conform_velocity(hipStream_t s, const Level &L, double *U);

// ---- energy --------------------------------------------------------------
// per-element min of the advective step and Tmax of T into c->minmax
void timestep_tmax(hipStream_t s, chip_ctx *c, const float *V, const float *T);
void tmax(hipStream_t s, chip_ctx *c, const float *T);
void predictor(hipStream_t s, const Level &L, float *T, float *Tdot, float mult);
void pg_residual(hipStream_t s, const chip_ctx *c, const float *T, const float *Tdot, const float *V, double *Eres);
void pg_gather_correct(hipStream_t s, const chip_ctx *c, const double *Eres, float *T, float *Tdot, float mult);
void conform_temperature(hipStream_t s, const chip_ctx *c, float *T);
void thermal_buoyancy(hipStream_t s, const chip_ctx *c, const float *T, float *buoy);
// volume integrals of v.v and of 1 (finest level) into out2[0], out2[1]
void vrms(hipStream_t s, chip_ctx *c, const float *V, double *out2);
void assemble_forces(hipStream_t s, const chip_ctx *c, const float *buoy, double *F);
void heat_flux(hipStream_t s, chip_ctx *c, const float *T, const float *V);
// averages (Process_velocity.c:179-224): Have[0..noz) T, [noz..2noz) |V|^2, [2noz..3noz) nodal viscosity
void averages(hipStream_t s, chip_ctx *c, float *Have, float *scratch);
void fcopy(hipStream_t s, float *dst, const float *src, long n);

// ---- the learned warm start's inputs and outputs (finest level)
void warm_dT(hipStream_t s, const float *T, const float *Th0, const float *Th1, long n, float *dT);
void e2n(hipStream_t s, const Level &L, const double *Pa, const double *Pb, float *out);
void warm_features(hipStream_t s, const Level &L, const float *T, const float *dT, const float *logeta,
                   const double *U, const double *Uprev, const double *D2, const float *pn, const float *dpn,
                   const float *d2pn, float inv_dT, float inv_u, float inv_p, float inv_2u, float inv_2p,
                   uint16_t *x);
void warm_apply(hipStream_t s, const Level &L, const float *out, double s_eu, double s_ep, double *U, double *P);
void warm_target(hipStream_t s, const Level &L, const double *U, const double *Ue, const float *dpn, double inv_eu,
                 double inv_ep, float *y);

}  // namespace chip
