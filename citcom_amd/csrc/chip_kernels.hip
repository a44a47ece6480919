// gfx950 kernels of the CitcomCU hot path.
//
// Every scatter-add of the reference (element loops adding into nodes) is
// written here as a per-node gather over the node's elements in increasing
// element number, which reproduces the reference's summation order and
// rounding points, so these kernels agree bit for bit with the reference
// wherever the reference itself is deterministic.  Built with
// -ffp-contract=off for that reason.  The stiffness operator is bandwidth
// bound (2 flops per 4-byte coefficient): coefficient streams are read once,
// coalesced, 256-byte aligned; vectors are gathered through L2.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "chip_kernels.h"

namespace chip {

__constant__ double cNv[64];   // N.vpt   (Shape_functions.c:104-110)
__constant__ double cNp[8];    // N.ppt at the element centre
__constant__ double cMv[16];   // M.vpt   (2-D surface shape functions)
__constant__ double cMxv[32];  // Mx.vpt
__constant__ int cOffz[9] = {0, 0, 0, 0, 0, 1, 1, 1, 1};
__constant__ int cOffx[9] = {0, 0, 1, 1, 0, 0, 1, 1, 0};
__constant__ int cOffy[9] = {0, 0, 0, 1, 1, 0, 0, 1, 1};
// local node of offsets (dz, dx, dy): cLoc[dz + 2 dx + 4 dy]
__constant__ int cLoc[8] = {1, 5, 2, 6, 4, 8, 3, 7};

namespace {
constexpr int TPB = 256;
inline int nblk(long n, int tpb = TPB) { return (int)std::max<long>(1, (n + tpb - 1) / tpb); }
}  // namespace

int upload_constants() {
    const double B = 0.57735026918962576451;
    const float b = (float)B;
    const float gp[9][3] = {{0, 0, 0}, {-b, -b, -b}, {b, -b, -b}, {b, b, -b}, {-b, b, -b},
                            {-b, -b, b}, {b, -b, b}, {b, b, b}, {-b, b, b}};
    const float sp[5][2] = {{0, 0}, {-b, -b}, {b, -b}, {b, b}, {-b, b}};
    const int bbv[3][9] = {{0, 1, 2, 2, 1, 1, 2, 2, 1}, {0, 1, 1, 2, 2, 1, 1, 2, 2}, {0, 1, 1, 1, 1, 2, 2, 2, 2}};
    auto lp = [](int p, double y) { return p == 1 ? 0.5 * (1 - y) : (p == 2 ? 0.5 * (1 + y) : 0.0); };
    auto ld = [](int p) { return p == 1 ? -0.5 : (p == 2 ? 0.5 : 0.0); };
    double Nv[64], Np[8], Mv[16], Mxv[32];
    for (int i = 1; i <= 8; i++) {
        for (int j = 1; j <= 8; j++) {
            double v = 1.0;
            for (int d = 1; d <= 3; d++) v *= lp(bbv[d - 1][i], gp[j][d - 1]);
            Nv[8 * (i - 1) + j - 1] = v;
        }
        double v = 1.0;
        for (int d = 1; d <= 3; d++) v *= lp(bbv[d - 1][i], 0.0);
        Np[i - 1] = v;
    }
    // construct_shape_functions, surface part (Shape_functions.c:130-158)
    for (int j = 1; j <= 4; j++)
        for (int k = 1; k <= 4; k++) {
            Mv[4 * (j - 1) + k - 1] = lp(bbv[0][j], sp[k][0]) * lp(bbv[1][j], sp[k][1]);
            Mxv[4 * (j - 1) + k - 1] = ld(bbv[0][j]) * lp(bbv[1][j], sp[k][1]);
            Mxv[16 + 4 * (j - 1) + k - 1] = ld(bbv[1][j]) * lp(bbv[0][j], sp[k][0]);
        }
    if (hipMemcpyToSymbol(HIP_SYMBOL(cNv), Nv, sizeof Nv) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(cNp), Np, sizeof Np) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(cMv), Mv, sizeof Mv) != hipSuccess) return -1;
    if (hipMemcpyToSymbol(HIP_SYMBOL(cMxv), Mxv, sizeof Mxv) != hipSuccess) return -1;
    return 0;
}

// ---------------------------------------------------------------------------
// index helpers

__device__ __forceinline__ int nid(const Geo &g, int k, int i, int j) { return k + g.noz * (i + g.nox * j); }

__device__ __forceinline__ void node_kij(const Geo &g, int n, int &k, int &i, int &j) {
    k = n % g.noz;
    const int t = n / g.noz;
    i = t % g.nox;
    j = t / g.nox;
}

__device__ __forceinline__ void colour_kij(const Geo &g, int c, int q, int &k, int &i, int &j) {
    const int r = q - g.cs[c];
    const int nz = g.cz[c], nx = g.cx[c];
    const int kk = r % nz, t = r / nz;
    k = 2 * kk + (c & 1);
    i = 2 * (t % nx) + ((c >> 1) & 1);
    j = 2 * (t / nx) + ((c >> 2) & 1);
}

__device__ __forceinline__ int colour_of(const Geo &g, int q) {
    int c = 0;
#pragma unroll
    for (int m = 1; m < 8; m++) c += (q >= g.cs[m]);
    return c;
}

// colour of q, or -1 for a padding slot
__device__ __forceinline__ int colour_valid(const Geo &g, int q) {
    if (q >= g.pitch) return -1;
    const int c = colour_of(g, q);
    return (q - g.cs[c] < g.cn[c]) ? c : -1;
}

// Stiffness storage: 64-node tiles in colour order, each tile holding its
// 243 coefficient runs of 64 floats back to back, so one wavefront reads
// 243 contiguous 256-byte lines (62 KB) per row product.
__device__ __forceinline__ size_t ktile(int q) { return (size_t)(q >> 6) * (KROW * 64) + (q & 63); }

// Workgroups are dispatched to the 8 XCDs round-robin.  This gives XCD x a
// contiguous range of the grid's tiles, so the vector entries that
// neighbouring tiles share are fetched into one L2 rather than several.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int x = b & 7, y = b >> 3, qn = nb >> 3, rn = nb & 7;
    return (x < rn ? x * (qn + 1) : rn * (qn + 1) + (x - rn) * qn) + y;
}

// full-row product of the node at colour index q, (k,i,j), with u.  The
// neighbour order (dy, dx, dz ascending) and the per-block expression are
// those of the CPU restatement.  Loads are issued one dy-plane (9 blocks) at
// a time so that each thread keeps 81 coefficients and 27 vector entries in
// flight.  Missing neighbours carry zero coefficients and read the node's
// own entries instead.
template <typename A, typename KA, typename UA>
__device__ __forceinline__ void row_product_t(const Geo &g, KA kat, int k, int i, int j, UA u, A out[3]) {
    const int self = nid(g, k, i, j);
    out[0] = out[1] = out[2] = A(0);
#pragma unroll
    for (int dy = -1; dy <= 1; dy++) {
        const bool oy = (unsigned)(j + dy) < (unsigned)g.noy;
        float kv[81];
        A uv[27];
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const bool ox = oy && (unsigned)(i + dx) < (unsigned)g.nox;
#pragma unroll
            for (int dz = -1; dz <= 1; dz++) {
                const int b = (dz + 1) + 3 * (dx + 1);
                const bool ok = ox && (unsigned)(k + dz) < (unsigned)g.noz;
                const int m = ok ? self + dz + g.noz * (dx + g.nox * dy) : self;
                uv[3 * b] = u(3 * m);
                uv[3 * b + 1] = u(3 * m + 1);
                uv[3 * b + 2] = u(3 * m + 2);
                const int s0 = 9 * (b + 9 * (dy + 1));
#pragma unroll
                for (int t = 0; t < 9; t++) kv[9 * b + t] = kat(s0 + t);
            }
        }
#pragma unroll
        for (int b = 0; b < 9; b++)
#pragma unroll
            for (int r = 0; r < 3; r++)
                out[r] += kv[9 * b + 3 * r] * uv[3 * b] + kv[9 * b + 3 * r + 1] * uv[3 * b + 1] +
                          kv[9 * b + 3 * r + 2] * uv[3 * b + 2];
    }
}

template <typename V>
__device__ __forceinline__ void row_product(const Geo &g, const float *__restrict__ K, int q, int k, int i, int j,
                                            const V *__restrict__ u, V out[3]) {
    const float *kp = K + ktile(q);
    row_product_t<V>(
        g, [kp](int ci) { return kp[ci * 64]; }, k, i, j, [u](int idx) { return u[idx]; }, out);
}

// the same with the FP16 (row-scaled) coefficients, widened to FP32
__device__ __forceinline__ void row_product(const Geo &g, const _Float16 *__restrict__ K, int q, int k, int i, int j,
                                            const float *__restrict__ u, float out[3]) {
    const _Float16 *kp = K + ktile(q);
    row_product_t<float>(
        g, [kp](int ci) { return (float)kp[ci * 64]; }, k, i, j, [u](int idx) { return u[idx]; }, out);
}

template <typename V>
__device__ __forceinline__ void strip3(unsigned f, V v[3]) {
    if (f & VBX) v[0] = V(0);
    if (f & VBY) v[1] = V(0);
    if (f & VBZ) v[2] = V(0);
}

// ---------------------------------------------------------------------------
// Smoother row sums.  The 8-colour smoother sums a row in a fixed order: the
// block terms t[b] = k0*u0 + k1*u1 + k2*u2 of each y-plane p = dy+1
// (b = (dz+1) + 3(dx+1); absent neighbours and b >= 9 give 0) are summed in a
// pairwise tree over 16 slots (t[q] += t[q+w], w = 1, 2, 4, 8, q a multiple
// of 2w) into P[p], and the row sum is (P[0] + P[1]) + P[2].  A thread forms
// a plane's tree in registers (PairSum16); a 16-lane row of a wavefront
// forms the same tree with DPP row shifts (IEEE addition is commutative, so
// both give the same bits).  The matvec keeps row_product's sequential order.

// tree over 16 slots formed term by term: push<B>(t[B]) for B = 0..15 keeps
// one partial sum per tree level; after B = 15, total is t[0] of the tree
template <typename V>
struct PairSum16 {
    V st[4];
    V total;
    template <int B>
    __device__ __forceinline__ void push(V v) {
#pragma unroll
        for (int l = 0; l < 4; l++) {
            if (B & (1 << l)) {
                v = st[l] + v;
            } else {
                st[l] = v;
                return;
            }
        }
        total = v;
    }
};

// lane i of each 16-lane row receives lane i+W's value (row_shl:W)
template <int W>
__device__ __forceinline__ double dpp_row_shl(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int rlo = __builtin_amdgcn_update_dpp(0, lo, 0x100 + W, 0xF, 0xF, false);
    const int rhi = __builtin_amdgcn_update_dpp(0, hi, 0x100 + W, 0xF, 0xF, false);
    return __hiloint2double(rhi, rlo);
}

// PairSum16 over the 16 lanes of a DPP row; lane 16k holds the row's total
__device__ __forceinline__ double tree16_lanes(double v) {
    v = v + dpp_row_shl<1>(v);
    v = v + dpp_row_shl<2>(v);
    v = v + dpp_row_shl<4>(v);
    v = v + dpp_row_shl<8>(v);
    return v;
}

__device__ __forceinline__ double lane_value(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// a vector entry: a plain load, or (COHERENT) one that reads the copy in L2
// rather than the CU's L1, for values other workgroups of the same kernel
// have written (k_gs_xcd)
template <bool COHERENT, typename V>
__device__ __forceinline__ V vload(const V *__restrict__ p) {
    if constexpr (COHERENT)
        return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *p;
}

// Plane P = dy+1 of the three rows of node (k, i, j), coefficients kp (its
// tile column), into P[r]; arithmetic in the vector type V.
// The three parts of plane_rows: the neighbours' vector entries (and which
// neighbours exist), the plane's 81 coefficients, and the tree sums
template <int P, typename V, bool COHERENT>
__device__ __forceinline__ void plane_gather(const Geo &g, int k, int i, int j, int self, const V *__restrict__ u,
                                             V uv[27], bool okb[9]) {
    constexpr int DY = P - 1;
    const bool oy = (unsigned)(j + DY) < (unsigned)g.noy;
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
        const bool ox = oy && (unsigned)(i + dx) < (unsigned)g.nox;
#pragma unroll
        for (int dz = -1; dz <= 1; dz++) {
            const int b = (dz + 1) + 3 * (dx + 1);
            const bool ok = ox && (unsigned)(k + dz) < (unsigned)g.noz;
            const int m = ok ? self + dz + g.noz * (dx + g.nox * DY) : self;
            okb[b] = ok;
            uv[3 * b] = vload<COHERENT>(u + 3 * m);
            uv[3 * b + 1] = vload<COHERENT>(u + 3 * m + 1);
            uv[3 * b + 2] = vload<COHERENT>(u + 3 * m + 2);
        }
    }
}

template <int P, typename KT>
__device__ __forceinline__ void plane_coefs(const KT *__restrict__ kp, float kv[81]) {
#pragma unroll
    for (int c = 0; c < 81; c++) kv[c] = (float)kp[(81 * P + c) * 64];
}

template <typename V>
__device__ __forceinline__ void plane_sums(const float kv[81], const V uv[27], const bool okb[9], V out[3]) {
    PairSum16<V> ps[3];
#define CHIP_TERM(B, R)                                                                                            \
    (okb[B] ? kv[9 * (B) + 3 * (R)] * uv[3 * (B)] + kv[9 * (B) + 3 * (R) + 1] * uv[3 * (B) + 1] +                  \
                  kv[9 * (B) + 3 * (R) + 2] * uv[3 * (B) + 2]                                                      \
            : V(0))
#define CHIP_PUSH(B)                                  \
    ps[0].template push<B>(CHIP_TERM(B, 0));          \
    ps[1].template push<B>(CHIP_TERM(B, 1));          \
    ps[2].template push<B>(CHIP_TERM(B, 2))
    CHIP_PUSH(0); CHIP_PUSH(1); CHIP_PUSH(2); CHIP_PUSH(3); CHIP_PUSH(4);
    CHIP_PUSH(5); CHIP_PUSH(6); CHIP_PUSH(7); CHIP_PUSH(8);
#undef CHIP_PUSH
#undef CHIP_TERM
#pragma unroll
    for (int r = 0; r < 3; r++) {
        ps[r].template push<9>(V(0));
        ps[r].template push<10>(V(0));
        ps[r].template push<11>(V(0));
        ps[r].template push<12>(V(0));
        ps[r].template push<13>(V(0));
        ps[r].template push<14>(V(0));
        ps[r].template push<15>(V(0));
        out[r] = ps[r].total;
    }
}

template <int P, typename V, typename KT>
__device__ __forceinline__ void plane_rows(const Geo &g, const KT *__restrict__ kp, int k, int i, int j, int self,
                                           const V *__restrict__ u, V out[3]) {
    float kv[81];
    V uv[27];
    bool okb[9];
    plane_gather<P, V, false>(g, k, i, j, self, u, uv, okb);
    plane_coefs<P, KT>(kp, kv);
    plane_sums<V>(kv, uv, okb, out);
}

// matvec: n_assemble_del2_u (Element_calculations.c:552-621) with strip;
// optionally res = rhs - Au.  One thread per node: the three rows share the
// 27 vector gathers.
template <typename V, typename KT>
__global__ void __launch_bounds__(256) k_matvec(Geo g, const KT *__restrict__ K, const unsigned *__restrict__ NODE,
                                                const V *__restrict__ u, V *__restrict__ Au,
                                                const V *__restrict__ rhs, V *__restrict__ res,
                                                const V *__restrict__ BIs) {
    const int q = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int col = colour_valid(g, q);
    if (col < 0) return;
    int k, i, j;
    colour_kij(g, col, q, k, i, j);
    V out[3];
    row_product(g, K, q, k, i, j, u, out);
    const int n = nid(g, k, i, j);
    if (BIs) {  // rows stored scaled by BI: K d = (BI K) d / BI
        out[0] = out[0] / BIs[3 * n];
        out[1] = out[1] / BIs[3 * n + 1];
        out[2] = out[2] / BIs[3 * n + 2];
    }
    strip3(NODE[n], out);
    Au[3 * n] = out[0];
    Au[3 * n + 1] = out[1];
    Au[3 * n + 2] = out[2];
    if (res) {
        res[3 * n] = rhs[3 * n] - out[0];
        res[3 * n + 1] = rhs[3 * n + 1] - out[1];
        res[3 * n + 2] = rhs[3 * n + 2] - out[2];
    }
}

// matvec of large levels with three wavefronts per 64-node tile: wavefront
// p loads y-plane p's 81 coefficients (coalesced 256-byte runs over the
// tile) and 27 vector entries and forms its 27 block terms (the expression
// of row_product_t); planes 1 and 2 hand their terms to wavefront 0 through
// LDS, which adds all 27 terms in row_product_t's sequential order, so the
// result has the same bits as k_matvec's while three planes' loads are in
// flight at once.
template <typename V, typename KT>
__global__ void __launch_bounds__(192) k_matvec3(Geo g, const KT *__restrict__ K, const unsigned *__restrict__ NODE,
                                                 const V *__restrict__ u, V *__restrict__ Au,
                                                 const V *__restrict__ rhs, V *__restrict__ res,
                                                 const V *__restrict__ BIs) {
    __shared__ V Ts[2][27][64];   // [plane - 1][block * 3 + row][lane]
    const int lane = threadIdx.x & 63, P = threadIdx.x >> 6;
    const int q = xcd_tile(blockIdx.x, gridDim.x) * 64 + lane;
    const int col = colour_valid(g, q);
    int k = 0, i = 0, j = 0;
    if (col >= 0) colour_kij(g, col, q, k, i, j);
    const int self = nid(g, k, i, j);
    V T[27];
    if (col >= 0) {
        const KT *kp = K + ktile(q);
        const int dy = P - 1;
        const bool oy = (unsigned)(j + dy) < (unsigned)g.noy;
        float kv[81];
        V uv[27];
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const bool ox = oy && (unsigned)(i + dx) < (unsigned)g.nox;
#pragma unroll
            for (int dz = -1; dz <= 1; dz++) {
                const int b = (dz + 1) + 3 * (dx + 1);
                const bool ok = ox && (unsigned)(k + dz) < (unsigned)g.noz;
                const int m = ok ? self + dz + g.noz * (dx + g.nox * dy) : self;
                uv[3 * b] = u[3 * m];
                uv[3 * b + 1] = u[3 * m + 1];
                uv[3 * b + 2] = u[3 * m + 2];
            }
        }
#pragma unroll
        for (int c = 0; c < 81; c++) kv[c] = (float)kp[(81 * P + c) * 64];
#pragma unroll
        for (int b = 0; b < 9; b++)
#pragma unroll
            for (int r = 0; r < 3; r++)
                T[3 * b + r] = kv[9 * b + 3 * r] * uv[3 * b] + kv[9 * b + 3 * r + 1] * uv[3 * b + 1] +
                               kv[9 * b + 3 * r + 2] * uv[3 * b + 2];
        if (P > 0)
#pragma unroll
            for (int t = 0; t < 27; t++) Ts[P - 1][t][lane] = T[t];
    }
    __syncthreads();
    if (P != 0 || col < 0) return;
    V out[3] = {V(0), V(0), V(0)};
#pragma unroll
    for (int b = 0; b < 9; b++)
#pragma unroll
        for (int r = 0; r < 3; r++) out[r] += T[3 * b + r];
#pragma unroll
    for (int pl = 0; pl < 2; pl++)
#pragma unroll
        for (int b = 0; b < 9; b++)
#pragma unroll
            for (int r = 0; r < 3; r++) out[r] += Ts[pl][3 * b + r][lane];
    const int n = self;
    if (BIs) {  // rows stored scaled by BI: K d = (BI K) d / BI
        out[0] = out[0] / BIs[3 * n];
        out[1] = out[1] / BIs[3 * n + 1];
        out[2] = out[2] / BIs[3 * n + 2];
    }
    strip3(NODE[n], out);
    Au[3 * n] = out[0];
    Au[3 * n + 1] = out[1];
    Au[3 * n + 2] = out[2];
    if (res) {
        res[3 * n] = rhs[3 * n] - out[0];
        res[3 * n + 1] = rhs[3 * n + 1] - out[1];
        res[3 * n + 2] = rhs[3 * n + 2] - out[2];
    }
}

// The same product with all three planes' coefficients and vector entries
// loaded before the first addition (one memory round trip instead of
// three; the sums keep row_product's order).  For small levels, whose
// matvec is latency-bound: 243 + 81 live values per thread, one wavefront
// per workgroup.
template <typename V>
__device__ __forceinline__ void row_product_all(const Geo &g, const float *__restrict__ kp, int k, int i, int j,
                                                const V *__restrict__ u, V out[3]) {
    const int self = nid(g, k, i, j);
    float kv[243];
    V uv[81];
#pragma unroll
    for (int dy = -1; dy <= 1; dy++) {
        const int P = dy + 1;
        const bool oy = (unsigned)(j + dy) < (unsigned)g.noy;
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const bool ox = oy && (unsigned)(i + dx) < (unsigned)g.nox;
#pragma unroll
            for (int dz = -1; dz <= 1; dz++) {
                const int b = (dz + 1) + 3 * (dx + 1);
                const bool ok = ox && (unsigned)(k + dz) < (unsigned)g.noz;
                const int m = ok ? self + dz + g.noz * (dx + g.nox * dy) : self;
                uv[27 * P + 3 * b] = u[3 * m];
                uv[27 * P + 3 * b + 1] = u[3 * m + 1];
                uv[27 * P + 3 * b + 2] = u[3 * m + 2];
                const int s0 = 9 * (b + 9 * P);
#pragma unroll
                for (int t = 0; t < 9; t++) kv[s0 + t] = kp[(s0 + t) * 64];
            }
        }
    }
    out[0] = out[1] = out[2] = V(0);
#pragma unroll
    for (int P = 0; P < 3; P++)
#pragma unroll
        for (int b = 0; b < 9; b++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const float *kb = kv + 81 * P + 9 * b + 3 * r;
                const V *ub = uv + 27 * P + 3 * b;
                out[r] += kb[0] * ub[0] + kb[1] * ub[1] + kb[2] * ub[2];
            }
}

constexpr int MV_SMALL_NNO = 20000;   // levels up to this size use k_matvec_small

template <typename V>
__global__ void __launch_bounds__(64) k_matvec_small(Geo g, const float *__restrict__ K,
                                                     const unsigned *__restrict__ NODE, const V *__restrict__ u,
                                                     V *__restrict__ Au, const V *__restrict__ rhs,
                                                     V *__restrict__ res, V *__restrict__ ucopy) {
    const int q = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int col = colour_valid(g, q);
    if (col < 0) return;
    int k, i, j;
    colour_kij(g, col, q, k, i, j);
    V out[3];
    row_product_all(g, K + ktile(q), k, i, j, u, out);
    const int n = nid(g, k, i, j);
    strip3(NODE[n], out);
    Au[3 * n] = out[0];
    Au[3 * n + 1] = out[1];
    Au[3 * n + 2] = out[2];
    if (res) {
        res[3 * n] = rhs[3 * n] - out[0];
        res[3 * n + 1] = rhs[3 * n + 1] - out[1];
        res[3 * n + 2] = rhs[3 * n + 2] - out[2];
    }
    if (ucopy) {   // u copied to ucopy in the same launch
        ucopy[3 * n] = u[3 * n];
        ucopy[3 * n + 1] = u[3 * n + 1];
        ucopy[3 * n + 2] = u[3 * n + 2];
    }
}

template <typename V>
void matvec(hipStream_t s, const Level &L, const V *u, V *Au, const V *rhs, V *res, hipEvent_t e0, hipEvent_t e1,
            V *ucopy) {
    if (!L.Kh && L.g.nno <= MV_SMALL_NNO) {
        auto kern = k_matvec_small<V>;
        const float *K = L.K;
        if (e0)
            hipExtLaunchKernelGGL(kern, dim3(nblk(L.g.pitch, 64)), dim3(64), 0, s, e0, e1, 0, L.g, K, L.NODE, u, Au,
                                  rhs, res, ucopy);
        else
            hipLaunchKernelGGL(kern, dim3(nblk(L.g.pitch, 64)), dim3(64), 0, s, L.g, K, L.NODE, u, Au, rhs, res,
                               ucopy);
        return;
    }
    if (ucopy) {   // large levels: a copy launch of its own
        if constexpr (sizeof(V) == 8)
            copy(s, (double *)ucopy, (const double *)u, 3L * L.g.nno);
        else
            fcopy(s, (float *)ucopy, (const float *)u, 3L * L.g.nno);
    }
    const int tiles = L.g.pitch / 64;   // the colour index space is whole 64-node tiles
    if constexpr (sizeof(V) == 4) {
        if (L.Kh) {
            const _Float16 *Kh = (const _Float16 *)L.Kh;
            const V *BIs = (const V *)L.BIf;
            auto kern = k_matvec3<V, _Float16>;
            if (e0)
                hipExtLaunchKernelGGL(kern, dim3(tiles), dim3(192), 0, s, e0, e1, 0, L.g, Kh, L.NODE, u, Au, rhs, res,
                                      BIs);
            else
                hipLaunchKernelGGL(kern, dim3(tiles), dim3(192), 0, s, L.g, Kh, L.NODE, u, Au, rhs, res, BIs);
            return;
        }
    }
    const V *none = nullptr;
    const float *K = L.K;
    auto kern = k_matvec3<V, float>;
    if (e0)
        hipExtLaunchKernelGGL(kern, dim3(tiles), dim3(192), 0, s, e0, e1, 0, L.g, K, L.NODE, u, Au, rhs, res, none);
    else
        hipLaunchKernelGGL(kern, dim3(tiles), dim3(192), 0, s, L.g, K, L.NODE, u, Au, rhs, res, none);
}
template void matvec<double>(hipStream_t, const Level &, const double *, double *, const double *, double *,
                             hipEvent_t, hipEvent_t, double *);
template void matvec<float>(hipStream_t, const Level &, const float *, float *, const float *, float *, hipEvent_t,
                            hipEvent_t, float *);

// Gauss-Seidel: gauss_seidel (General_matrix_functions.c:1160-1361) as eight
// colour passes; the node-block update t = (F - (K d)_node) * BI, rounded to
// float as the reference's temp[] is, then d += t.  One workgroup per 64-node
// tile of one colour: wavefront p forms y-plane p of the three rows of its 64
// nodes (81 coalesced coefficient loads and 27 vector gathers, all issued
// together), the partial sums meet in LDS, and wavefront r updates row r.
// Every row is formed before any unknown is written (workgroup barrier), as
// the reference updates all three unknowns of a node from the same Ad.
template <typename V, typename KT>
__global__ void __launch_bounds__(192) k_gs_colour(Geo g, int c, const KT *__restrict__ K,
                                                   const V *__restrict__ BI, V *d, const V *__restrict__ F) {
    __shared__ V Ps[3][3][64];   // [plane][row][node]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int q = g.cs[c] + xcd_tile(blockIdx.x, gridDim.x) * 64 + lane;
    const bool act = q < g.cs[c] + g.cn[c];
    int k = 0, i = 0, j = 0;
    if (act) colour_kij(g, c, q, k, i, j);
    const int self = nid(g, k, i, j);
    // row w's F, BI and d, read before the barrier: only this thread
    // writes d[3 self + w] in this pass
    V Fr = V(0), Br = V(0), dr = V(0);
    if (act) {
        Fr = F[3 * self + w];
        Br = BI[3 * self + w];
        dr = d[3 * self + w];
    }
    if (act) {
        V out[3];
        const KT *kp = K + ktile(q);
        if (w == 0)
            plane_rows<0>(g, kp, k, i, j, self, (const V *)d, out);
        else if (w == 1)
            plane_rows<1>(g, kp, k, i, j, self, (const V *)d, out);
        else
            plane_rows<2>(g, kp, k, i, j, self, (const V *)d, out);
        Ps[w][0][lane] = out[0];
        Ps[w][1][lane] = out[1];
        Ps[w][2][lane] = out[2];
    }
    __syncthreads();
    if (act) {
        const int r = w;
        const V o = (Ps[0][r][lane] + Ps[1][r][lane]) + Ps[2][r][lane];
        // FP16 rows are stored scaled by BI: t = BI F - (BI K) d
        const float t = sizeof(KT) == 2 ? (float)(Fr * Br - o) : (float)((Fr - o) * Br);
        d[3 * self + r] = dr + t;
    }
}

template <typename V>
void gs_colour(hipStream_t s, const Level &L, const V *BI, int colour, V *d, const V *F, hipEvent_t e0, hipEvent_t e1) {
    const int tiles = (L.g.cn[colour] + 63) / 64;
    if constexpr (sizeof(V) == 4) {
        if (L.Kh) {
            const _Float16 *Kh = (const _Float16 *)L.Kh;
            auto kern = k_gs_colour<V, _Float16>;
            if (e0)
                hipExtLaunchKernelGGL(kern, dim3(tiles), dim3(192), 0, s, e0, e1, 0, L.g, colour, Kh, BI, d, F);
            else
                hipLaunchKernelGGL(kern, dim3(tiles), dim3(192), 0, s, L.g, colour, Kh, BI, d, F);
            return;
        }
    }
    const float *K = L.K;
    auto kern = k_gs_colour<V, float>;
    if (e0)
        hipExtLaunchKernelGGL(kern, dim3(tiles), dim3(192), 0, s, e0, e1, 0, L.g, colour, K, BI, d, F);
    else
        hipLaunchKernelGGL(kern, dim3(tiles), dim3(192), 0, s, L.g, colour, K, BI, d, F);
}
template void gs_colour<double>(hipStream_t, const Level &, const double *, int, double *, const double *, hipEvent_t,
                                hipEvent_t);
template void gs_colour<float>(hipStream_t, const Level &, const float *, int, float *, const float *, hipEvent_t,
                               hipEvent_t);

// the FP16 stiffness of the mixed = 2 form: each row scaled by its BI, so
// the coefficients are O(1) (the diagonal block's diagonal is 1) whatever the
// viscosity, then rounded to FP16.  Same tiling as K.
__global__ void __launch_bounds__(TPB) k_scale_half(Geo g, const float *__restrict__ K, const double *__restrict__ BI,
                                                    _Float16 *__restrict__ Kh) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.pitch) return;
    const int col = colour_valid(g, q);
    const size_t base = ktile(q);
    if (col < 0) {
        for (int ci = 0; ci < KROW; ci++) Kh[base + (size_t)ci * 64] = (_Float16)0.0f;
        return;
    }
    int k, i, j;
    colour_kij(g, col, q, k, i, j);
    const int n = nid(g, k, i, j);
    const float b[3] = {(float)BI[3 * n], (float)BI[3 * n + 1], (float)BI[3 * n + 2]};
    for (int ci = 0; ci < KROW; ci++) Kh[base + (size_t)ci * 64] = (_Float16)(K[base + (size_t)ci * 64] * b[(ci / 3) % 3]);
}

void scale_half(hipStream_t s, const Level &L) {
    hipLaunchKernelGGL(k_scale_half, dim3(nblk(L.g.pitch)), dim3(TPB), 0, s, L.g, (const float *)L.K,
                       (const double *)L.BI, (_Float16 *)L.Kh);
}

// ---------------------------------------------------------------------------
// A level whose stiffness fits in LDS (the coarsest, 4x4x2 elements at the
// workloads here) runs every sweep of a gauss_seidel call, and the final
// Ad = K d, in one workgroup.  One wavefront per node: lane 16p + b forms
// block b of y-plane p for the three rows, each 16-lane row sums its plane
// with DPP shifts, lane 0 combines the planes, lanes 0..2 update.

constexpr int SMALL_TPB = 1024;

__host__ __device__ inline size_t kstage_bytes(int nno) { return (((size_t)KROW * nno * 4 + 15) / 16) * 16; }
__host__ __device__ inline size_t vstage_bytes(int nno) { return (size_t)9 * nno * 8 + (size_t)27 * nno * 4; }

__device__ __forceinline__ int compact_base(const Geo &g, int col) {
    int base = 0;
    for (int m = 0; m < col; m++) base += g.cn[m];
    return base;
}

// unit = 1: workgroup j smooths the unit right-hand side e_j into column j
// of the row-major [3 nno][3 nno] matrix d (the coarse operator), without Ad.
__global__ void __launch_bounds__(SMALL_TPB) k_gs_lds(Geo g, const float *__restrict__ K, const double *__restrict__ BI,
                                                      const unsigned *__restrict__ NODE, double *d,
                                                      const double *__restrict__ F, double *__restrict__ Ad, int sweeps,
                                                      int guess, int unit) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nno = g.nno, tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = blockDim.x >> 6;
    float *Ks = (float *)smem;                                   // [243][nno], compact colour order
    double *ds = (double *)(smem + kstage_bytes(nno));           // [3 nno] natural order
    double *Fs = ds + 3 * nno, *Bs = Fs + 3 * nno;
    int *nb = (int *)(Bs + 3 * nno);                             // [nno][27] neighbour node or -1
    for (int q = tid; q < g.pitch; q += blockDim.x) {
        const int col = colour_valid(g, q);
        if (col < 0) continue;
        const int cq = compact_base(g, col) + (q - g.cs[col]);
        int k, i, j;
        colour_kij(g, col, q, k, i, j);
        const int self = nid(g, k, i, j);
        const float *kp = K + ktile(q);
#pragma unroll 27
        for (int ci = 0; ci < KROW; ci++) Ks[ci * nno + cq] = kp[ci * 64];
        for (int sl = 0; sl < 27; sl++) {
            const int dz = sl % 3 - 1, dx = (sl / 3) % 3 - 1, dy = sl / 9 - 1;
            const bool ok = (unsigned)(k + dz) < (unsigned)g.noz && (unsigned)(i + dx) < (unsigned)g.nox &&
                            (unsigned)(j + dy) < (unsigned)g.noy;
            nb[cq * 27 + sl] = ok ? self + dz + g.noz * (dx + g.nox * dy) : -1;
        }
    }
    for (int idx = tid; idx < 3 * nno; idx += blockDim.x) {
        ds[idx] = guess ? d[idx] : 0.0;
        Fs[idx] = unit ? (idx == (int)blockIdx.x ? 1.0 : 0.0) : F[idx];
        Bs[idx] = BI[idx];
    }
    __syncthreads();
    const int p = lane >> 4, b = lane & 15;
    const int sl = (p < 3 && b < 9) ? 9 * p + b : -1;   // neighbour slot of this lane
    for (int cyc = 0; cyc < sweeps; cyc++) {
        int cb = 0;
        for (int c = 0; c < 8; c++) {
            const int cn = g.cn[c];
            for (int t = w; t < cn; t += nw) {   // wave-uniform
                const int cq = cb + t;
                double v0 = 0.0, v1 = 0.0, v2 = 0.0;
                if (sl >= 0) {
                    const int m = nb[cq * 27 + sl];
                    if (m >= 0) {
                        const double u0 = ds[3 * m], u1 = ds[3 * m + 1], u2 = ds[3 * m + 2];
                        const float *kc = Ks + (sl * 9) * nno + cq;
                        v0 = kc[0] * u0 + kc[nno] * u1 + kc[2 * nno] * u2;
                        v1 = kc[3 * nno] * u0 + kc[4 * nno] * u1 + kc[5 * nno] * u2;
                        v2 = kc[6 * nno] * u0 + kc[7 * nno] * u1 + kc[8 * nno] * u2;
                    }
                }
                v0 = tree16_lanes(v0);
                v1 = tree16_lanes(v1);
                v2 = tree16_lanes(v2);
                const double o0 = (lane_value(v0, 0) + lane_value(v0, 16)) + lane_value(v0, 32);
                const double o1 = (lane_value(v1, 0) + lane_value(v1, 16)) + lane_value(v1, 32);
                const double o2 = (lane_value(v2, 0) + lane_value(v2, 16)) + lane_value(v2, 32);
                if (lane < 3) {
                    const int n = nb[cq * 27 + 13];
                    const double o = lane == 0 ? o0 : (lane == 1 ? o1 : o2);
                    const float tt = (float)((Fs[3 * n + lane] - o) * Bs[3 * n + lane]);
                    ds[3 * n + lane] += tt;
                }
            }
            cb += cn;
            __syncthreads();
        }
    }
    // Ad = K d in row_product's sequential order, BC rows stripped
    for (int cq = tid; Ad && cq < nno; cq += blockDim.x) {
        const int n = nb[cq * 27 + 13];
        double out[3] = {0.0, 0.0, 0.0};
        for (int s2 = 0; s2 < 27; s2++) {
            const int m = nb[cq * 27 + s2];
            const int mm = m >= 0 ? m : n;   // absent neighbours have zero coefficients
            const double u0 = ds[3 * mm], u1 = ds[3 * mm + 1], u2 = ds[3 * mm + 2];
            const float *kc = Ks + (s2 * 9) * nno + cq;
#pragma unroll
            for (int r = 0; r < 3; r++)
                out[r] += kc[(3 * r) * nno] * u0 + kc[(3 * r + 1) * nno] * u1 + kc[(3 * r + 2) * nno] * u2;
        }
        strip3(NODE[n], out);
        Ad[3 * n] = out[0];
        Ad[3 * n + 1] = out[1];
        Ad[3 * n + 2] = out[2];
    }
    __syncthreads();
    if (unit)
        for (int idx = tid; idx < 3 * nno; idx += blockDim.x) d[(size_t)idx * 3 * nno + blockIdx.x] = ds[idx];
    else
        for (int idx = tid; idx < 3 * nno; idx += blockDim.x) d[idx] = ds[idx];
}

// smoother calls applied as matrices on levels of at most this many dofs
// (two FP64 matrices of 18 MB at most)
constexpr long AFFINE_MAX_DOF = 1500;

bool affine_ok(const Level &L) { return 3L * L.g.nno <= AFFINE_MAX_DOF; }

// one-workgroup smoother call (k_gs_wg, below): tiles in flight, threads,
// and the largest level it takes.  One CU reads the stiffness through L2 at
// 50-70 GB/s, so at 2601 nodes a call takes 156 us against 111 us as
// per-colour launches of 4.3 us each; below about 1500 nodes it is faster.
constexpr int WG_SLOTS = 4;
constexpr int WG_TPB = WG_SLOTS * 3 * 64;
constexpr int WG_MAX_NNO = 1024;
// largest level of the one-XCD smoother (k_gs_xcd, below): its FP32
// stiffness (972 B a node) stays within about one XCD's 4 MB L2; polls per
// barrier before a launch gives up
constexpr int XCD_MAX_NNO = 6000;
constexpr unsigned XCD_SPIN = 1u << 20;
template <typename V>
__global__ void __launch_bounds__(WG_TPB) k_gs_wg(Geo g, const float *__restrict__ K,
                                                  const unsigned *__restrict__ NODE, const V *__restrict__ BI, V *d,
                                                  const V *__restrict__ F, V *__restrict__ Ad,
                                                  const V *__restrict__ rhs, V *__restrict__ res, int sweeps,
                                                  int guess);

size_t lds_limit() {
    static size_t lim = 0;
    if (!lim) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            lim = (size_t)v;
        else
            lim = 65536;
        if (lim > 65536) {
            (void)hipFuncSetAttribute((const void *)k_gs_lds, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim);
            const size_t ps = (size_t)WG_SLOTS * 9 * 64;   // k_gs_wg's static partial sums
            (void)hipFuncSetAttribute((const void *)k_gs_wg<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(lim - ps * sizeof(double)));
            (void)hipFuncSetAttribute((const void *)k_gs_wg<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(lim - ps * sizeof(float)));
        }
    }
    return lim;
}

int gs_small_kind(const Level &L) {
    const size_t nno = L.g.nno;
    return kstage_bytes(nno) + vstage_bytes(nno) <= lds_limit() ? 1 : 0;
}

void gs_small(hipStream_t s, const Level &L, double *d, const double *F, double *Ad, int sweeps, int guess) {
    const int nno = L.g.nno;
    hipLaunchKernelGGL(k_gs_lds, dim3(1), dim3(SMALL_TPB), kstage_bytes(nno) + vstage_bytes(nno), s, L.g, L.K, L.BI,
                       L.NODE, d, F, Ad, sweeps, guess, 0);
}

// ---------------------------------------------------------------------------
// A small level whose stiffness does not fit in LDS (8x8x8 elements, 729
// nodes, 708 KB of stiffness) runs a whole gauss_seidel call in one
// workgroup: d in LDS, the stiffness read through L2 (resident in the one
// XCD's L2 after the first sweep).  Each colour pass is
// k_gs_colour's, with workgroup barriers in place of kernel boundaries:
// three wavefronts per 64-node tile form the three y-planes (same trees,
// same update expression), WG_SLOTS tiles at a time; then Ad = K d in
// row_product's order, BC rows stripped, and res = rhs - Ad when asked.
// A colour pass of such a level is latency-bound as a kernel of its own.

template <typename V>
__global__ void __launch_bounds__(WG_TPB) k_gs_wg(Geo g, const float *__restrict__ K,
                                                  const unsigned *__restrict__ NODE, const V *__restrict__ BI, V *d,
                                                  const V *__restrict__ F, V *__restrict__ Ad,
                                                  const V *__restrict__ rhs, V *__restrict__ res, int sweeps,
                                                  int guess) {
    extern __shared__ __align__(16) unsigned char smem[];
    V *ds = (V *)smem;   // [3 nno]
    __shared__ V Ps[WG_SLOTS][3][3][64];   // [slot][plane][row][node]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, slot = w / 3, pl = w - 3 * slot;
    const int n3 = 3 * g.nno;
    for (int idx = tid; idx < n3; idx += blockDim.x) ds[idx] = guess ? d[idx] : V(0);
    __syncthreads();
    for (int cyc = 0; cyc < sweeps; cyc++)
        for (int c = 0; c < 8; c++) {
            const int tiles = (g.cn[c] + 63) >> 6;
            for (int t0 = 0; t0 < tiles; t0 += WG_SLOTS) {   // uniform over the workgroup
                const int q = g.cs[c] + (t0 + slot) * 64 + lane;
                const bool act = t0 + slot < tiles && q < g.cs[c] + g.cn[c];
                int k = 0, i = 0, j = 0;
                if (act) colour_kij(g, c, q, k, i, j);
                const int self = nid(g, k, i, j);
                V Fr = V(0), Br = V(0);
                if (act) {
                    Fr = F[3 * self + pl];
                    Br = BI[3 * self + pl];
                    V out[3];
                    const float *kp = K + ktile(q);
                    if (pl == 0)
                        plane_rows<0>(g, kp, k, i, j, self, (const V *)ds, out);
                    else if (pl == 1)
                        plane_rows<1>(g, kp, k, i, j, self, (const V *)ds, out);
                    else
                        plane_rows<2>(g, kp, k, i, j, self, (const V *)ds, out);
                    Ps[slot][pl][0][lane] = out[0];
                    Ps[slot][pl][1][lane] = out[1];
                    Ps[slot][pl][2][lane] = out[2];
                }
                __syncthreads();
                if (act) {
                    const int r = pl;
                    const V o = (Ps[slot][0][r][lane] + Ps[slot][1][r][lane]) + Ps[slot][2][r][lane];
                    const float t = (float)((Fr - o) * Br);
                    ds[3 * self + r] = ds[3 * self + r] + t;
                }
                __syncthreads();
            }
        }
    for (int q = tid; q < g.pitch; q += blockDim.x) {
        const int col = colour_valid(g, q);
        if (col < 0) continue;
        int k, i, j;
        colour_kij(g, col, q, k, i, j);
        const float *kp = K + ktile(q);
        const V *u = ds;
        V out[3];
        row_product_t<V>(
            g, [kp](int ci) { return kp[ci * 64]; }, k, i, j, [u](int idx) { return u[idx]; }, out);
        const int n = nid(g, k, i, j);
        strip3(NODE[n], out);
        Ad[3 * n] = out[0];
        Ad[3 * n + 1] = out[1];
        Ad[3 * n + 2] = out[2];
        if (res) {
            res[3 * n] = rhs[3 * n] - out[0];
            res[3 * n + 1] = rhs[3 * n + 1] - out[1];
            res[3 * n + 2] = rhs[3 * n + 2] - out[2];
        }
    }
    for (int idx = tid; idx < n3; idx += blockDim.x) d[idx] = ds[idx];
}

template <typename V>
size_t wg_stage_bytes(const Level &L) {
    return (size_t)3 * L.g.nno * sizeof(V);
}

template <typename V>
bool gs_wg_ok(const Level &L) {
    return !L.Kh && L.g.nno <= WG_MAX_NNO &&
           wg_stage_bytes<V>(L) + sizeof(V) * WG_SLOTS * 9 * 64 <= lds_limit();
}
template bool gs_wg_ok<double>(const Level &);
template bool gs_wg_ok<float>(const Level &);

template <typename V>
void gs_wg(hipStream_t s, const Level &L, const V *BI, V *d, const V *F, V *Ad, const V *rhs, V *res, int sweeps,
           int guess, hipEvent_t e0, hipEvent_t e1) {
    auto kern = k_gs_wg<V>;
    const float *K = L.K;
    if (e0)
        hipExtLaunchKernelGGL(kern, dim3(1), dim3(WG_TPB), wg_stage_bytes<V>(L), s, e0, e1, 0, L.g, K, L.NODE, BI, d,
                              F, Ad, rhs, res, sweeps, guess);
    else
        hipLaunchKernelGGL(kern, dim3(1), dim3(WG_TPB), wg_stage_bytes<V>(L), s, L.g, K, L.NODE, BI, d, F, Ad, rhs,
                           res, sweeps, guess);
}
template void gs_wg<double>(hipStream_t, const Level &, const double *, double *, const double *, double *,
                            const double *, double *, int, int, hipEvent_t, hipEvent_t);
template void gs_wg<float>(hipStream_t, const Level &, const float *, float *, const float *, float *, const float *,
                           float *, int, int, hipEvent_t, hipEvent_t);

// ---------------------------------------------------------------------------
// A mid-size level (a few thousand nodes, a few MB of stiffness) runs a
// whole gauss_seidel call as one launch of nwg workgroups that share one XCD.
// Workgroups are dispatched to the XCDs round-robin, so blocks 8m (m = 0 ..
// nwg-1) sit on one XCD and share its L2; the grid's other blocks return at
// once.  Each colour pass is k_gs_colour's (same trees, same update
// expression; workgroup m takes the colour's tiles m, m + nwg, ...), and a
// barrier through that L2 stands in for each kernel boundary: every wave
// waits until its stores are acknowledged, then one thread per workgroup
// adds one to a 64-bit arrival counter and polls it (relaxed device-scope
// atomics).  Vector entries that other workgroups write are read with
// device-scope loads, which take the L2 copy rather than the CU's L1; the
// stiffness, F and BI are read-only and stay cached.  The counter is never
// reset: every launch adds a multiple of nwg, so a launch's first arrival
// returns base + (arrivals before it) with base a multiple of nwg.  The first
// barrier also checks that the workgroups share an XCD (the XCC_ID hardware
// register); if they do not, or a poll runs past its limit, the launch stops
// early and writes a nonzero status (2: not on one XCD, 1: poll limit) for
// the host, which reports the multigrid call as failed.  Then Ad = K d in
// row_product's order (and res = rhs - Ad), as after the per-colour passes.

__device__ __forceinline__ unsigned xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20); }

// polls until the counter reaches target; false past the poll limit
__device__ __forceinline__ bool xcd_wait(unsigned long long *cnt, unsigned long long target) {
    for (unsigned s = 0; __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; s++)
        if (s >= XCD_SPIN) return false;
    return true;
}

template <typename V>
__global__ void __launch_bounds__(192) k_gs_xcd(Geo g, const float *__restrict__ K, const unsigned *__restrict__ NODE,
                                                const V *__restrict__ BI, V *d, const V *__restrict__ F,
                                                V *__restrict__ Ad, const V *__restrict__ rhs, V *__restrict__ res,
                                                int sweeps, int guess, XcdBar *bar, double *status) {
    if (blockIdx.x & 7) return;
    const int me = blockIdx.x >> 3, nwg = gridDim.x >> 3;
    __shared__ V Ps[3][3][64];   // [plane][row][node]
    __shared__ unsigned long long base_s;
    __shared__ int stop_s;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (!guess)
        for (int idx = me * 192 + tid; idx < 3 * g.nno; idx += nwg * 192) d[idx] = V(0);
    // first barrier: the launch's base and the one-XCD check
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0) {
        const unsigned seen = __hip_atomic_fetch_or(&bar->mask, 1u << xcc_id(), __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        (void)seen;
        __builtin_amdgcn_s_waitcnt(0);   // the mask bit lands before the arrival
        const unsigned long long old =
            __hip_atomic_fetch_add(&bar->count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long base = old - old % (unsigned long long)nwg;
        base_s = base;
        int stop = xcd_wait(&bar->count, base + nwg) ? 0 : 1;
        if (!stop) {
            const unsigned m = __hip_atomic_load(&bar->mask, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (m & (m - 1)) stop = 2;
        }
        stop_s = stop;
    }
    __syncthreads();
    if (stop_s) {
        if (tid == 0) *status = (double)stop_s;
        return;
    }
    unsigned long long target = base_s + nwg;
    for (int cyc = 0; cyc < sweeps; cyc++)
        for (int c = 0; c < 8; c++) {
            const int tiles = (g.cn[c] + 63) >> 6;
            for (int t = me; t < tiles; t += nwg) {   // uniform over the workgroup
                const int q = g.cs[c] + t * 64 + lane;
                const bool act = q < g.cs[c] + g.cn[c];
                int k = 0, i = 0, j = 0;
                if (act) colour_kij(g, c, q, k, i, j);
                const int self = nid(g, k, i, j);
                V Fr = V(0), Br = V(0), dr = V(0);
                if (act) {
                    Fr = F[3 * self + w];
                    Br = BI[3 * self + w];
                    dr = vload<true>(d + 3 * self + w);
                    const float *kp = K + ktile(q);
                    float kv[81];
                    V uv[27], out[3];
                    bool okb[9];
                    if (w == 0) {
                        plane_gather<0, V, true>(g, k, i, j, self, d, uv, okb);
                        plane_coefs<0, float>(kp, kv);
                    } else if (w == 1) {
                        plane_gather<1, V, true>(g, k, i, j, self, d, uv, okb);
                        plane_coefs<1, float>(kp, kv);
                    } else {
                        plane_gather<2, V, true>(g, k, i, j, self, d, uv, okb);
                        plane_coefs<2, float>(kp, kv);
                    }
                    plane_sums<V>(kv, uv, okb, out);
                    Ps[w][0][lane] = out[0];
                    Ps[w][1][lane] = out[1];
                    Ps[w][2][lane] = out[2];
                }
                __syncthreads();
                if (act) {
                    const int r = w;
                    const V o = (Ps[0][r][lane] + Ps[1][r][lane]) + Ps[2][r][lane];
                    const float tt = (float)((Fr - o) * Br);
                    d[3 * self + r] = dr + tt;
                }
                __syncthreads();
            }
            // every wave's updates are acknowledged before the arrival
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            target += nwg;
            if (tid == 0) {
                __hip_atomic_fetch_add(&bar->count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                stop_s = xcd_wait(&bar->count, target) ? 0 : 1;
            }
            __syncthreads();
            if (stop_s) {
                if (tid == 0) *status = 1.0;
                return;
            }
        }
    // every workgroup has read the mask (after the first barrier) and passed
    // the last barrier: clear it for the next launch
    if (me == 0 && tid == 0) __hip_atomic_store(&bar->mask, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int q = me * 192 + tid; q < g.pitch; q += nwg * 192) {
        const int col = colour_valid(g, q);
        if (col < 0) continue;
        int k, i, j;
        colour_kij(g, col, q, k, i, j);
        const float *kp = K + ktile(q);
        V out[3];
        row_product_t<V>(
            g, [kp](int ci) { return kp[ci * 64]; }, k, i, j, [d](int idx) { return vload<true>(d + idx); }, out);
        const int n = nid(g, k, i, j);
        strip3(NODE[n], out);
        Ad[3 * n] = out[0];
        Ad[3 * n + 1] = out[1];
        Ad[3 * n + 2] = out[2];
        if (res) {
            res[3 * n] = rhs[3 * n] - out[0];
            res[3 * n + 1] = rhs[3 * n + 1] - out[1];
            res[3 * n + 2] = rhs[3 * n + 2] - out[2];
        }
    }
}

// workgroups of a level's k_gs_xcd launch: one per tile of its largest
// colour (at most 16, which gs_xcd_ok checks)
int xcd_groups(const Level &L) {
    int t = 1;
    for (int c = 0; c < 8; c++) t = std::max(t, (L.g.cn[c] + 63) / 64);
    return std::min(t, 16);
}

bool gs_xcd_ok(const Level &L) {
    int t = 1;
    for (int c = 0; c < 8; c++) t = std::max(t, (L.g.cn[c] + 63) / 64);
    return !L.Kh && L.xbar && L.g.nno > WG_MAX_NNO && L.g.nno <= XCD_MAX_NNO && t <= 16;
}

template <typename V>
void gs_xcd(hipStream_t s, const Level &L, const V *BI, V *d, const V *F, V *Ad, const V *rhs, V *res, int sweeps,
            int guess, double *status, hipEvent_t e0, hipEvent_t e1) {
    auto kern = k_gs_xcd<V>;
    const dim3 grid(8 * xcd_groups(L));
    if (e0)
        hipExtLaunchKernelGGL(kern, grid, dim3(192), 0, s, e0, e1, 0, L.g, (const float *)L.K, L.NODE, BI, d, F, Ad,
                              rhs, res, sweeps, guess, L.xbar, status);
    else
        hipLaunchKernelGGL(kern, grid, dim3(192), 0, s, L.g, (const float *)L.K, L.NODE, BI, d, F, Ad, rhs, res,
                           sweeps, guess, L.xbar, status);
}
template void gs_xcd<double>(hipStream_t, const Level &, const double *, double *, const double *, double *,
                             const double *, double *, int, int, double *, hipEvent_t, hipEvent_t);
template void gs_xcd<float>(hipStream_t, const Level &, const float *, float *, const float *, float *, const float *,
                            float *, int, int, double *, hipEvent_t, hipEvent_t);

// The XCC_ID of the active blocks of a k_gs_xcd-shaped launch (8 nwg
// blocks, every eighth one active), for the set-up check that they share
// one XCD.
__global__ void k_xcd_ids(unsigned *ids) {
    if ((blockIdx.x & 7) == 0 && threadIdx.x == 0) ids[blockIdx.x >> 3] = xcc_id();
}

int xcd_colocated(hipStream_t s, int nwg, unsigned *ids_dev, bool *same) {
    *same = false;
    if (nwg < 1 || nwg > 16) return -1;
    unsigned h[16];
    for (int rep = 0; rep < 4; rep++) {
        hipLaunchKernelGGL(k_xcd_ids, dim3(8 * nwg), dim3(64), 0, s, ids_dev);
        if (hipGetLastError() != hipSuccess) return -1;
        if (hipMemcpyAsync(h, ids_dev, nwg * sizeof(unsigned), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return -1;
        for (int m = 1; m < nwg; m++)
            if (h[m] != h[0]) return 0;
    }
    *same = true;
    return 0;
}

// ---------------------------------------------------------------------------
// Box smoother of the FP32 cycles.  A level's nodes are cut into boxes of at
// most BOX nodes a side; one workgroup takes one box, copies it with its
// one-node halo from din into LDS, runs the 8 colour passes of the
// node-block update on the box's own nodes in LDS (the update and row sums
// of k_gs_colour), and writes the box to dout.  Inside a box each colour
// sees the colours before it as updated; across a box face the neighbour's
// values are those of din.  One launch replaces the 8 colour launches of a
// sweep, whose cost on the small levels is launch latency, not bytes.

// segments of at most BOX nodes, as equal as they come
static bool axis_split(int n, int &nseg, unsigned short *s) {
    nseg = (n + BOX - 1) / BOX;
    if (nseg < 1 || nseg > BOX_MAXSEG) return false;
    int pos = 0;
    for (int t = 0; t < nseg; t++) {
        s[t] = (unsigned short)pos;
        pos += n / nseg + (t < n % nseg ? 1 : 0);
    }
    s[nseg] = (unsigned short)n;
    return true;
}

bool box_part(int nz, int nx, int ny, BoxPart &bp) {
    return axis_split(nz, bp.n[0], bp.s[0]) && axis_split(nx, bp.n[1], bp.s[1]) && axis_split(ny, bp.n[2], bp.s[2]);
}

// the nodes of colour c in box (bz, bx, by): first index and count per axis
struct BoxColour {
    int k0, i0, j0, nz, nx, ny;
};
__device__ __forceinline__ BoxColour box_colour(const BoxPart &bp, int bz, int bx, int by, int c) {
    BoxColour r;
    const int z0 = bp.s[0][bz], z1 = bp.s[0][bz + 1];
    const int x0 = bp.s[1][bx], x1 = bp.s[1][bx + 1];
    const int y0 = bp.s[2][by], y1 = bp.s[2][by + 1];
    r.k0 = z0 + ((z0 ^ c) & 1);
    r.i0 = x0 + ((x0 ^ (c >> 1)) & 1);
    r.j0 = y0 + ((y0 ^ (c >> 2)) & 1);
    r.nz = max(0, (z1 - r.k0 + 1) >> 1);
    r.nx = max(0, (x1 - r.i0 + 1) >> 1);
    r.ny = max(0, (y1 - r.j0 + 1) >> 1);
    return r;
}

// colour-ordered index of node (k, i, j) of colour c (inverse of colour_kij)
__device__ __forceinline__ int colour_index(const Geo &g, int c, int k, int i, int j) {
    return g.cs[c] + (k >> 1) + g.cz[c] * ((i >> 1) + g.cx[c] * (j >> 1));
}

// Box tiles: for the (box, colour) tile b 8 + c and y-plane P, each node's
// 81 coefficients of that plane (K[ktile(q)][81 P + cc]) contiguous in
// 16-byte groups of G = 4 (FP32) or 8 (FP16): [tile][P][cc / G][lane][cc % G],
// so a wavefront reads its plane in 21 or 11 coalesced 16-byte loads per
// lane rather than 81 four- or two-byte ones (a wavefront can have at most
// 63 vector loads in flight).  Lanes past the colour's count and the
// padding past cc = 80 hold zeros.
template <typename KT>
struct BoxTile {
    static constexpr int G = 16 / (int)sizeof(KT);
    static constexpr int NG = (81 + G - 1) / G;
    static constexpr size_t PLANE = (size_t)NG * 64 * G;
    static constexpr size_t TILE = 3 * PLANE;
};

size_t box_tile_elems(const Level &L) {
    return (size_t)L.nbox * 8 * (L.Kbox_half ? BoxTile<_Float16>::TILE : BoxTile<float>::TILE);
}

template <typename KT>
__global__ void __launch_bounds__(64) k_box_tiles(Geo g, BoxPart bp, const KT *__restrict__ K, KT *__restrict__ Kb) {
    using BT = BoxTile<KT>;
    const int tile = blockIdx.x, lane = threadIdx.x;
    const int b = tile >> 3, c = tile & 7;
    const int bz = b % bp.n[0], bx = (b / bp.n[0]) % bp.n[1], by = b / (bp.n[0] * bp.n[1]);
    const BoxColour bc = box_colour(bp, bz, bx, by, c);
    KT *dst = Kb + (size_t)tile * BT::TILE + lane * BT::G;
    const bool act = lane < bc.nz * bc.nx * bc.ny;
    const KT *src = K;
    if (act) {
        const int kk = lane % bc.nz, t = lane / bc.nz;
        const int k = bc.k0 + 2 * kk, i = bc.i0 + 2 * (t % bc.nx), j = bc.j0 + 2 * (t / bc.nx);
        src = K + ktile(colour_index(g, c, k, i, j));
    }
    for (int P = 0; P < 3; P++)
        for (int cc = 0; cc < BT::NG * BT::G; cc++)
            dst[P * BT::PLANE + (size_t)(cc / BT::G) * 64 * BT::G + cc % BT::G] =
                act && cc < 81 ? src[(81 * P + cc) * 64] : KT(0);
}

void box_tiles(hipStream_t s, const Level &L) {
    if (!L.Kbox) return;
    const dim3 grid(8 * L.nbox);
    if (L.Kbox_half)
        hipLaunchKernelGGL(k_box_tiles<_Float16>, grid, dim3(64), 0, s, L.g, L.bp, (const _Float16 *)L.Kh,
                           (_Float16 *)L.Kbox);
    else
        hipLaunchKernelGGL(k_box_tiles<float>, grid, dim3(64), 0, s, L.g, L.bp, (const float *)L.K, (float *)L.Kbox);
}

constexpr int BOXH = BOX + 2;   // box side with its halo
static_assert(BOX <= 8, "a box holds at most 64 nodes of one colour");

// one plane's coefficients of one lane, as loaded (16-byte groups)
template <typename KT>
struct BoxCoefs {
    uint4 q[BoxTile<KT>::NG];
    __device__ __forceinline__ void load(const uint4 *__restrict__ p) {
#pragma unroll
        for (int t = 0; t < BoxTile<KT>::NG; t++) q[t] = p[t * 64];
    }
    __device__ __forceinline__ void widen(float kv[81]) const {
        const KT *e = reinterpret_cast<const KT *>(q);
#pragma unroll
        for (int ci = 0; ci < 81; ci++) kv[ci] = (float)e[ci];
    }
};

// One colour pass of a box: wavefront w forms y-plane w of the three rows
// of the colour's (at most 64) nodes with the coefficients kc, then requests
// the coefficients of the colour two ahead (kn) into kc before the workgroup
// barriers, so their latency overlaps this colour's update and the next
// colour's pass.
template <typename KT>
__device__ __forceinline__ void box_pass(const Geo &g, const BoxPart &bp, int bz, int bx, int by, int c, int z0,
                                         int x0, int y0, int hz, int hx, int nz, int nx, float *sd,
                                         const float *sF, const float *sB, float (*Ps)[3][64], BoxCoefs<KT> &kc,
                                         const uint4 *__restrict__ kn) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const BoxColour bc = box_colour(bp, bz, bx, by, c);
    const bool act = lane < bc.nz * bc.nx * bc.ny;
    int k = 0, i = 0, j = 0;
    if (act) {
        const int kk = lane % bc.nz, t = lane / bc.nz;
        k = bc.k0 + 2 * kk;
        i = bc.i0 + 2 * (t % bc.nx);
        j = bc.j0 + 2 * (t / bc.nx);
    }
    const int lself = (k - z0 + 1) + hz * ((i - x0 + 1) + hx * (j - y0 + 1));
    const int lown = (k - z0) + nz * ((i - x0) + nx * (j - y0));
    if (act) {
        // plane w: the 9 neighbours (dz, dx) at dy = w - 1
        float uv[27];
        bool okb[9];
        const int dy = w - 1;
        const bool oy = (unsigned)(j + dy) < (unsigned)g.noy;
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const bool ox = oy && (unsigned)(i + dx) < (unsigned)g.nox;
#pragma unroll
            for (int dz = -1; dz <= 1; dz++) {
                const int bb = (dz + 1) + 3 * (dx + 1);
                okb[bb] = ox && (unsigned)(k + dz) < (unsigned)g.noz;
                const int m = 3 * (lself + dz + hz * (dx + hx * dy));
                uv[3 * bb] = sd[m];
                uv[3 * bb + 1] = sd[m + 1];
                uv[3 * bb + 2] = sd[m + 2];
            }
        }
        float kv[81], out[3];
        kc.widen(kv);
        plane_sums<float>(kv, uv, okb, out);
        Ps[w][0][lane] = out[0];
        Ps[w][1][lane] = out[1];
        Ps[w][2][lane] = out[2];
    }
    // unconditional (the last colour's again near the end), so that no wait
    // is placed on these loads before their use two passes later
    kc.load(kn);
    __syncthreads();
    if (act) {
        const float Fr = sF[3 * lown + w], Br = sB[3 * lown + w];
        const float o = (Ps[0][w][lane] + Ps[1][w][lane]) + Ps[2][w][lane];
        // FP16 rows are stored scaled by BI: t = BI F - (BI K) d
        const float t = sizeof(KT) == 2 ? Fr * Br - o : (Fr - o) * Br;
        sd[3 * lself + w] += t;
    }
    __syncthreads();
}

// One sweep, a workgroup per box.  ZIN: din is zero (not read).
template <typename KT, bool ZIN>
__global__ void __launch_bounds__(192) k_gs_box(Geo g, BoxPart bp, const KT *__restrict__ Kb,
                                                const float *__restrict__ BI, const float *__restrict__ F,
                                                const float *__restrict__ din, float *__restrict__ dout) {
    using BT = BoxTile<KT>;
    __shared__ float sd[BOXH * BOXH * BOXH * 3];
    __shared__ float sF[BOX * BOX * BOX * 3], sB[BOX * BOX * BOX * 3];   // the box's own F and BI
    __shared__ float Ps[3][3][64];   // [plane][row][node]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b = blockIdx.x;
    const int bz = b % bp.n[0], bx = (b / bp.n[0]) % bp.n[1], by = b / (bp.n[0] * bp.n[1]);
    const int z0 = bp.s[0][bz], x0 = bp.s[1][bx], y0 = bp.s[2][by];
    const int hz = bp.s[0][bz + 1] - z0 + 2, hx = bp.s[1][bx + 1] - x0 + 2, hy = bp.s[2][by + 1] - y0 + 2;
    const int nz = hz - 2, nx = hx - 2, ny = hy - 2;
    // this wavefront's plane of colour c's tile, lane's first group
    auto plane = [&](int c) {
        return reinterpret_cast<const uint4 *>(Kb + ((size_t)b * 8 + c) * BT::TILE + w * BT::PLANE) + lane;
    };
    // the first two colours' coefficients, in flight while the box loads
    BoxCoefs<KT> ka, kb;
    ka.load(plane(0));
    kb.load(plane(1));
    for (int t = threadIdx.x; t < nz * nx * ny; t += 192) {
        const int lz = t % nz, r = t / nz, lx = r % nx, ly = r / nx;
        const int n = nid(g, z0 + lz, x0 + lx, y0 + ly);
#pragma unroll
        for (int e = 0; e < 3; e++) {
            sF[3 * t + e] = F[3 * n + e];
            sB[3 * t + e] = BI[3 * n + e];
        }
    }
    // the box and its halo; entries outside the mesh are zero and never read
    for (int t = threadIdx.x; t < hz * hx * hy; t += 192) {
        const int lz = t % hz, r = t / hz, lx = r % hx, ly = r / hx;
        const int k = z0 - 1 + lz, i = x0 - 1 + lx, j = y0 - 1 + ly;
        float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
        if (!ZIN && (unsigned)k < (unsigned)g.noz && (unsigned)i < (unsigned)g.nox && (unsigned)j < (unsigned)g.noy) {
            const int n = nid(g, k, i, j);
            v0 = din[3 * n];
            v1 = din[3 * n + 1];
            v2 = din[3 * n + 2];
        }
        sd[3 * t] = v0;
        sd[3 * t + 1] = v1;
        sd[3 * t + 2] = v2;
    }
    __syncthreads();
#pragma unroll 1
    for (int c = 0; c < 8; c += 2) {
        box_pass<KT>(g, bp, bz, bx, by, c, z0, x0, y0, hz, hx, nz, nx, sd, sF, sB, Ps, ka, plane(min(c + 2, 7)));
        box_pass<KT>(g, bp, bz, bx, by, c + 1, z0, x0, y0, hz, hx, nz, nx, sd, sF, sB, Ps, kb, plane(min(c + 3, 7)));
    }
    // the box's own nodes
    for (int t = threadIdx.x; t < nz * nx * ny; t += 192) {
        const int lz = t % nz, r = t / nz, lx = r % nx, ly = r / nx;
        const int n = nid(g, z0 + lz, x0 + lx, y0 + ly);
        const int m = 3 * ((lz + 1) + hz * ((lx + 1) + hx * (ly + 1)));
        dout[3 * n] = sd[m];
        dout[3 * n + 1] = sd[m + 1];
        dout[3 * n + 2] = sd[m + 2];
    }
}

void gs_box(hipStream_t s, const Level &L, const float *BI, const float *din, float *dout, const float *F,
            hipEvent_t e0, hipEvent_t e1) {
    const dim3 grid(L.nbox), block(192);
    auto launch = [&](auto kern, const auto *Kb) {
        if (e0)
            hipExtLaunchKernelGGL(kern, grid, block, 0, s, e0, e1, 0, L.g, L.bp, Kb, BI, F, din, dout);
        else
            hipLaunchKernelGGL(kern, grid, block, 0, s, L.g, L.bp, Kb, BI, F, din, dout);
    };
    if (L.Kbox_half) {
        const _Float16 *Kb = (const _Float16 *)L.Kbox;
        if (din) launch(k_gs_box<_Float16, false>, Kb);
        else launch(k_gs_box<_Float16, true>, Kb);
    } else {
        const float *Kb = (const float *)L.Kbox;
        if (din) launch(k_gs_box<float, false>, Kb);
        else launch(k_gs_box<float, true>, Kb);
    }
}

// ---------------------------------------------------------------------------
// Coarsest level as a matrix.  The gauss_seidel call on the coarsest level of
// ---------------------------------------------------------------------------
// Exact solve on the FP32 cycles' coarsest level (mixed forms): the level's
// stiffness as a dense symmetric matrix, factored and inverted in FP64 by
// rocSOLVER (potrf, potri) when the operators are rebuilt, applied as FP16
// rows scaled by their largest entry (FP64 sums).
// Fixed components (the rows strip3 zeroes) get an identity row and column,
// so the matrix is SPD and the solve leaves them at the right-hand side's
// value, which the restriction has already zeroed.

// A (column-major n x n, zeroed) <- K of the level: node q's row r, block
// of neighbour m, column cc at A[(3n+r) + (3m+cc) n]
__global__ void __launch_bounds__(TPB) k_dense_assemble(Geo g, const float *__restrict__ K,
                                                        const unsigned *__restrict__ NODE, double *A) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = colour_valid(g, q);
    if (col < 0) return;
    int k, i, j;
    colour_kij(g, col, q, k, i, j);
    const long n3 = 3L * g.nno;
    const int self = nid(g, k, i, j);
    const unsigned fs = NODE[self];
    const unsigned fixbit[3] = {VBX, VBY, VBZ};
    const float *kp = K + ktile(q);
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++)
            for (int dz = -1; dz <= 1; dz++) {
                if ((unsigned)(j + dy) >= (unsigned)g.noy || (unsigned)(i + dx) >= (unsigned)g.nox ||
                    (unsigned)(k + dz) >= (unsigned)g.noz)
                    continue;
                const int m = self + dz + g.noz * (dx + g.nox * dy);
                const unsigned fm = NODE[m];
                const int b = (dz + 1) + 3 * (dx + 1), P = dy + 1;
                for (int r = 0; r < 3; r++)
                    for (int cc = 0; cc < 3; cc++) {
                        const long row = 3L * self + r, cl = 3L * m + cc;
                        double v = (double)kp[(81 * P + 9 * b + 3 * r + cc) * 64];
                        if ((fs & fixbit[r]) || (fm & fixbit[cc])) v = (row == cl) ? 1.0 : 0.0;
                        A[row + cl * n3] = v;
                    }
            }
}

// M (row-major, full) <- the inverse held in A's lower triangle, in FP32
__global__ void __launch_bounds__(TPB) k_sym_to_f32(const double *__restrict__ A, float *__restrict__ M, int n) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * n) return;
    const long r = idx / n, c = idx % n;
    const long lo = r > c ? r : c, hi = r > c ? c : r;   // lower triangle: row >= column
    M[idx] = (float)A[lo + hi * (long)n];
}

// H, s <- M (row-major FP32) with each row scaled by its largest entry and
// rounded to FP16 (entries within 1, half the bytes to apply), one
// wavefront per row
__global__ void __launch_bounds__(256) k_rowscale_f16(const float *__restrict__ M, _Float16 *__restrict__ H,
                                                      float *__restrict__ s, int n) {
    const int t = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const float *row = M + (size_t)i * n;
    float m = 0.0f;
    for (int j = t; j < n; j += 64) m = fmaxf(m, fabsf(row[j]));
#pragma unroll
    for (int w = 1; w < 64; w *= 2) m = fmaxf(m, __shfl_xor(m, w, 64));
    const float sc = m > 0.0f ? m : 1.0f;
    for (int j = t; j < n; j += 64) H[(size_t)i * n + j] = (_Float16)(row[j] / sc);
    if (t == 0) s[i] = sc;
}

// out = s M' F (M' the scaled FP16 rows), one wavefront per row, FP64 sums
// in a fixed order (lane t the terms j = t, t+64, ..., then a butterfly
// over the wavefront)
template <typename V>
__global__ void __launch_bounds__(256) k_dense_apply(const _Float16 *__restrict__ H, const float *__restrict__ s,
                                                     const V *__restrict__ F, V *__restrict__ out, int n) {
    const int t = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    double S = 0.0;
    if (i < n) {
        const _Float16 *row = H + (size_t)i * n;
#pragma unroll 8
        for (int j = t; j < n; j += 64) S = S + (double)(float)row[j] * (double)F[j];
    }
#pragma unroll
    for (int w = 1; w < 64; w *= 2) S = S + __shfl_xor(S, w, 64);
    if (i < n && t == 0) out[i] = (V)(S * (double)s[i]);
}

void dense_assemble(hipStream_t s, const Level &L, double *A) {
    const long n3 = 3L * L.g.nno;
    (void)hipMemsetAsync(A, 0, sizeof(double) * n3 * n3, s);
    hipLaunchKernelGGL(k_dense_assemble, dim3(nblk(L.g.pitch)), dim3(TPB), 0, s, L.g, L.K, L.NODE, A);
}

void sym_to_f32(hipStream_t s, const double *A, float *M, int n) {
    hipLaunchKernelGGL(k_sym_to_f32, dim3((unsigned)(((long)n * n + TPB - 1) / TPB)), dim3(TPB), 0, s, A, M, n);
}

void rowscale_f16(hipStream_t s, const float *M, uint16_t *H, float *sc, int n) {
    hipLaunchKernelGGL(k_rowscale_f16, dim3((n + 3) / 4), dim3(256), 0, s, M, (_Float16 *)H, sc, n);
}

template <typename V>
void dense_apply(hipStream_t s, const uint16_t *H, const float *sc, const V *F, V *out, int n) {
    hipLaunchKernelGGL(k_dense_apply<V>, dim3((n + 3) / 4), dim3(256), 0, s, (const _Float16 *)H, sc, F, out, n);
}
template void dense_apply<double>(hipStream_t, const uint16_t *, const float *, const double *, double *, int);
template void dense_apply<float>(hipStream_t, const uint16_t *, const float *, const float *, float *, int);

// multi_grid (v_steps_low sweeps from a zero guess, General_matrix_functions.c
// :574, :613) is linear in its right-hand side apart from the FP32 rounding
// of each increment, and its Ad is not read afterwards.  Column j of Mc is
// the smoother's result for e_j (one workgroup per column, rebuilt with the
// stiffness); a call is then d = Mc F.

__global__ void k_gs_unit(Geo g, const float *__restrict__ K, const double *__restrict__ BI, double *__restrict__ Ma,
                          double *__restrict__ Mb, int sweeps);

void coarse_build(hipStream_t s, const Level &L, int sweeps) {
    const int nno = L.g.nno;
    if (gs_small_kind(L))
        hipLaunchKernelGGL(k_gs_lds, dim3(3 * nno), dim3(SMALL_TPB), kstage_bytes(nno) + vstage_bytes(nno), s, L.g,
                           L.K, L.BI, L.NODE, L.Mc, nullptr, nullptr, sweeps, 0, 1);
    else
        hipLaunchKernelGGL(k_gs_unit, dim3(3 * nno), dim3(192), (size_t)3 * nno * 8, s, L.g, L.K, L.BI, nullptr, L.Mc,
                           sweeps);
}

// ---------------------------------------------------------------------------
// The level above the coarsest as an affine operator.  Its gauss_seidel calls
// in multi_grid (down_heavy = up_heavy sweeps) are, apart from the FP32
// rounding of each increment, d_out = Ma d_in + Mb F (guess) or Mb F (zero
// guess).  Column j of Ma is the smoother's result for d_in = e_j, F = 0;
// column j of Mb for d_in = 0, F = e_j; one workgroup per column (three
// wavefronts per 64-node tile, as k_gs_colour), d kept in LDS.
// (Ma null: the Mb columns only, one workgroup each)
__global__ void __launch_bounds__(192) k_gs_unit(Geo g, const float *__restrict__ K, const double *__restrict__ BI,
                                                 double *__restrict__ Ma, double *__restrict__ Mb, int sweeps) {
    extern __shared__ __align__(16) unsigned char smem[];
    double *ds = (double *)smem;   // [3 nno]
    __shared__ double Ps[3][3][64];
    const int n3 = 3 * g.nno, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool forA = Ma && (int)blockIdx.x < n3;
    const int jcol = (forA || !Ma) ? blockIdx.x : blockIdx.x - n3;
    for (int idx = threadIdx.x; idx < n3; idx += blockDim.x) ds[idx] = (forA && idx == jcol) ? 1.0 : 0.0;
    __syncthreads();
    for (int cyc = 0; cyc < sweeps; cyc++)
        for (int c = 0; c < 8; c++) {
            const int tiles = (g.cn[c] + 63) >> 6;
            for (int t = 0; t < tiles; t++) {
                const int q = g.cs[c] + t * 64 + lane;
                const bool act = q < g.cs[c] + g.cn[c];
                int k = 0, i = 0, j = 0;
                if (act) colour_kij(g, c, q, k, i, j);
                const int self = nid(g, k, i, j);
                if (act) {
                    double out[3];
                    const float *kp = K + ktile(q);
                    if (w == 0)
                        plane_rows<0>(g, kp, k, i, j, self, (const double *)ds, out);
                    else if (w == 1)
                        plane_rows<1>(g, kp, k, i, j, self, (const double *)ds, out);
                    else
                        plane_rows<2>(g, kp, k, i, j, self, (const double *)ds, out);
                    Ps[w][0][lane] = out[0];
                    Ps[w][1][lane] = out[1];
                    Ps[w][2][lane] = out[2];
                }
                __syncthreads();
                if (act) {
                    const int r = w, e = 3 * self + r;
                    const double Fr = (!forA && e == jcol) ? 1.0 : 0.0;
                    const double o = (Ps[0][r][lane] + Ps[1][r][lane]) + Ps[2][r][lane];
                    const float tt = (float)((Fr - o) * BI[e]);
                    ds[e] = ds[e] + tt;
                }
                __syncthreads();
            }
        }
    double *M = forA ? Ma : Mb;
    for (int idx = threadIdx.x; idx < n3; idx += blockDim.x) M[(size_t)idx * n3 + jcol] = ds[idx];
}

void affine_build(hipStream_t s, const Level &L, int sweeps) {
    const int n3 = 3 * L.g.nno;
    hipLaunchKernelGGL(k_gs_unit, dim3(2 * n3), dim3(192), (size_t)n3 * 8, s, L.g, L.K, L.BI, L.Ma, L.Mb, sweeps);
    if (L.Maf) {   // the FP32 cycles read the maps rounded to FP32: half the bytes
        to_f32(s, L.Maf, L.Ma, (long)n3 * n3);
        to_f32(s, L.Mbf, L.Mb, (long)n3 * n3);
    }
}

// out = Ma din + Mb F (din may be null: zero guess).  One wavefront per row
// i: lane t sums the Ma terms j = t, t+64, ... in order, then the Mb terms
// likewise, into S[t]; the S[t] meet in a pairwise tree (S[q] += S[q+w],
// w = 1 .. 32) formed by a butterfly over the wavefront.
template <typename V, typename M>
__global__ void __launch_bounds__(256) k_affine_apply(const M *__restrict__ Ma, const M *__restrict__ Mb,
                                                      const V *__restrict__ din, const V *__restrict__ F,
                                                      V *__restrict__ out, int n3) {
    const int t = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    double S = 0.0;
    if (i < n3) {   // unrolled so that the loads of several terms are in flight; the sum keeps its order
        if (din)
#pragma unroll 8
            for (int j = t; j < n3; j += 64) S = S + (double)Ma[(size_t)i * n3 + j] * (double)din[j];
#pragma unroll 8
        for (int j = t; j < n3; j += 64) S = S + (double)Mb[(size_t)i * n3 + j] * (double)F[j];
    }
#pragma unroll
    for (int w = 1; w < 64; w *= 2) S = S + __shfl_xor(S, w, 64);
    if (i < n3 && t == 0) out[i] = (V)S;
}

template <typename V>
void affine_apply(hipStream_t s, const Level &L, const V *din, const V *F, V *out) {
    const int n3 = 3 * L.g.nno;
    if (sizeof(V) == 4 && L.Maf)
        hipLaunchKernelGGL((k_affine_apply<V, float>), dim3((n3 + 3) / 4), dim3(256), 0, s, (const float *)L.Maf,
                           (const float *)L.Mbf, din, F, out, n3);
    else
        hipLaunchKernelGGL((k_affine_apply<V, double>), dim3((n3 + 3) / 4), dim3(256), 0, s, (const double *)L.Ma,
                           (const double *)L.Mb, din, F, out, n3);
}
template void affine_apply<double>(hipStream_t, const Level &, const double *, const double *, double *);
template void affine_apply<float>(hipStream_t, const Level &, const float *, const float *, float *);

// d = Mc F: sixteen lanes per row i, lane t summing j = t, t+16, ... in
// order into S[t]; the S[t] meet in a pairwise tree (S[q] += S[q+w],
// w = 1, 2, 4, 8) formed by a butterfly over the 16 lanes
template <typename V>
__global__ void __launch_bounds__(256) k_coarse_apply(const double *__restrict__ M, const V *__restrict__ F,
                                                      V *__restrict__ d, int n3) {
    const int t = threadIdx.x & 15, i = blockIdx.x * 16 + (threadIdx.x >> 4);
    double S = 0.0;
    if (i < n3)
#pragma unroll 8
        for (int j = t; j < n3; j += 16) S = S + M[(size_t)i * n3 + j] * (double)F[j];
    S = S + __shfl_xor(S, 1, 64);
    S = S + __shfl_xor(S, 2, 64);
    S = S + __shfl_xor(S, 4, 64);
    S = S + __shfl_xor(S, 8, 64);
    if (i < n3 && t == 0) d[i] = (V)S;
}

template <typename V>
void coarse_apply(hipStream_t s, const Level &L, const V *F, V *d) {
    const int n3 = 3 * L.g.nno;
    hipLaunchKernelGGL(k_coarse_apply<V>, dim3((n3 + 15) / 16), dim3(256), 0, s, L.Mc, F, d, n3);
}
template void coarse_apply<double>(hipStream_t, const Level &, const double *, double *);
template void coarse_apply<float>(hipStream_t, const Level &, const float *, float *);

// ---------------------------------------------------------------------------
// grid transfers

// project_vector (Solver_multigrid.c:72-159), optionally followed by the
// strip_bcs_from_residual that multi_grid applies to its result
template <typename V>
__global__ void __launch_bounds__(TPB) k_project(Geo gf, Geo gc, const float *__restrict__ TWW,
                                                 const float *__restrict__ MASS, const unsigned *__restrict__ NODE,
                                                 const V *__restrict__ AU, V *__restrict__ AD, int do_strip,
                                                 V *__restrict__ zero) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gc.nno) return;
    int k, i, j;
    node_kij(gc, n, k, i, j);
    // every load first: the 27 fine nodes 2k-1..2k+1 (x, y likewise) that the
    // coarse node's elements reach, and the element weights; then the sums
    // in the reference's order (elements ascending, their 8 fine nodes in
    // local order)
    V fv[27][3];
#pragma unroll
    for (int t = 0; t < 27; t++) {
        const int fz = 2 * k - 1 + t % 3, fx = 2 * i - 1 + (t / 3) % 3, fy = 2 * j - 1 + t / 9;
        const bool in = (unsigned)fz < (unsigned)gf.noz && (unsigned)fx < (unsigned)gf.nox &&
                        (unsigned)fy < (unsigned)gf.noy;
        const int fn = in ? nid(gf, fz, fx, fy) : nid(gf, 2 * k, 2 * i, 2 * j);
        fv[t][0] = AU[3 * fn];
        fv[t][1] = AU[3 * fn + 1];
        fv[t][2] = AU[3 * fn + 2];
    }
    float w[8];
    bool ok[8];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int ez = k - 1 + (m & 1), ex = i - 1 + ((m >> 1) & 1), ey = j - 1 + (m >> 2);
        ok[m] = (unsigned)ez < (unsigned)gc.elz && (unsigned)ex < (unsigned)gc.elx && (unsigned)ey < (unsigned)gc.ely;
        const int a = cLoc[(k - ez) + 2 * (i - ex) + 4 * (j - ey)];
        w[m] = ok[m] ? TWW[(size_t)(ez + gc.elz * (ex + gc.elx * ey)) * TWW_SZ + a] : 0.0f;
    }
    // offsets of local node b (the cOff tables, as compile-time constants)
    constexpr int OZ[9] = {0, 0, 0, 0, 0, 1, 1, 1, 1};
    constexpr int OX[9] = {0, 0, 1, 1, 0, 0, 1, 1, 0};
    constexpr int OY[9] = {0, 0, 0, 1, 1, 0, 0, 1, 1};
    V acc[3] = {V(0), V(0), V(0)};
#pragma unroll
    for (int m = 0; m < 8; m++) {
        if (!ok[m]) continue;
        // element m = (k-1+mz, i-1+mx, j-1+my) holds the coarse node at local
        // offset (1-mz, 1-mx, 1-my); its fine node b is 2 e + that offset +
        // off(b), i.e. (mz + OZ[b], ...) from 2k - 1
        const int oz = m & 1, ox = (m >> 1) & 1, oy = m >> 2;
        V av0 = V(0), av1 = V(0), av2 = V(0);
#pragma unroll
        for (int b = 1; b <= 8; b++) {
            const int t = (oz + OZ[b]) + 3 * (ox + OX[b]) + 9 * (oy + OY[b]);
            av0 += fv[t][0];
            av1 += fv[t][1];
            av2 += fv[t][2];
        }
        acc[0] += w[m] * av0;
        acc[1] += w[m] * av1;
        acc[2] += w[m] * av2;
    }
    const float m = MASS[n];
    acc[0] = acc[0] * m;
    acc[1] = acc[1] * m;
    acc[2] = acc[2] * m;
    if (do_strip) strip3(NODE[n], acc);
    AD[3 * n] = acc[0];
    AD[3 * n + 1] = acc[1];
    AD[3 * n + 2] = acc[2];
    if (zero) {   // the coarse level's smoother starts from zero
        zero[3 * n] = V(0);
        zero[3 * n + 1] = V(0);
        zero[3 * n + 2] = V(0);
    }
}

template <typename V>
void project(hipStream_t s, const Level &fine, const Level &coarse, const V *AU, V *AD, bool do_strip, V *zero) {
    hipLaunchKernelGGL(k_project<V>, dim3(nblk(coarse.g.nno)), dim3(TPB), 0, s, fine.g, coarse.g, coarse.TWW,
                       coarse.MASS, coarse.NODE, AU, AD, (int)do_strip, zero);
}
template void project<double>(hipStream_t, const Level &, const Level &, const double *, double *, bool, double *);
template void project<float>(hipStream_t, const Level &, const Level &, const float *, float *, bool, float *);

// interp_vector + un_inject_vector (Solver_multigrid.c:173-298, 581-634):
// injection, then linear fills along x, z, y in the reference's order,
// optionally followed by multi_grid's strip_bcs_from_residual
template <typename V>
__global__ void __launch_bounds__(TPB) k_interp(Geo gc, Geo gf, const float *__restrict__ iw,
                                                const unsigned *__restrict__ NODE, const V *__restrict__ AD,
                                                V *__restrict__ AU, int do_strip) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gf.nno) return;
    int k, i, j;
    node_kij(gf, n, k, i, j);
    const float *wx = iw, *wz = iw + 2 * gf.nox, *wy = iw + 2 * (gf.nox + gf.noz);
    // the (up to) 8 coarse nodes around the fine node are loaded together,
    // then the linear fills along x, then z, then y: an odd fine index
    // interpolates its two coarse neighbours n1 a + n2 b, an even one takes
    // its single coarse value
    const int kc[2] = {k >> 1, (k + 1) >> 1}, ic[2] = {i >> 1, (i + 1) >> 1}, jc[2] = {j >> 1, (j + 1) >> 1};
    V c[2][2][2][3];   // [y][z][x][component]
#pragma unroll
    for (int yy = 0; yy < 2; yy++)
#pragma unroll
        for (int zz = 0; zz < 2; zz++)
#pragma unroll
            for (int xx = 0; xx < 2; xx++) {
                const int cn = nid(gc, kc[zz], ic[xx], jc[yy]);
#pragma unroll
                for (int d = 0; d < 3; d++) c[yy][zz][xx][d] = AD[3 * cn + d];
            }
    const float nx1 = wx[2 * i], nx2 = wx[2 * i + 1], nz1 = wz[2 * k], nz2 = wz[2 * k + 1];
    const float ny1 = wy[2 * j], ny2 = wy[2 * j + 1];
    V v[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        V lx[2][2];
#pragma unroll
        for (int yy = 0; yy < 2; yy++)
#pragma unroll
            for (int zz = 0; zz < 2; zz++)
                lx[yy][zz] = (i & 1) ? nx1 * c[yy][zz][0][d] + nx2 * c[yy][zz][1][d] : c[yy][zz][0][d];
        V pz[2];
#pragma unroll
        for (int yy = 0; yy < 2; yy++) pz[yy] = (k & 1) ? nz1 * lx[yy][0] + nz2 * lx[yy][1] : lx[yy][0];
        v[d] = (j & 1) ? ny1 * pz[0] + ny2 * pz[1] : pz[0];
    }
    if (do_strip) strip3(NODE[n], v);
    AU[3 * n] = v[0];
    AU[3 * n + 1] = v[1];
    AU[3 * n + 2] = v[2];
}

template <typename V>
void interp(hipStream_t s, const Level &coarse, const Level &fine, const V *AD, V *AU, bool do_strip) {
    hipLaunchKernelGGL(k_interp<V>, dim3(nblk(fine.g.nno)), dim3(TPB), 0, s, coarse.g, fine.g, fine.iw, fine.NODE, AD,
                       AU, (int)do_strip);
}
template void interp<double>(hipStream_t, const Level &, const Level &, const double *, double *, bool);
template void interp<float>(hipStream_t, const Level &, const Level &, const float *, float *, bool);

// ---------------------------------------------------------------------------
// divergence and gradient (Element_calculations.c:691-769)

__global__ void __launch_bounds__(TPB) k_div(Geo g, const float *__restrict__ EG, const double *__restrict__ U,
                                             double *__restrict__ divU) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
    const float *eg = EG + (size_t)e * EG_SZ;
    double acc = 0.0;
#pragma unroll
    for (int a = 1; a <= 8; a++) {
        const int n = nid(g, ez + cOffz[a], ex + cOffx[a], ey + cOffy[a]);
        const int p = 3 * (a - 1);
        acc += eg[p] * U[3 * n] + eg[p + 1] * U[3 * n + 1] + eg[p + 2] * U[3 * n + 2];
    }
    divU[e] = acc;
}

void div(hipStream_t s, const Level &L, const double *U, double *divU) {
    hipLaunchKernelGGL(k_div, dim3(nblk(L.g.nel)), dim3(TPB), 0, s, L.g, L.EG, U, divU);
}

__global__ void __launch_bounds__(TPB) k_grad(Geo g, const float *__restrict__ EG, const unsigned *__restrict__ NODE,
                                              const double *__restrict__ P, double *__restrict__ gradP) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    int k, i, j;
    node_kij(g, n, k, i, j);
    // all eight elements' P and g entries are loaded first (absent elements
    // read element 0 and are masked), then summed in ascending element order
    double pe[8];
    float eg[8][3];
    bool ok[8];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int ez = k - 1 + (m & 1), ex = i - 1 + ((m >> 1) & 1), ey = j - 1 + (m >> 2);
        ok[m] = (unsigned)ez < (unsigned)g.elz && (unsigned)ex < (unsigned)g.elx && (unsigned)ey < (unsigned)g.ely;
        const int e = ok[m] ? ez + g.elz * (ex + g.elx * ey) : 0;
        const int a = cLoc[(k - ez) + 2 * (i - ex) + 4 * (j - ey)];
        const float *ge = EG + (size_t)e * EG_SZ + 3 * ((ok[m] ? a : 1) - 1);
        pe[m] = P[e];
        eg[m][0] = ge[0];
        eg[m][1] = ge[1];
        eg[m][2] = ge[2];
    }
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < 8; m++) {
        if (!ok[m] || 0.0 == pe[m]) continue;   // assemble_grad_p skips P[e] == 0 (:746)
        acc[0] += eg[m][0] * pe[m];
        acc[1] += eg[m][1] * pe[m];
        acc[2] += eg[m][2] * pe[m];
    }
    strip3(NODE[n], acc);
    gradP[3 * n] = acc[0];
    gradP[3 * n + 1] = acc[1];
    gradP[3 * n + 2] = acc[2];
}

void grad(hipStream_t s, const Level &L, const double *P, double *gradP) {
    hipLaunchKernelGGL(k_grad, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, L.EG, L.NODE, P, gradP);
}

__global__ void k_strip(Geo g, const unsigned *__restrict__ NODE, double *v) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    const unsigned f = NODE[n];
    if (f & VBX) v[3 * n] = 0.0;
    if (f & VBY) v[3 * n + 1] = 0.0;
    if (f & VBZ) v[3 * n + 2] = 0.0;
}

void strip(hipStream_t s, const Level &L, double *v) {
    hipLaunchKernelGGL(k_strip, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, L.NODE, v);
}

// ---------------------------------------------------------------------------
// reductions: one launch, per-block partial sums, the last block to finish
// sums the partials in block order (deterministic for a given size)

constexpr int DOT_TPB = 256;
// at most 256 workgroups: the last-workgroup handoff is one atomic per
// workgroup on a single counter
constexpr int DOT_MAXB = 256;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

template <int ND>
__device__ void block_sum(double v[ND], double *sh /* [ND][DOT_TPB/64] */) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int m = 0; m < ND; m++) v[m] = wave_sum(v[m]);
    if (lane == 0)
#pragma unroll
        for (int m = 0; m < ND; m++) sh[m * 16 + w] = v[m];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int m = 0; m < ND; m++) {
            double t = 0.0;
            for (int q = 0; q < nw; q++) t += sh[m * 16 + q];
            v[m] = t;
        }
}

template <int ND, typename V>
__global__ void __launch_bounds__(DOT_TPB) k_dots(int n, const V *__restrict__ a0, const V *__restrict__ b0,
                                                  const V *__restrict__ a1, const V *__restrict__ b1,
                                                  const V *__restrict__ a2, const V *__restrict__ b2,
                                                  const V *__restrict__ a3, const V *__restrict__ b3,
                                                  double *partial, unsigned *counter, double *out) {
    __shared__ double sh[4 * 16];
    __shared__ bool last;
    // four independent partial sums per thread (u = 0..3 over consecutive
    // grid strides), combined as (s0 + s1) + (s2 + s3)
    double v[ND], w[ND][4];
#pragma unroll
    for (int m = 0; m < ND; m++)
#pragma unroll
        for (int u = 0; u < 4; u++) w[m][u] = 0.0;
    const int stride = gridDim.x * blockDim.x;
    for (int i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * stride;
            if (i < n) {
                w[0][u] += (double)a0[i] * (double)b0[i];
                if (ND > 1) w[1][u] += (double)a1[i] * (double)b1[i];
                if (ND > 2) w[2][u] += (double)a2[i] * (double)b2[i];
                if (ND > 3) w[3][u] += (double)a3[i] * (double)b3[i];
            }
        }
    }
#pragma unroll
    for (int m = 0; m < ND; m++) v[m] = (w[m][0] + w[m][1]) + (w[m][2] + w[m][3]);
    block_sum<ND>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < ND; m++) partial[blockIdx.x * 4 + m] = v[m];
        __threadfence();
        const unsigned t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
#pragma unroll
    for (int m = 0; m < ND; m++) v[m] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
        for (int m = 0; m < ND; m++) v[m] += __builtin_nontemporal_load(&partial[b * 4 + m]);
    __syncthreads();
    block_sum<ND>(v, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < ND; m++) out[m] = v[m];
        *counter = 0u;
    }
}

template <typename V>
void dots(hipStream_t s, chip_ctx *c, int n, int nd, const V *a0, const V *b0, typename same_t<const V *>::type a1,
          typename same_t<const V *>::type b1, typename same_t<const V *>::type a2, typename same_t<const V *>::type b2,
          typename same_t<const V *>::type a3, typename same_t<const V *>::type b3, double *out) {
    const int nb = std::min(DOT_MAXB, std::max(1, (n + DOT_TPB * 4 - 1) / (DOT_TPB * 4)));
    switch (nd) {
    case 1: hipLaunchKernelGGL((k_dots<1, V>), dim3(nb), dim3(DOT_TPB), 0, s, n, a0, b0, a1, b1, a2, b2, a3, b3, c->partial, c->counter, out); break;
    case 2: hipLaunchKernelGGL((k_dots<2, V>), dim3(nb), dim3(DOT_TPB), 0, s, n, a0, b0, a1, b1, a2, b2, a3, b3, c->partial, c->counter, out); break;
    case 3: hipLaunchKernelGGL((k_dots<3, V>), dim3(nb), dim3(DOT_TPB), 0, s, n, a0, b0, a1, b1, a2, b2, a3, b3, c->partial, c->counter, out); break;
    default: hipLaunchKernelGGL((k_dots<4, V>), dim3(nb), dim3(DOT_TPB), 0, s, n, a0, b0, a1, b1, a2, b2, a3, b3, c->partial, c->counter, out); break;
    }
}
template void dots<double>(hipStream_t, chip_ctx *, int, int, const double *, const double *, const double *,
                           const double *, const double *, const double *, const double *, const double *, double *);
template void dots<float>(hipStream_t, chip_ctx *, int, int, const float *, const float *, const float *,
                          const float *, const float *, const float *, const float *, const float *, double *);

// multi_grid's line search on a small level in one launch of one workgroup
// (General_matrix_functions.c:626-638): sc = (AU.AU, AU.res), alpha =
// sc[1] / sc[0], vel += alpha del and, with resout, resout -= alpha AU
// (resout may be res: every read of res precedes the barrier).  In place of
// dots + mg_update where the level's vectors are short enough that the two
// launches, not the bytes, set the time: one CU streams about 100 GB/s, so
// at 8192 entries (about 160 KB moved) the one launch takes 2 us.
constexpr int LS_TPB = 1024;
constexpr long LS_MAX_N = 8192;

template <typename V>
__global__ void __launch_bounds__(LS_TPB) k_line_search(long n, V *__restrict__ vel, const V *__restrict__ del,
                                                        V *resout, const V *__restrict__ AU, const V *res,
                                                        double *__restrict__ sc) {
    __shared__ double sh[4 * 16];
    __shared__ double alpha_s;
    double v[2] = {0.0, 0.0};
    for (long i = threadIdx.x; i < n; i += LS_TPB) {
        const double a = (double)AU[i];
        v[0] += a * a;
        v[1] += a * (double)res[i];
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0) {
        sc[0] = v[0];
        sc[1] = v[1];
        alpha_s = v[1] / v[0];
    }
    __syncthreads();
    const V alpha = (V)alpha_s;
    for (long i = threadIdx.x; i < n; i += LS_TPB) {
        vel[i] += alpha * del[i];
        if (resout) resout[i] -= alpha * AU[i];
    }
}

bool line_search_ok(long n) { return n <= LS_MAX_N; }

template <typename V>
void line_search(hipStream_t s, long n, V *vel, const V *del, V *resout, const V *AU, const V *res, double *sc) {
    hipLaunchKernelGGL(k_line_search<V>, dim3(1), dim3(LS_TPB), 0, s, n, vel, del, resout, AU, res, sc);
}
template void line_search<double>(hipStream_t, long, double *, const double *, double *, const double *,
                                  const double *, double *);
template void line_search<float>(hipStream_t, long, float *, const float *, float *, const float *, const float *,
                                 double *);

// ---------------------------------------------------------------------------
// elementwise

template <typename V>
__global__ void k_fill(V *v, V val, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = val;
}
void fill(hipStream_t s, double *v, double val, long n) { hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n)), dim3(TPB), 0, s, v, val, n); }
void fill(hipStream_t s, float *v, float val, long n) { hipLaunchKernelGGL(k_fill<float>, dim3(nblk(n)), dim3(TPB), 0, s, v, val, n); }

// FP64 -> FP32 (the residual handed to the FP32 multigrid)
__global__ void k_to_f32(float *d, const double *a, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)a[i];
}
void to_f32(hipStream_t s, float *dst, const double *src, long n) { hipLaunchKernelGGL(k_to_f32, dim3(nblk(n)), dim3(TPB), 0, s, dst, src, n); }

// y += x with an FP32 correction x (FP64 accumulation of the refinement)
__global__ void k_add_f32(double *y, const float *x, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += (double)x[i];
}
void add_f32(hipStream_t s, double *y, const float *x, long n) { hipLaunchKernelGGL(k_add_f32, dim3(nblk(n)), dim3(TPB), 0, s, y, x, n); }

__global__ void k_copy(double *d, const double *a, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = a[i];
}
void copy(hipStream_t s, double *dst, const double *src, long n) { hipLaunchKernelGGL(k_copy, dim3(nblk(n)), dim3(TPB), 0, s, dst, src, n); }

__global__ void k_fcopy(float *d, const float *a, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = a[i];
}
void fcopy(hipStream_t s, float *dst, const float *src, long n) { hipLaunchKernelGGL(k_fcopy, dim3(nblk(n)), dim3(TPB), 0, s, dst, src, n); }

__global__ void k_axpy(double *y, const double *x, double a, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
void axpy_scalar(hipStream_t s, double *y, const double *x, double a, long n) { hipLaunchKernelGGL(k_axpy, dim3(nblk(n)), dim3(TPB), 0, s, y, x, a, n); }

// the warm start's extrapolation along the sequence of solves: with x the
// last solution and h1, h2 the two before it, x <- x + r1 (x - h1) +
// r2 (h1 - h2); d2 (optional) <- x - 2 h1 + h2, the second difference; then
// h2 <- h1, h1 <- the old x
__global__ void k_extrapolate(double *x, double *h1, double *h2, double r1, double r2, double *d2, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xn = x[i], a = h1[i], b = h2[i];
    x[i] = xn + r1 * (xn - a) + r2 * (a - b);
    if (d2) d2[i] = (xn - a) - (a - b);
    h2[i] = a;
    h1[i] = xn;
}
void extrapolate(hipStream_t s, double *x, double *h1, double *h2, double r1, double r2, double *d2, long n) {
    hipLaunchKernelGGL(k_extrapolate, dim3(nblk(n)), dim3(TPB), 0, s, x, h1, h2, r1, r2, d2, n);
}

template <typename V>
__global__ void k_sub(V *r, const V *a, const V *b, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = a[i] - b[i];
}
void sub(hipStream_t s, double *r, const double *a, const double *b, long n) { hipLaunchKernelGGL(k_sub<double>, dim3(nblk(n)), dim3(TPB), 0, s, r, a, b, n); }
void sub(hipStream_t s, float *r, const float *a, const float *b, long n) { hipLaunchKernelGGL(k_sub<float>, dim3(nblk(n)), dim3(TPB), 0, s, r, a, b, n); }

// multi_grid line search update (General_matrix_functions.c:626-638)
template <typename V>
__global__ void k_mg_update(V *vel, const V *del, V *res, const V *AU, const double *sc, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const V alpha = (V)(sc[1] / sc[0]);
    vel[i] += alpha * del[i];
    if (res) res[i] -= alpha * AU[i];
}
void mg_update(hipStream_t s, double *vel, const double *del, double *res, const double *AU, const double *sc, long n) {
    hipLaunchKernelGGL(k_mg_update<double>, dim3(nblk(n)), dim3(TPB), 0, s, vel, del, res, AU, sc, n);
}
void mg_update(hipStream_t s, float *vel, const float *del, float *res, const float *AU, const double *sc, long n) {
    hipLaunchKernelGGL(k_mg_update<float>, dim3(nblk(n)), dim3(TPB), 0, s, vel, del, res, AU, sc, n);
}

__global__ void k_uz_z(double *z1, const double *BPI, const double *r1, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z1[i] = BPI[i] * r1[i];
}
void uz_z(hipStream_t s, double *z1, const double *BPI, const double *r1, long n) { hipLaunchKernelGGL(k_uz_z, dim3(nblk(n)), dim3(TPB), 0, s, z1, BPI, r1, n); }

// s2 = z1 + delta s1, delta = *num / *den (Stokes_flow_Incomp.c:401-411)
__global__ void k_uz_search(double *s2, const double *z1, const double *s1, const double *num, const double *den,
                            int first, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (first) s2[i] = z1[i];
    else {
        const double delta = *num / *den;
        s2[i] = z1[i] + delta * s1[i];
    }
}
void uz_search(hipStream_t s, double *s2, const double *z1, const double *s1, const double *num, const double *den,
               int first, long n) {
    hipLaunchKernelGGL(k_uz_search, dim3(nblk(n)), dim3(TPB), 0, s, s2, z1, s1, num, den, first, n);
}

// r2 = r1 - alpha Ah, P += alpha s2 with alpha = valid ? *num / *den : 0
// (Stokes_flow_Incomp.c:420-431)
__global__ void k_uz_pupdate(double *r2, const double *r1, const double *Ah, double *P, const double *s2,
                             const double *num, const double *den, int valid, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = valid ? *num / *den : 0.0;
    r2[i] = r1[i] - alpha * Ah[i];
    P[i] += alpha * s2[i];
}
void uz_pupdate(hipStream_t s, double *r2, const double *r1, const double *Ah, double *P, const double *s2,
                const double *num, const double *den, int valid, long n) {
    hipLaunchKernelGGL(k_uz_pupdate, dim3(nblk(n)), dim3(TPB), 0, s, r2, r1, Ah, P, s2, num, den, valid, n);
}

// V -= alpha u1 (Stokes_flow_Incomp.c:433-434)
__global__ void k_uz_vupdate(double *V, const double *u1, const double *num, const double *den, int valid, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double alpha = valid ? *num / *den : 0.0;
    V[i] -= alpha * u1[i];
}
void uz_vupdate(hipStream_t s, double *V, const double *u1, const double *num, const double *den, int valid, long n) {
    hipLaunchKernelGGL(k_uz_vupdate, dim3(nblk(n)), dim3(TPB), 0, s, V, u1, num, den, valid, n);
}

// Ah = F - gradP - K V (Stokes_flow_Incomp.c:357-358)
__global__ void k_init_residual(double *Ah, const double *F, const double *gP, const double *KV, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Ah[i] = F[i] - gP[i] - KV[i];
}
void init_residual(hipStream_t s, double *Ah, const double *F, const double *gradP, const double *KV, long n) {
    hipLaunchKernelGGL(k_init_residual, dim3(nblk(n)), dim3(TPB), 0, s, Ah, F, gradP, KV, n);
}

// v_from_vector (Stokes_flow_Incomp.c:530-552); prescribed velocities are zero
__global__ void k_v_from_vector(Geo g, const unsigned *__restrict__ NODE, const double *__restrict__ U, float *V) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    const unsigned f = NODE[n];
    V[n] = (f & VBX) ? 0.0f : (float)U[3 * n];
    V[g.nno + n] = (f & VBY) ? 0.0f : (float)U[3 * n + 1];
    V[2 * g.nno + n] = (f & VBZ) ? 0.0f : (float)U[3 * n + 2];
}
void v_from_vector(hipStream_t s, const Level &L, const double *U, float *V) {
    hipLaunchKernelGGL(k_v_from_vector, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, L.NODE, U, V);
}

// velocities_conform_bcs (Boundary_conditions.c:993-1023) with VB = 0
void conform_velocity(hipStream_t s, const Level &L, double *U) { strip(s, L, U); }

// ---------------------------------------------------------------------------
// stiffness rebuild

// visc_from_T (Viscosity_structures.c:491-737, rheol 0, 3 and 11) or
// visc_from_mat (:477-489), then the MAX / MIN clip of get_system_viscosity
struct ViscLaw {
    int rheol, tdepv, vmin_on, vmax_on;
    float visc0[4], viscE[4], viscT[4];
    float visc_min, visc_max;
};

__global__ void __launch_bounds__(TPB) k_visc_from_T(Geo g, ViscLaw law, const int *__restrict__ mat,
                                                     const float *__restrict__ T, float *__restrict__ evi) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
    const int l = mat[e] - 1;
    const float tempa = law.visc0[l];
    float TT[9];
#pragma unroll
    for (int a = 1; a <= 8; a++) TT[a] = T[nid(g, ez + cOffz[a], ex + cOffx[a], ey + cOffy[a])];
    for (int jj = 1; jj <= 8; jj++) {
        float v;
        if (!law.tdepv) {
            v = tempa;
        } else {
            float temp = 1.0e-32f;
#pragma unroll
            for (int kk = 1; kk <= 8; kk++) temp = (float)(temp + fmaxf(0.0f, TT[kk]) * cNv[8 * (kk - 1) + jj - 1]);
            if (law.rheol == 11) {
                // eta0 exp(E/(T+T0) - E/(0.5+T0)): the first quotient in
                // float, the second in double (the literal 0.5), :697-719
                const double arg = (double)(law.viscE[l] / (temp + law.viscT[l])) -
                                   (double)law.viscE[l] / (0.5 + (double)law.viscT[l]);
                v = (float)(tempa * exp(arg));
            } else {
                float arg;
                if (law.rheol == 3) arg = law.viscE[l] * (law.viscT[l] - temp);
                else arg = law.viscE[l] * (1.0f - temp);
                v = (float)(tempa * exp((double)arg));
            }
        }
        if (law.vmax_on && v > law.visc_max) v = law.visc_max;
        if (law.vmin_on && v < law.visc_min) v = law.visc_min;
        evi[e * 8 + jj - 1] = v;
    }
}

void visc_from_T(hipStream_t s, const chip_ctx *c, const float *T, float *evi) {
    ViscLaw law{};
    law.rheol = c->p.rheol;
    law.tdepv = c->p.tdepv;
    law.vmin_on = c->p.vmin_on;
    law.vmax_on = c->p.vmax_on;
    law.visc_min = c->p.visc_min;
    law.visc_max = c->p.visc_max;
    for (int i = 0; i < 4; i++) {
        law.visc0[i] = c->p.visc0[i];
        law.viscE[i] = c->p.viscE[i];
        law.viscT[i] = c->p.viscT[i];
    }
    const Geo &g = c->L[c->levmax].g;
    hipLaunchKernelGGL(k_visc_from_T, dim3(nblk(g.nel)), dim3(TPB), 0, s, g, law, c->mat, T, evi);
}

// visc_from_gint_to_nodes (Nodal_mesh.c:583-615)
__global__ void __launch_bounds__(TPB) k_visc_to_nodes(Geo g, const float *__restrict__ EVI,
                                                       const float *__restrict__ TWW, const float *__restrict__ MASS,
                                                       float *__restrict__ VN) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    int k, i, j;
    node_kij(g, n, k, i, j);
    float acc = 0.0f;
    for (int ey = j - 1; ey <= j; ey++) {
        if ((unsigned)ey >= (unsigned)g.ely) continue;
        for (int ex = i - 1; ex <= i; ex++) {
            if ((unsigned)ex >= (unsigned)g.elx) continue;
            for (int ez = k - 1; ez <= k; ez++) {
                if ((unsigned)ez >= (unsigned)g.elz) continue;
                const int e = ez + g.elz * (ex + g.elx * ey);
                double tv = 0.0;
#pragma unroll
                for (int q = 0; q < 8; q++) tv += EVI[e * 8 + q];
                tv = tv / 8;
                const int a = cLoc[(k - ez) + 2 * (i - ex) + 4 * (j - ey)];
                acc = (float)(acc + TWW[(size_t)e * TWW_SZ + a] * tv);
            }
        }
    }
    VN[n] = acc * MASS[n];
}

// project_scalar (Solver_multigrid.c:344-388), result into the coarse VN
__global__ void __launch_bounds__(TPB) k_visc_project(Geo gf, Geo gc, const float *__restrict__ VNf,
                                                      const float *__restrict__ TWW, const float *__restrict__ MASS,
                                                      float *__restrict__ VNc) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= gc.nno) return;
    int k, i, j;
    node_kij(gc, n, k, i, j);
    float acc = 0.0f;
    for (int ey = j - 1; ey <= j; ey++) {
        if ((unsigned)ey >= (unsigned)gc.ely) continue;
        for (int ex = i - 1; ex <= i; ex++) {
            if ((unsigned)ex >= (unsigned)gc.elx) continue;
            for (int ez = k - 1; ez <= k; ez++) {
                if ((unsigned)ez >= (unsigned)gc.elz) continue;
                const int a = cLoc[(k - ez) + 2 * (i - ex) + 4 * (j - ey)];
                const int el = ez + gc.elz * (ex + gc.elx * ey);
                const int fz = 2 * ez + cOffz[a], fx = 2 * ex + cOffx[a], fy = 2 * ey + cOffy[a];
                float average = 0.0f;
#pragma unroll
                for (int b = 1; b <= 8; b++) average += VNf[nid(gf, fz + cOffz[b], fx + cOffx[b], fy + cOffy[b])];
                const float w = (float)((1.0 / 8) * average);
                acc += w * TWW[(size_t)el * TWW_SZ + a];
            }
        }
    }
    VNc[n] = acc * MASS[n];
}

// visc_from_nodes_to_gint (Nodal_mesh.c:617-640)
__global__ void __launch_bounds__(TPB) k_visc_to_gint(Geo g, const float *__restrict__ VN, float *__restrict__ EVI) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
    float vn[9];
#pragma unroll
    for (int a = 1; a <= 8; a++) vn[a] = VN[nid(g, ez + cOffz[a], ex + cOffx[a], ey + cOffy[a])];
#pragma unroll
    for (int i = 1; i <= 8; i++) {
        double tv = 0.0;
#pragma unroll
        for (int j = 1; j <= 8; j++) tv += cNv[8 * (j - 1) + i - 1] * vn[j];
        EVI[e * 8 + i - 1] = (float)tv;
    }
}

void project_viscosity(hipStream_t s, const Level &fine, const Level &coarse) {
    hipLaunchKernelGGL(k_visc_to_nodes, dim3(nblk(fine.g.nno)), dim3(TPB), 0, s, fine.g, fine.EVI, fine.TWW, fine.MASS,
                       fine.VN);
    hipLaunchKernelGGL(k_visc_project, dim3(nblk(coarse.g.nno)), dim3(TPB), 0, s, fine.g, coarse.g, fine.VN,
                       coarse.TWW, coarse.MASS, coarse.VN);
    hipLaunchKernelGGL(k_visc_to_gint, dim3(nblk(coarse.g.nel)), dim3(TPB), 0, s, coarse.g, coarse.VN, coarse.EVI);
}

// get_elt_g (Element_calculations.c:831-934, CART3D)
__global__ void k_elt_g(Geo g, const float *__restrict__ GNX, const float *__restrict__ GDA, float *__restrict__ EG) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const float *gnx = GNX + (size_t)e * GNX_SZ + 192;
    const float tf = 8.0f * GDA[(size_t)e * GDA_SZ + 10];
    const double temp = tf;
#pragma unroll
    for (int a = 1; a <= 8; a++) {
        const int p = 3 * (a - 1);
        EG[(size_t)e * EG_SZ + p] = (float)(-gnx[a - 1] * temp);
        EG[(size_t)e * EG_SZ + p + 1] = (float)(-gnx[8 + a - 1] * temp);
        EG[(size_t)e * EG_SZ + p + 2] = (float)(-gnx[16 + a - 1] * temp);
    }
}

void elt_g(hipStream_t s, const Level &L) {
    hipLaunchKernelGGL(k_elt_g, dim3(nblk(L.g.nel)), dim3(TPB), 0, s, L.g, L.GNX, L.GDA, L.EG);
}

// 3x3 block of the element stiffness elt_k (get_elt_k, Element_calculations.c:
// 133-293, cartesian isotropic) plus get_aug_k (:1127-1163): rows of local
// node ar, columns of local node bc.  The reference evaluates the (a,b) block
// for a <= b and stores its transpose below the diagonal; the aug term is
// added per entry in the (row, column) orientation.
__device__ void elt_block(const float *__restrict__ gnx, const float *__restrict__ gda, const float *__restrict__ evi,
                          const float *__restrict__ eg, int aug_on, double aug, int ar, int bc, double blk[3][3]) {
    const int a = ar < bc ? ar : bc, b = ar < bc ? bc : ar;
    double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, temp = 0.0;
#pragma unroll
    for (int k = 1; k <= 8; k++) {
        float wf = 1.0f * gda[k];
        wf = wf * evi[k - 1];
        const double W = wf;
        const float ax = gnx[8 * (a - 1) + k - 1], ay = gnx[64 + 8 * (a - 1) + k - 1], az = gnx[128 + 8 * (a - 1) + k - 1];
        const float bx = gnx[8 * (b - 1) + k - 1], by = gnx[64 + 8 * (b - 1) + k - 1], bz = gnx[128 + 8 * (b - 1) + k - 1];
        m[0][0] += W * ax * bx;
        m[0][1] += W * ay * bx;
        m[0][2] += W * az * bx;
        m[1][0] += W * ax * by;
        m[1][1] += W * ay * by;
        m[1][2] += W * az * by;
        m[2][0] += W * ax * bz;
        m[2][1] += W * ay * bz;
        m[2][2] += W * az * bz;
    }
#pragma unroll
    for (int k = 1; k <= 8; k++) {
        float wf = 1.0f * gda[k];
        wf = wf * evi[k - 1];
        const double W = wf;
        const float s = gnx[8 * (a - 1) + k - 1] * gnx[8 * (b - 1) + k - 1] +
                        gnx[64 + 8 * (a - 1) + k - 1] * gnx[64 + 8 * (b - 1) + k - 1] +
                        gnx[128 + 8 * (a - 1) + k - 1] * gnx[128 + 8 * (b - 1) + k - 1];
        temp += W * s;
    }
    m[0][0] += temp;
    m[1][1] += temp;
    m[2][2] += temp;
    const bool tr = ar > bc;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) blk[r][c] = tr ? m[c][r] : m[r][c];
    if (aug_on) {
        double visc = 0.0;
#pragma unroll
        for (int q = 0; q < 8; q++) visc += evi[q];
        visc = visc / 8;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) blk[r][c] += visc * aug * eg[3 * (ar - 1) + r] * eg[3 * (bc - 1) + c];
    }
}

// construct_node_ks (Construct_arrays.c:359-535) restated per node: the block
// coupling node n with neighbour m is the float accumulation, in element
// order, of w_r ww_c elt_k[rows of max(n,m)][cols of min(n,m)] (BC weights
// zero the prescribed rows and columns), transposed when m > n; also
// build_diagonal_of_K (Element_calculations.c:624-652) into BI = 1/diag.
__global__ void __launch_bounds__(128) k_node_ks(Geo g, const float *__restrict__ GNX, const float *__restrict__ GDA,
                                                 const float *__restrict__ EVI, const float *__restrict__ EG,
                                                 const unsigned *__restrict__ NODE, int aug_on, double aug,
                                                 float *__restrict__ K, double *__restrict__ BI) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = colour_valid(g, q);
    if (col < 0) return;
    int k, i, j;
    colour_kij(g, col, q, k, i, j);
    const int n = nid(g, k, i, j);
    float *kq = K + ktile(q);
    double diag[3] = {0.0, 0.0, 0.0};
    for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++)
            for (int dz = -1; dz <= 1; dz++) {
                const int s = (dz + 1) + 3 * (dx + 1) + 9 * (dy + 1);
                const int km = k + dz, im = i + dx, jm = j + dy;
                float acc[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                const bool ok = (unsigned)km < (unsigned)g.noz && (unsigned)im < (unsigned)g.nox &&
                                (unsigned)jm < (unsigned)g.noy;
                if (ok) {
                    const int m = nid(g, km, im, jm);
                    const bool nbig = (m <= n);
                    const int big = nbig ? n : m, small = nbig ? m : n;
                    const unsigned fb = NODE[big], fs = NODE[small];
                    const double w[3] = {(fb & VBX) ? 0.0 : 1.0, (fb & VBY) ? 0.0 : 1.0, (fb & VBZ) ? 0.0 : 1.0};
                    const double ww[3] = {(fs & VBX) ? 0.0 : 1.0, (fs & VBY) ? 0.0 : 1.0, (fs & VBZ) ? 0.0 : 1.0};
                    const int kb = nbig ? k : km, ib = nbig ? i : im, jb = nbig ? j : jm;
                    const int ks = nbig ? km : k, is = nbig ? im : i, js = nbig ? jm : j;
                    // elements holding both nodes, ascending element number
                    for (int ey = max(j, jm) - 1; ey <= min(j, jm); ey++) {
                        if ((unsigned)ey >= (unsigned)g.ely) continue;
                        for (int ex = max(i, im) - 1; ex <= min(i, im); ex++) {
                            if ((unsigned)ex >= (unsigned)g.elx) continue;
                            for (int ez = max(k, km) - 1; ez <= min(k, km); ez++) {
                                if ((unsigned)ez >= (unsigned)g.elz) continue;
                                const int e = ez + g.elz * (ex + g.elx * ey);
                                const int ab = cLoc[(kb - ez) + 2 * (ib - ex) + 4 * (jb - ey)];
                                const int as = cLoc[(ks - ez) + 2 * (is - ex) + 4 * (js - ey)];
                                double blk[3][3];
                                elt_block(GNX + (size_t)e * GNX_SZ, GDA + (size_t)e * GDA_SZ, EVI + (size_t)e * 8,
                                          EG + (size_t)e * EG_SZ, aug_on, aug, ab, as, blk);
                                if (s == 13) {
                                    diag[0] += blk[0][0];
                                    diag[1] += blk[1][1];
                                    diag[2] += blk[2][2];
                                }
                                // reference order: column direction outer, row inner
#pragma unroll
                                for (int c = 0; c < 3; c++)
#pragma unroll
                                    for (int r = 0; r < 3; r++)
                                        acc[r][c] = (float)(acc[r][c] + w[r] * ww[c] * blk[r][c]);
                            }
                        }
                    }
                    // m > n: the lower node's stored block, transposed.  m == n:
                    // the reference's matvec and smoother read a node's own block
                    // through the column loop (Element_calculations.c:604-605,
                    // General_matrix_functions.c:1254-1255), i.e. transposed too
                    if (!nbig || s == 13) {
                        float t[3][3];
#pragma unroll
                        for (int r = 0; r < 3; r++)
#pragma unroll
                            for (int c = 0; c < 3; c++) t[r][c] = acc[c][r];
#pragma unroll
                        for (int r = 0; r < 3; r++)
#pragma unroll
                            for (int c = 0; c < 3; c++) acc[r][c] = t[r][c];
                    }
                }
#pragma unroll
                for (int r = 0; r < 3; r++)
#pragma unroll
                    for (int c = 0; c < 3; c++) kq[(s * 9 + r * 3 + c) * 64] = acc[r][c];
            }
    BI[3 * n] = (float)1.0 / diag[0];
    BI[3 * n + 1] = (float)1.0 / diag[1];
    BI[3 * n + 2] = (float)1.0 / diag[2];
}

void node_ks(hipStream_t s, const Level &L, const chip_problem &p) {
    hipLaunchKernelGGL(k_node_ks, dim3(nblk(L.g.pitch, 128)), dim3(128), 0, s, L.g, L.GNX, L.GDA, L.EVI, L.EG, L.NODE,
                       p.aug_lagr, p.aug_number, L.K, L.BI);
}

// assemble_dAhatp_entry / build_diagonal_of_Ahat (Element_calculations.c:654-685, 771-824)
__global__ void k_bpi(Geo g, const float *__restrict__ EG, const double *__restrict__ BI, int precond,
                      double *__restrict__ BPI) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    if (!precond) {
        BPI[e] = 1.0;
        return;
    }
    const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
    const float *eg = EG + (size_t)e * EG_SZ;
    double divU = 0.0;
#pragma unroll
    for (int a = 1; a <= 8; a++) {
        const int n = nid(g, ez + cOffz[a], ex + cOffx[a], ey + cOffy[a]);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double gp = BI[3 * n + d] * eg[3 * (a - 1) + d];
            divU += eg[3 * (a - 1) + d] * gp;
        }
    }
    BPI[e] = (divU != 0.0) ? 1.0 / divU : 1.0;
}

void build_bpi(hipStream_t s, const Level &L, int precond) {
    hipLaunchKernelGGL(k_bpi, dim3(nblk(L.g.nel)), dim3(TPB), 0, s, L.g, L.EG, L.BI, precond, L.BPI);
}

// ---------------------------------------------------------------------------
// energy equation

// order-preserving float <-> unsigned maps for atomic min / max
__device__ __forceinline__ unsigned ford(float f) {
    const unsigned b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// std_timestep (Advection_diffusion.c:737-810): per-element advective limit
// min into minmax[0]; Tmax (Global_operations.c:394-406) into minmax[1]
__global__ void __launch_bounds__(TPB) k_timestep_tmax(Geo g, const float *__restrict__ ECO, const float *__restrict__ V,
                                                       const float *__restrict__ T, unsigned *minmax) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned mn = ford(1.0e8f), mx = ford(-10.0f);
    if (V && t < g.nel) {
        const int e = t;
        const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
        float uc1 = 0.0f, uc2 = 0.0f, uc3 = 0.0f;
#pragma unroll
        for (int a = 1; a <= 8; a++) {
            const int nd = nid(g, ez + cOffz[a], ex + cOffx[a], ey + cOffy[a]);
            uc1 = (float)(uc1 + cNp[a - 1] * V[nd]);
            uc2 = (float)(uc2 + cNp[a - 1] * V[g.nno + nd]);
            uc3 = (float)(uc3 + cNp[a - 1] * V[2 * g.nno + nd]);
        }
        const float *eco = ECO + (size_t)e * ECO_SZ;
        const float uc = (float)(fabs((double)uc1) / eco[5] + fabs((double)uc2) / eco[6] + fabs((double)uc3) / eco[7]);
        const float step = (float)(0.5 / uc);
        mn = ford(step);
    }
    if (t < g.nno) mx = ford(T[t]);
    // wave reductions, then one atomic per wave
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mn = min(mn, (unsigned)__shfl_down((int)mn, o, 64));
        mx = max(mx, (unsigned)__shfl_down((int)mx, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&minmax[0], mn);
        atomicMax(&minmax[1], mx);
    }
}

__global__ void k_minmax_init(unsigned *minmax) {
    minmax[0] = ford(1.0e8f);
    minmax[1] = ford(-10.0f);
}

void timestep_tmax(hipStream_t s, chip_ctx *c, const float *V, const float *T) {
    const Level &L = c->L[c->levmax];
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(1), 0, s, c->minmax);
    hipLaunchKernelGGL(k_timestep_tmax, dim3(nblk(std::max(L.g.nel, L.g.nno))), dim3(TPB), 0, s, L.g, L.ECO, V, T,
                       c->minmax);
}

void tmax(hipStream_t s, chip_ctx *c, const float *T) { timestep_tmax(s, c, nullptr, T); }

// predictor (Advection_diffusion.c:352-366)
__global__ void k_predictor(Geo g, const unsigned *__restrict__ NODE, float *T, float *Tdot, float mult) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    if (!(NODE[n] & (TBX | TBZ | TBY))) T[n] += mult * Tdot[n];
    Tdot[n] = 0.0f;
}

void predictor(hipStream_t s, const Level &L, float *T, float *Tdot, float mult) {
    hipLaunchKernelGGL(k_predictor, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, L.NODE, T, Tdot, mult);
}

// pg_shape_fn + element_residual (Advection_diffusion.c:448-727, CART3D,
// diffusion on, no FBZ flux nodes): Eres[e][a] for the 8 nodes of element e
__global__ void __launch_bounds__(TPB) k_pg_residual(Geo g, const float *__restrict__ GNX, const float *__restrict__ GDA,
                                                     const float *__restrict__ ECO, const unsigned *__restrict__ NODE,
                                                     const float *__restrict__ diffusivity, double Q,
                                                     const float *__restrict__ T, const float *__restrict__ Tdot,
                                                     const float *__restrict__ V, double *__restrict__ Eres) {
    const int el = blockIdx.x * blockDim.x + threadIdx.x;
    if (el >= g.nel) return;
    const int ez = el % g.elz, ex = (el / g.elz) % g.elx, ey = el / (g.elz * g.elx);
    const float diff = (float)((diffusivity[ez] + diffusivity[ez + 1]) * 0.5);
    const float *gnx = GNX + (size_t)el * GNX_SZ;
    const float *gda = GDA + (size_t)el * GDA_SZ;
    const float *eco = ECO + (size_t)el * ECO_SZ;
    int nodes[9];
    float v1n[9], v2n[9], v3n[9];
#pragma unroll
    for (int a = 1; a <= 8; a++) {
        nodes[a] = nid(g, ez + cOffz[a], ex + cOffx[a], ey + cOffy[a]);
        v1n[a] = V[nodes[a]];
        v2n[a] = V[g.nno + nodes[a]];
        v3n[a] = V[2 * g.nno + nodes[a]];
    }
    // pg_shape_fn
    double uc1 = 0.0, uc2 = 0.0, uc3 = 0.0;
#pragma unroll
    for (int a = 1; a <= 8; a++) {
        uc1 += cNp[a - 1] * v1n[a];
        uc2 += cNp[a - 1] * v2n[a];
        uc3 += cNp[a - 1] * v3n[a];
    }
    const double twodiff = 2.0 * diff;
    const double uxse = fabs(uc1 * (double)eco[5]), ueta = fabs(uc2 * (double)eco[6]), ufai = fabs(uc3 * (double)eco[7]);
    const double xse = (uxse > twodiff) ? (1.0 - twodiff / uxse) : 0.0;
    const double eta = (ueta > twodiff) ? (1.0 - twodiff / ueta) : 0.0;
    const double fai = (ufai > twodiff) ? (1.0 - twodiff / ufai) : 0.0;
    const double unorm = uc1 * uc1 + uc2 * uc2 + uc3 * uc3;
    const double adiff = (unorm > 0.000001) ? ((uxse * xse + ueta * eta + ufai * fai) / (2.0 * unorm)) : 0.0;
    // element_residual: gauss-point values
    double dT[9], tx1[9], tx2[9], tx3[9], v1[9], v2[9], v3[9];
#pragma unroll
    for (int i = 1; i <= 8; i++) dT[i] = tx1[i] = tx2[i] = tx3[i] = v1[i] = v2[i] = v3[i] = 0.0;
#pragma unroll
    for (int j = 1; j <= 8; j++) {
        const int node = nodes[j];
        const double Tn = T[node];
        const double DT = (NODE[node] & (TBX | TBY | TBZ)) ? 0.0 : (double)Tdot[node];
#pragma unroll
        for (int i = 1; i <= 8; i++) {
            const double sfn = cNv[8 * (j - 1) + i - 1];
            dT[i] += DT * sfn;
            tx1[i] += gnx[8 * (j - 1) + i - 1] * Tn;
            tx2[i] += gnx[64 + 8 * (j - 1) + i - 1] * Tn;
            tx3[i] += gnx[128 + 8 * (j - 1) + i - 1] * Tn;
            v1[i] += v1n[j] * sfn;
            v2[i] += v2n[j] * sfn;
            v3[i] += v3n[j] * sfn;
        }
    }
    double u1g[9], u2g[9], u3g[9];
#pragma unroll
    for (int i = 1; i <= 8; i++) {
        double a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
        for (int j = 1; j <= 8; j++) {
            const double sfn = cNv[8 * (j - 1) + i - 1];
            a1 += v1n[j] * sfn;
            a2 += v2n[j] * sfn;
            a3 += v3n[j] * sfn;
        }
        u1g[i] = a1;
        u2g[i] = a2;
        u3g[i] = a3;
    }
#pragma unroll
    for (int j = 1; j <= 8; j++) {
        double r = 0.0;
#pragma unroll
        for (int i = 1; i <= 8; i++) {
            const float gx = gnx[8 * (j - 1) + i - 1], gy = gnx[64 + 8 * (j - 1) + i - 1], gz = gnx[128 + 8 * (j - 1) + i - 1];
            const double prod1 = (u1g[i] * gx + u2g[i] * gy + u3g[i] * gz);
            const double PG = cNv[8 * (j - 1) + i - 1] + adiff * prod1;
            const float dl = diff * 1.0f * gda[i];
            r -= PG * gda[i] * (dT[i] - Q + v1[i] * tx1[i] + v2[i] * tx2[i] + v3[i] * tx3[i]) +
                 dl * (gx * tx1[i] + gy * tx2[i] + gz * tx3[i]);
        }
        Eres[(size_t)el * 8 + j - 1] = r;
    }
}

void pg_residual(hipStream_t s, const chip_ctx *c, const float *T, const float *Tdot, const float *V, double *Eres) {
    const Level &L = c->L[c->levmax];
    // heating_adi = heating_visc = 0 and heating_latent = 1 (Instructions.c:376-381)
    const double Q = (c->p.Q0 - 0.0f + 0.0f) * 1.0f;
    hipLaunchKernelGGL(k_pg_residual, dim3(nblk(L.g.nel)), dim3(TPB), 0, s, L.g, L.GNX, L.GDA, L.ECO, L.NODE,
                       c->diffusivity, Q, T, Tdot, V, Eres);
}

// pg_solver's scatter and lumped-mass scaling (Advection_diffusion.c:423-437)
// fused with corrector (:369-387)
__global__ void __launch_bounds__(TPB) k_pg_gather(Geo g, const unsigned *__restrict__ NODE,
                                                   const float *__restrict__ MASS, const double *__restrict__ Eres,
                                                   float *T, float *Tdot, float mult) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    int k, i, j;
    node_kij(g, n, k, i, j);
    float dt = 0.0f;
    for (int ey = j - 1; ey <= j; ey++) {
        if ((unsigned)ey >= (unsigned)g.ely) continue;
        for (int ex = i - 1; ex <= i; ex++) {
            if ((unsigned)ex >= (unsigned)g.elx) continue;
            for (int ez = k - 1; ez <= k; ez++) {
                if ((unsigned)ez >= (unsigned)g.elz) continue;
                const int e = ez + g.elz * (ex + g.elx * ey);
                const int a = cLoc[(k - ez) + 2 * (i - ex) + 4 * (j - ey)];
                dt = (float)(dt + Eres[(size_t)e * 8 + a - 1]);
            }
        }
    }
    dt *= MASS[n];
    if (!(NODE[n] & (TBX | TBZ | TBY))) T[n] += mult * dt;
    Tdot[n] += dt;
}

void pg_gather_correct(hipStream_t s, const chip_ctx *c, const double *Eres, float *T, float *Tdot, float mult) {
    const Level &L = c->L[c->levmax];
    hipLaunchKernelGGL(k_pg_gather, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, L.NODE, L.MASS, Eres, T, Tdot, mult);
}

// temperatures_conform_bcs (Boundary_conditions.c:949-990); the side walls
// carry zero-flux flags, so only TBZ nodes are set
__global__ void k_conform_T(Geo g, const unsigned *__restrict__ NODE, const float *__restrict__ TB, float *T) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    if (NODE[n] & TBZ) T[n] = TB[n];
}

void conform_temperature(hipStream_t s, const chip_ctx *c, float *T) {
    const Level &L = c->L[c->levmax];
    hipLaunchKernelGGL(k_conform_T, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, L.NODE, c->TB, T);
}

// thermal_buoyancy (Pan_problem_misc_functions.c:83-153) with remove_horiz_ave
// / return_horiz_ave (Global_operations.c:55-80, 133-251)
__global__ void k_buoy_raw(Geo g, float Ra, const float *__restrict__ expansivity, const float *__restrict__ T,
                           float *__restrict__ b) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    b[n] = Ra * T[n] * expansivity[n % g.noz];
}

// horizontal average of a nodal field on node layer kz = blockIdx.x, with
// the bilinear surface quadrature of get_global_1d_shape_fn
__global__ void __launch_bounds__(TPB) k_horiz_ave(Geo g, const float *__restrict__ xc, const float *__restrict__ yc,
                                                   const float *__restrict__ X, float *__restrict__ H) {
    __shared__ double sh[4 * 16];
    const int kz = blockIdx.x;
    double v[2] = {0.0, 0.0};
    for (int f = threadIdx.x; f < g.elx * g.ely; f += blockDim.x) {
        const int ex = f / g.ely, ey = f % g.ely;
        const int n1 = nid(g, kz, ex, ey), n2 = nid(g, kz, ex + 1, ey), n3 = nid(g, kz, ex + 1, ey + 1),
                  n4 = nid(g, kz, ex, ey + 1);
        const double xx1[5] = {0, xc[ex], xc[ex + 1], xc[ex + 1], xc[ex]};
        const double xx2[5] = {0, yc[ey], yc[ey], yc[ey + 1], yc[ey + 1]};
        double dG[5];
#pragma unroll
        for (int nint = 1; nint <= 4; nint++) {
            double d11 = 0, d12 = 0, d21 = 0, d22 = 0;
#pragma unroll
            for (int ii = 1; ii <= 4; ii++) {
                const double mx0 = cMxv[4 * (ii - 1) + nint - 1], mx1 = cMxv[16 + 4 * (ii - 1) + nint - 1];
                d11 += xx1[ii] * mx0;
                d12 += xx2[ii] * mx0;
                d21 += xx1[ii] * mx1;
                d22 += xx2[ii] * mx1;
            }
            dG[nint] = d11 * d22 - d12 * d21;
        }
        const float xv[5] = {0, X[n1], X[n2], X[n3], X[n4]};
#pragma unroll
        for (int d = 1; d <= 4; d++)
#pragma unroll
            for (int nint = 1; nint <= 4; nint++) {
                v[0] += xv[d] * cMv[4 * (d - 1) + nint - 1] * dG[nint];
                v[1] += cMv[4 * (d - 1) + nint - 1] * dG[nint];
            }
    }
    block_sum<2>(v, sh);
    if (threadIdx.x == 0 && v[1] != 0.0) H[kz] = (float)(v[0] / v[1]);
}

// averages (Process_velocity.c:179-224): per-layer horizontal averages of
// T, of |V|^2 (square-rooted on the host) and of the nodal viscosity
// (visc_from_gint_to_nodes of the finest EVI); |V|^2 in FP32 as the
// reference forms it
__global__ void k_vsq(long nno, const float *__restrict__ V, float *__restrict__ out) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nno) return;
    const float a = V[n], b = V[nno + n], c = V[2 * nno + n];
    out[n] = a * a + b * b + c * c;
}

void averages(hipStream_t s, chip_ctx *c, float *Have /* [3 noz] */, float *scratch /* [nno] */) {
    Level &L = c->L[c->levmax];
    hipLaunchKernelGGL(k_horiz_ave, dim3(L.g.noz), dim3(TPB), 0, s, L.g, c->xc_d, c->yc_d, c->T, Have);
    hipLaunchKernelGGL(k_vsq, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, (long)L.g.nno, c->V, scratch);
    hipLaunchKernelGGL(k_horiz_ave, dim3(L.g.noz), dim3(TPB), 0, s, L.g, c->xc_d, c->yc_d, scratch, Have + L.g.noz);
    hipLaunchKernelGGL(k_visc_to_nodes, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, L.EVI, L.TWW, L.MASS, scratch);
    hipLaunchKernelGGL(k_horiz_ave, dim3(L.g.noz), dim3(TPB), 0, s, L.g, c->xc_d, c->yc_d, scratch,
                       Have + 2 * L.g.noz);
}

__global__ void k_sub_layer(Geo g, const float *__restrict__ H, float *b) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    b[n] -= H[n % g.noz];
}

void thermal_buoyancy(hipStream_t s, const chip_ctx *c, const float *T, float *buoy) {
    const Level &L = c->L[c->levmax];
    hipLaunchKernelGGL(k_buoy_raw, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, c->p.rayleigh, c->expansivity, T, buoy);
    hipLaunchKernelGGL(k_horiz_ave, dim3(L.g.noz), dim3(TPB), 0, s, L.g, c->xc_d, c->yc_d, buoy, c->H);
    hipLaunchKernelGGL(k_sub_layer, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, c->H, buoy);
}

// assemble_forces + get_elt_f (Element_calculations.c:74-125, 989-1121,
// CART3D, no imposed velocities): element z-forces, then a per-node gather
__global__ void __launch_bounds__(TPB) k_elt_f(Geo g, const float *__restrict__ GDA, const float *__restrict__ buoy,
                                               double *__restrict__ Ef) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
    const float *gda = GDA + (size_t)e * GDA_SZ;
    double force[9], fg[9];
#pragma unroll
    for (int a = 1; a <= 8; a++) force[a] = buoy[nid(g, ez + cOffz[a], ex + cOffx[a], ey + cOffy[a])];
#pragma unroll
    for (int j = 1; j <= 8; j++) {
        double t = 0.0;
#pragma unroll
        for (int k = 1; k <= 8; k++) t += force[k] * cNv[8 * (k - 1) + j - 1];
        fg[j] = t;
    }
#pragma unroll
    for (int a = 1; a <= 8; a++) {
        double f = 0.0;
#pragma unroll
        for (int j = 1; j <= 8; j++) f += fg[j] * cNv[8 * (a - 1) + j - 1] * gda[j] * 1.0f;
        Ef[(size_t)e * 8 + a - 1] = f;
    }
}

__global__ void __launch_bounds__(TPB) k_forces_gather(Geo g, const unsigned *__restrict__ NODE,
                                                       const double *__restrict__ Ef, double *__restrict__ F) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    int k, i, j;
    node_kij(g, n, k, i, j);
    double acc = 0.0;
    for (int ey = j - 1; ey <= j; ey++) {
        if ((unsigned)ey >= (unsigned)g.ely) continue;
        for (int ex = i - 1; ex <= i; ex++) {
            if ((unsigned)ex >= (unsigned)g.elx) continue;
            for (int ez = k - 1; ez <= k; ez++) {
                if ((unsigned)ez >= (unsigned)g.elz) continue;
                const int e = ez + g.elz * (ex + g.elx * ey);
                acc += Ef[(size_t)e * 8 + cLoc[(k - ez) + 2 * (i - ex) + 4 * (j - ey)] - 1];
            }
        }
    }
    const unsigned f = NODE[n];
    F[3 * n] = 0.0;
    F[3 * n + 1] = 0.0;
    F[3 * n + 2] = (f & VBZ) ? 0.0 : acc;
}

void assemble_forces(hipStream_t s, const chip_ctx *c, const float *buoy, double *F) {
    const Level &L = c->L[c->levmax];
    hipLaunchKernelGGL(k_elt_f, dim3(nblk(L.g.nel)), dim3(TPB), 0, s, L.g, L.GDA, buoy, c->Eres);
    hipLaunchKernelGGL(k_forces_gather, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, L.NODE, c->Eres, F);
}

// heat_flux (Process_buoyancy.c:63-203)
__global__ void __launch_bounds__(TPB) k_hf_elem(Geo g, const float *__restrict__ GNX, const float *__restrict__ GDA,
                                                 const float *__restrict__ ECO, const float *__restrict__ diffusivity,
                                                 const float *__restrict__ T, const float *__restrict__ V,
                                                 double *__restrict__ uTe) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
    const double diff = (diffusivity[ez] + diffusivity[ez + 1]) * 0.5;
    const float *gnx = GNX + (size_t)e * GNX_SZ;
    const float *gda = GDA + (size_t)e * GDA_SZ;
    double VZ[9];
    float Tn[9];
#pragma unroll
    for (int a = 1; a <= 8; a++) {
        const int n = nid(g, ez + cOffz[a], ex + cOffx[a], ey + cOffy[a]);
        VZ[a] = V[2 * g.nno + n];
        Tn[a] = T[n];
    }
    double uT = 0.0;
#pragma unroll
    for (int i = 1; i <= 8; i++) {
        double u = 0.0, Tg = 0.0, dTdz = 0.0;
#pragma unroll
        for (int j = 1; j <= 8; j++) {
            u += VZ[j] * cNv[8 * (j - 1) + i - 1];
            Tg += Tn[j] * cNv[8 * (j - 1) + i - 1];
            dTdz = dTdz + Tn[j] * gnx[128 + 8 * (j - 1) + i - 1];
        }
        uT = uT + (u * Tg - diff * dTdz) * gda[i];
    }
    uTe[e] = uT / ECO[(size_t)e * ECO_SZ + 12];
}

// Volume-RMS velocity (no counterpart in the reference): per element the Q1
// quadrature of v.v and of 1 with the reference's shape functions and
// Jacobian weights (Shape_function_dA), in FP64.
__global__ void __launch_bounds__(TPB) k_vrms_elem(Geo g, const float *__restrict__ GDA, const float *__restrict__ V,
                                                   double *__restrict__ num, double *__restrict__ den,
                                                   double *__restrict__ ones) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= g.nel) return;
    const int ez = e % g.elz, ex = (e / g.elz) % g.elx, ey = e / (g.elz * g.elx);
    const float *gda = GDA + (size_t)e * GDA_SZ;
    double Vn[3][9];
#pragma unroll
    for (int a = 1; a <= 8; a++) {
        const int n = nid(g, ez + cOffz[a], ex + cOffx[a], ey + cOffy[a]);
#pragma unroll
        for (int d = 0; d < 3; d++) Vn[d][a] = V[d * g.nno + n];
    }
    double vv = 0.0, vol = 0.0;
#pragma unroll
    for (int i = 1; i <= 8; i++) {
        double s2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double u = 0.0;
#pragma unroll
            for (int j = 1; j <= 8; j++) u += Vn[d][j] * cNv[8 * (j - 1) + i - 1];
            s2 += u * u;
        }
        vv += s2 * gda[i];
        vol += gda[i];
    }
    num[e] = vv;
    den[e] = vol;
    ones[e] = 1.0;
}

void vrms(hipStream_t s, chip_ctx *c, const float *V, double *out2) {
    const Level &L = c->L[c->levmax];
    const long nel = L.g.nel;
    double *num = c->Eres, *den = c->Eres + nel, *one = c->Eres + 2 * nel;
    hipLaunchKernelGGL(k_vrms_elem, dim3(nblk(nel)), dim3(TPB), 0, s, L.g, L.GDA, V, num, den, one);
    dots(s, c, (int)nel, 2, (const double *)num, (const double *)one, (const double *)den, (const double *)one,
         nullptr, nullptr, nullptr, nullptr, out2);
}

__global__ void __launch_bounds__(TPB) k_hf_gather(Geo g, const float *__restrict__ TWW, const float *__restrict__ MASS,
                                                   const double *__restrict__ uTe, float *__restrict__ hf) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    int k, i, j;
    node_kij(g, n, k, i, j);
    float acc = 0.0f;
    for (int ey = j - 1; ey <= j; ey++) {
        if ((unsigned)ey >= (unsigned)g.ely) continue;
        for (int ex = i - 1; ex <= i; ex++) {
            if ((unsigned)ex >= (unsigned)g.elx) continue;
            for (int ez = k - 1; ez <= k; ez++) {
                if ((unsigned)ez >= (unsigned)g.elz) continue;
                const int e = ez + g.elz * (ex + g.elx * ey);
                const int a = cLoc[(k - ez) + 2 * (i - ex) + 4 * (j - ey)];
                acc = (float)(acc + TWW[(size_t)e * TWW_SZ + a] * uTe[e]);
            }
        }
    }
    hf[n] = acc * MASS[n];
}

// surface extrapolation and surface-element sums (Process_buoyancy.c:158-193)
__global__ void __launch_bounds__(TPB) k_hf_surface(Geo g, const float *__restrict__ ECO, const float *__restrict__ hf,
                                                    double *__restrict__ out) {
    __shared__ double sh[4 * 16];
    double v[4] = {0, 0, 0, 0};
    for (int f = threadIdx.x; f < g.elx * g.ely; f += blockDim.x) {
        const int ex = f % g.elx, ey = f / g.elx;
        double tb = 0.0, tt = 0.0;
        const int cols[4] = {ex + g.nox * ey, ex + 1 + g.nox * ey, ex + 1 + g.nox * (ey + 1), ex + g.nox * (ey + 1)};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int bot = cols[q] * g.noz, top = bot + g.noz - 1;
            const float sh_ = 2 * hf[top] - hf[top - 1];
            const float bh_ = 2 * hf[bot] - hf[bot + 1];
            tb += bh_;
            tt += sh_;
        }
        const int eb = g.elz * (ex + g.elx * ey), et = eb + g.elz - 1;
        const float ab = ECO[(size_t)eb * ECO_SZ + 12], at = ECO[(size_t)et * ECO_SZ + 12];
        v[0] += tb * ab;
        v[1] += ab;
        v[2] += tt * at;
        v[3] += at;
    }
    block_sum<4>(v, sh);
    if (threadIdx.x == 0)
        for (int m = 0; m < 4; m++) out[m] = v[m];
}

void heat_flux(hipStream_t s, chip_ctx *c, const float *T, const float *V) {
    const Level &L = c->L[c->levmax];
    hipLaunchKernelGGL(k_hf_elem, dim3(nblk(L.g.nel)), dim3(TPB), 0, s, L.g, L.GNX, L.GDA, L.ECO, c->diffusivity, T, V,
                       c->Eres);
    hipLaunchKernelGGL(k_hf_gather, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, L.TWW, L.MASS, c->Eres, c->hf);
    hipLaunchKernelGGL(k_hf_surface, dim3(1), dim3(TPB), 0, s, L.g, L.ECO, c->hf, c->dots + 24);
}

// ---------------------------------------------------------------------------
// Learned coarse-level correction (no reference counterpart).  The coarsest
// level's smoother call of multi_grid (General_matrix_functions.c:574, :613)
// is replaced by a small 3-D CNN (trained by the BF16 MFMA kernels of
// learn_kernels.hip): inputs F / rms(F) (three channels) and log10
// viscosity, output the correction / rms(F).  The inner solve still
// iterates to its own residual criterion.

// one workgroup of CNET_TPB threads runs the whole correction; its two
// activation buffers ([nv][16] bf16 each) take 64 nv bytes of LDS
constexpr int CNET_TPB = 512;
constexpr int CNET_MAX_NV = 1024;
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(TPB) k_cnet_logeta(Geo g, const float *__restrict__ EVI, float *__restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    const int k = n % g.noz, i = (n / g.noz) % g.nox, j = n / (g.noz * g.nox);
    float acc = 0.f;
    int cnt = 0;
    for (int dj = -1; dj <= 0; dj++)
        for (int di = -1; di <= 0; di++)
            for (int dk = -1; dk <= 0; dk++) {
                const int ez = k + dk, ex = i + di, ey = j + dj;
                if (ez < 0 || ex < 0 || ey < 0 || ez >= g.elz || ex >= g.elx || ey >= g.ely) continue;
                const size_t e = (size_t)ez + (size_t)g.elz * (ex + (size_t)g.elx * ey);
                float m = 0.f;
                for (int q = 0; q < 8; q++) m += log10f(EVI[e * 8 + q]);
                acc += 0.125f * m;
                cnt++;
            }
    out[n] = acc / (float)cnt;
}

void cnet_logeta(hipStream_t s, const Level &L, float *out) {
    hipLaunchKernelGGL(k_cnet_logeta, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, (const float *)L.EVI, out);
}

// round to the nearest bf16, ties to even (finite values)
__device__ __forceinline__ uint16_t bf16_rne(float f) {
    const unsigned u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}


__device__ __forceinline__ bf16x4_t cnet_load4(const uint16_t *p) {
    const uint2 r = *reinterpret_cast<const uint2 *>(p);
    bf16x4_t a;
    a[0] = (short)(r.x & 0xFFFF);
    a[1] = (short)(r.x >> 16);
    a[2] = (short)(r.y & 0xFFFF);
    a[3] = (short)(r.y >> 16);
    return a;
}

// The whole correction in one workgroup: rms(F) and the input (F / rms,
// log10 eta) into LDS, the layers with activations in LDS (the products,
// operand layout and rounding of k_conv_fwd in learn_kernels.hip: one
// 16x16x16 bf16 MFMA per tap, FP32 accumulation, bias, ReLU, bf16), the
// last layer's three output channels times out_scale times rms into d,
// BC rows zero.
template <typename V>
__global__ void __launch_bounds__(CNET_TPB) k_cnet(Geo g, const V *__restrict__ F, const float *__restrict__ logeta,
                                                   const uint16_t *__restrict__ W, const float *__restrict__ bias,
                                                   int layers, const float *__restrict__ ys,
                                                   const unsigned *__restrict__ NODE, V *__restrict__ d) {
    extern __shared__ __align__(16) uint16_t act[];   // [2][nv][16]
    __shared__ double sh[16];
    __shared__ double sc_sh[2];
    const int nv = g.nno, l = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double v[1] = {0.0};
    for (int i = threadIdx.x; i < 3 * nv; i += blockDim.x) v[0] += (double)F[i] * (double)F[i];
    block_sum<1>(v, sh);
    if (threadIdx.x == 0) {
        const double rms = sqrt(v[0] / (3.0 * nv));
        sc_sh[0] = rms;
        sc_sh[1] = rms > 0.0 ? 1.0 / rms : 0.0;
    }
    __syncthreads();
    const double rms = sc_sh[0], inv = sc_sh[1];
    for (int idx = threadIdx.x; idx < nv * 16; idx += blockDim.x) {
        const int n = idx >> 4, ch = idx & 15;
        float x = 0.f;
        if (ch < 3)
            x = (float)((double)F[3 * n + ch] * inv);
        else if (ch == 3)
            x = logeta[n];
        act[idx] = bf16_rne(x);
    }
    __syncthreads();
    const int o = l & 15, kq = 4 * (l >> 4);
    for (int li = 0; li < layers; li++) {
        const uint16_t *X = act + (size_t)(li & 1) * nv * 16;
        uint16_t *Y = act + (size_t)((li + 1) & 1) * nv * 16;
        const uint16_t *Wl = W + (size_t)li * 16 * 27 * 16;
        const float bo = bias[li * 16 + o];
        const bool last = li == layers - 1;
        for (int v0 = wv * 16; v0 < nv; v0 += nw * 16) {   // uniform per wavefront
            const int m = v0 + (l & 15);
            const bool in = m < nv;
            const int k = m % g.noz, i = (m / g.noz) % g.nox, j = m / (g.noz * g.nox);
            f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int t = 0; t < 27; t++) {   // tap t = (dz+1) + 3(dx+1) + 9(dy+1)
                const int dz = t % 3 - 1, dx = (t / 3) % 3 - 1, dy = t / 9 - 1;
                const bool ok = in && (unsigned)(k + dz) < (unsigned)g.noz && (unsigned)(i + dx) < (unsigned)g.nox &&
                                (unsigned)(j + dy) < (unsigned)g.noy;
                bf16x4_t a = {0, 0, 0, 0};
                if (ok) a = cnet_load4(X + (size_t)(m + dz + g.noz * (dx + g.nox * dy)) * 16 + kq);
                const bf16x4_t w = cnet_load4(Wl + ((size_t)o * 27 + t) * 16 + kq);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, acc, 0, 0, 0);
            }
            for (int i = 0; i < 4; i++) {
                const int vv = v0 + kq + i;
                if (vv >= nv) continue;
                float y = acc[i] + bo;
                if (!last) {
                    Y[(size_t)vv * 16 + o] = bf16_rne(y > 0.f ? y : 0.f);
                } else if (o < 3) {
                    V out = (V)((double)y * (double)ys[o] * rms);
                    const unsigned f = NODE[vv];
                    if ((o == 0 && (f & VBX)) || (o == 1 && (f & VBY)) || (o == 2 && (f & VBZ))) out = V(0);
                    d[3 * vv + o] = out;
                }
            }
        }
        __syncthreads();
    }
}

bool cnet_fits(const Geo &g) { return g.nno <= CNET_MAX_NV; }

template <typename V>
int cnet_apply(hipStream_t s, const chip_ctx *c, const Level &L, const V *F, V *d) {
    if (!cnet_fits(L.g)) return -1;
    auto kern = k_cnet<V>;
    hipLaunchKernelGGL(kern, dim3(1), dim3(CNET_TPB), (size_t)2 * L.g.nno * 16 * sizeof(uint16_t), s, L.g, F,
                       (const float *)c->cnet_logeta, (const uint16_t *)c->cnet_w, (const float *)c->cnet_b,
                       c->cnet_layers, (const float *)c->cnet_ys, L.NODE, d);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
template int cnet_apply<double>(hipStream_t, const chip_ctx *, const Level &, const double *, double *);
template int cnet_apply<float>(hipStream_t, const chip_ctx *, const Level &, const float *, float *);

}  // namespace chip

// ---------------------------------------------------------------------------
// the learned warm start's inputs and outputs on the finest level (no
// counterpart in the reference: the network's output is only the Uzawa
// solve's initial guess)

namespace chip {

__device__ __forceinline__ uint16_t warm_bf16(float f) {
    const unsigned u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// dT = T - (2 Th0 - Th1): the temperature's departure from its own linear
// extrapolation over the last two solves
__global__ void k_warm_dT(const float *__restrict__ T, const float *__restrict__ Th0, const float *__restrict__ Th1,
                          long n, float *__restrict__ dT) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dT[i] = T[i] - (2.0f * Th0[i] - Th1[i]);
}
void warm_dT(hipStream_t s, const float *T, const float *Th0, const float *Th1, long n, float *dT) {
    hipLaunchKernelGGL(k_warm_dT, dim3(nblk(n)), dim3(TPB), 0, s, T, Th0, Th1, n, dT);
}

// nodal value of an element field: the mean of (Pa - Pb) over the node's
// elements (Pb optional)
__global__ void k_e2n(Geo g, const double *__restrict__ Pa, const double *__restrict__ Pb, float *__restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    const int k = n % g.noz, i = (n / g.noz) % g.nox, j = n / (g.noz * g.nox);
    double acc = 0.0;
    int cnt = 0;
    for (int dj = -1; dj <= 0; dj++)
        for (int di = -1; di <= 0; di++)
            for (int dk = -1; dk <= 0; dk++) {
                const int ez = k + dk, ex = i + di, ey = j + dj;
                if (ez < 0 || ex < 0 || ey < 0 || ez >= g.elz || ex >= g.elx || ey >= g.ely) continue;
                const size_t e = (size_t)ez + (size_t)g.elz * (ex + (size_t)g.elx * ey);
                acc += Pb ? Pa[e] - Pb[e] : Pa[e];
                cnt++;
            }
    out[n] = (float)(acc / cnt);
}
void e2n(hipStream_t s, const Level &L, const double *Pa, const double *Pb, float *out) {
    hipLaunchKernelGGL(k_e2n, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, Pa, Pb, out);
}

// the network's input, 16 bf16 channels per node: T - 1/2, dT / s_dT,
// log10(eta) / 2.5, the extrapolated U / s_u (3) and nodal P / s_p, the
// last step's increments U - U_prev / s_u (3) and nodal P - P_prev / s_p,
// the second differences of U / s_2u (3) and of the nodal P / s_2p, z
// (node index over noz - 1).  U_prev here is the last solution.
__global__ void k_warm_features(Geo g, const float *__restrict__ T, const float *__restrict__ dT,
                                const float *__restrict__ logeta, const double *__restrict__ U,
                                const double *__restrict__ Uprev, const double *__restrict__ D2,
                                const float *__restrict__ pn, const float *__restrict__ dpn,
                                const float *__restrict__ d2pn, float inv_dT, float inv_u, float inv_p,
                                float inv_2u, float inv_2p, uint16_t *__restrict__ x) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    float f[16];
    f[0] = T[n] - 0.5f;
    f[1] = dT[n] * inv_dT;
    f[2] = logeta[n] * 0.4f;
    for (int d = 0; d < 3; d++) {
        const size_t i = 3 * (size_t)n + d;
        f[3 + d] = (float)U[i] * inv_u;
        f[7 + d] = (float)(U[i] - Uprev[i]) * inv_u;
        f[11 + d] = (float)D2[i] * inv_2u;
    }
    f[6] = pn[n] * inv_p;
    f[10] = dpn[n] * inv_p;
    f[14] = d2pn[n] * inv_2p;
    f[15] = (float)(n % g.noz) / (float)(g.noz - 1);
    uint4 *o = reinterpret_cast<uint4 *>(x + (size_t)n * 16);
    for (int q = 0; q < 2; q++) {
        uint4 w;
        w.x = warm_bf16(f[8 * q + 0]) | ((uint32_t)warm_bf16(f[8 * q + 1]) << 16);
        w.y = warm_bf16(f[8 * q + 2]) | ((uint32_t)warm_bf16(f[8 * q + 3]) << 16);
        w.z = warm_bf16(f[8 * q + 4]) | ((uint32_t)warm_bf16(f[8 * q + 5]) << 16);
        w.w = warm_bf16(f[8 * q + 6]) | ((uint32_t)warm_bf16(f[8 * q + 7]) << 16);
        o[q] = w;
    }
}
void warm_features(hipStream_t s, const Level &L, const float *T, const float *dT, const float *logeta,
                   const double *U, const double *Uprev, const double *D2, const float *pn, const float *dpn,
                   const float *d2pn, float inv_dT, float inv_u, float inv_p, float inv_2u, float inv_2p,
                   uint16_t *x) {
    hipLaunchKernelGGL(k_warm_features, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, T, dT, logeta, U, Uprev, D2, pn,
                       dpn, d2pn, inv_dT, inv_u, inv_p, inv_2u, inv_2p, x);
}

// the network's correction: U += s_eu out[n][0..2] at every node; P[e] +=
// s_ep (mean of out[.][3] over the element's eight nodes)
__global__ void k_warm_apply(Geo g, const float *__restrict__ out, double s_eu, double s_ep, double *__restrict__ U,
                             double *__restrict__ P) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3L * g.nno) U[i] += s_eu * (double)out[(i / 3) * 16 + i % 3];
    if (i < g.nel) {
        const int ez = (int)(i % g.elz), ex = (int)((i / g.elz) % g.elx), ey = (int)(i / ((long)g.elz * g.elx));
        double m = 0.0;
        for (int q = 0; q < 8; q++) {
            const int k = ez + (q & 1), ii = ex + ((q >> 1) & 1), j = ey + (q >> 2);
            m += (double)out[((size_t)k + (size_t)g.noz * (ii + (size_t)g.nox * j)) * 16 + 3];
        }
        P[i] += s_ep * 0.125 * m;
    }
}
void warm_apply(hipStream_t s, const Level &L, const float *out, double s_eu, double s_ep, double *U, double *P) {
    hipLaunchKernelGGL(k_warm_apply, dim3(nblk(3L * L.g.nno)), dim3(TPB), 0, s, L.g, out, s_eu, s_ep, U, P);
}

// the training target of one solve: (U - Ue) / s_eu per node (3 channels)
// and the nodal mean of (P - Pe) / s_ep, Ue, Pe the extrapolated guess
__global__ void k_warm_target(Geo g, const double *__restrict__ U, const double *__restrict__ Ue,
                              const float *__restrict__ dpn, double inv_eu, double inv_ep, float *__restrict__ y) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nno) return;
    for (int d = 0; d < 3; d++) y[4 * (size_t)n + d] = (float)((U[3 * (size_t)n + d] - Ue[3 * (size_t)n + d]) * inv_eu);
    y[4 * (size_t)n + 3] = (float)(dpn[n] * inv_ep);
}
void warm_target(hipStream_t s, const Level &L, const double *U, const double *Ue, const float *dpn, double inv_eu,
                 double inv_ep, float *y) {
    hipLaunchKernelGGL(k_warm_target, dim3(nblk(L.g.nno)), dim3(TPB), 0, s, L.g, U, Ue, dpn, inv_eu, inv_ep, y);
}

}  // namespace chip
