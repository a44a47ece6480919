// BF16 MFMA kernels of the learned warm start's U-Net: 3x3x3 convolutions
// with any channel counts that are multiples of 16 (forward, data gradient
// through transposed weights, weight gradient), the grid transfers between
// node-coincident levels (full-weighting restriction, trilinear
// interpolation, each the other's adjoint up to a factor 8), a BF16 sum and
// a weighted mean-squared-error loss.  No counterpart in the reference.
//
// Layouts: activations X[b][v][C] bf16 (channels fastest, C a multiple of
// 16), v = k + nz*(i + nx*j); weights W[o][t][c] bf16 (FP32 masters kept by
// the caller); tap t = (dz+1) + 3(dx+1) + 9(dy+1).  Operand layout of one
// v_mfma_f32_16x16x16_bf16 on a wavefront: lane l holds A[l%16][4(l/16)+i],
// B[4(l/16)+i][l%16] and D[4(l/16)+i][l%16], i = 0..3.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/citcom_learn.h"
#include "learn_common.h"

using namespace learn;

namespace {

constexpr int TAPS = 27;

// ---- forward: Y[v][o] = act(sum_t sum_c X[v+off_t][c] W[o][t][c] + bias[o] + R[v][o])
// One wavefront computes 16 voxels x (NBO*16) output channels; the K loop
// runs over the taps and the 16-channel blocks of the input.  mirror reads
// tap t at the mirrored offset -off_t (the data gradient, with transposed
// weights).
template <int NBO>
__global__ void __launch_bounds__(256) k_conv_fwd_c(Vol g, int ci, int co, const uint16_t *__restrict__ X,
                                                    const uint16_t *__restrict__ W, const float *__restrict__ bias,
                                                    const uint16_t *__restrict__ R, int relu, int mirror,
                                                    uint16_t *__restrict__ Y, float *__restrict__ Yf) {
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int groups = co / (16 * NBO);
    const int b = blockIdx.y / groups, og = blockIdx.y % groups;
    const int v0 = (blockIdx.x * 4 + wv) * 16;
    if (v0 >= g.nv) return;
    const uint16_t *Xb = X + (size_t)b * g.nv * ci;
    const int m = v0 + (l & 15), kq = 4 * (l >> 4);
    const Cell cm = cell(g, m);
    const int ncb = ci >> 4;
    f32x4 acc[NBO];
#pragma unroll
    for (int ob = 0; ob < NBO; ob++) acc[ob] = f32x4{0.f, 0.f, 0.f, 0.f};
    // weight row of this lane's output channel (column l%16 of B) per block
    const uint16_t *Wl = W + (size_t)(og * NBO * 16 + (l & 15)) * TAPS * ci + kq;
    const size_t wstride = (size_t)16 * TAPS * ci;   // next block of 16 output channels
#pragma unroll 3
    for (int t = 0; t < TAPS; t++) {
        const int vn = nbr(g, cm, mirror ? 26 - t : t);
        const uint16_t *xa = Xb + (size_t)(vn < 0 ? 0 : vn) * ci + kq;
        for (int cb = 0; cb < ncb; cb++) {
            bf16x4 a = {0, 0, 0, 0};
            if (vn >= 0) a = load4(xa + cb * 16);
#pragma unroll
            for (int ob = 0; ob < NBO; ob++) {
                const bf16x4 w = load4(Wl + ob * wstride + (size_t)t * ci + cb * 16);
                acc[ob] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, acc[ob], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int ob = 0; ob < NBO; ob++) {
        const int o = (og * NBO + ob) * 16 + (l & 15);
        const float bo = bias ? bias[o] : 0.f;
        for (int i = 0; i < 4; i++) {
            const int v = v0 + kq + i;
            if (v >= g.nv) continue;
            const size_t at = ((size_t)b * g.nv + v) * co + o;
            float y = acc[ob][i] + bo;
            if (R) y += bf2f(R[at]);
            if (relu) y = y > 0.f ? y : 0.f;
            Y[at] = f2bf(y);
            if (Yf) Yf[at] = y;
        }
    }
}

// WT[c][t][o] = W[o][t][c]
__global__ void k_transpose_w(const uint16_t *__restrict__ W, int ci, int co, uint16_t *__restrict__ WT) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)co * TAPS * ci) return;
    const int c = (int)(idx % ci), t = (int)((idx / ci) % TAPS), o = (int)(idx / ((long)ci * TAPS));
    WT[((size_t)c * TAPS + t) * co + o] = W[idx];
}

// ---- weight gradient: dW[o][t][c] = sum_b sum_v dY[b][v][o] X[b][v+off_t][c],
// db[o] = sum dY[b][v][o].  One wavefront per (voxel chunk, 16 output
// channels, 16 input channels): 27 tap tiles, K = 16 voxels per MFMA, its
// partial sums in its own slot; a second kernel sums the slots in chunk
// order (a fixed order: the result does not depend on scheduling).
constexpr int WCH = 512;   // voxels per chunk
__global__ void __launch_bounds__(256) k_conv_bwd_weight_c(Vol g, int nb, int ci, int co,
                                                           const uint16_t *__restrict__ X,
                                                           const uint16_t *__restrict__ dY, float *__restrict__ part) {
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int per = (g.nv + WCH - 1) / WCH;
    const int nob = co >> 4, ncb = ci >> 4;
    const long w = (long)blockIdx.x * 4 + wv;   // over nb * per * nob * ncb
    if (w >= (long)nb * per * nob * ncb) return;
    const int cb = (int)(w % ncb), ob = (int)((w / ncb) % nob);
    const long chunk = w / ((long)ncb * nob);
    const int b = (int)(chunk / per), vbeg = (int)(chunk % per) * WCH;
    const int vend = min(g.nv, vbeg + WCH);
    const size_t slot = (size_t)co * TAPS * ci + co;
    float *P = part + chunk * slot;
    const uint16_t *Xb = X + (size_t)b * g.nv * ci, *dYb = dY + (size_t)b * g.nv * co;
    const int r = l & 15, kq = 4 * (l >> 4);
    const int o = ob * 16 + r, c = cb * 16 + r;
    float dbias = 0.f;
    for (int t = 0; t < TAPS; t++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int v0 = vbeg; v0 < vend; v0 += 16) {
            bf16x4 a, bb;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int v = v0 + kq + i;
                uint16_t av = 0, bv = 0;
                if (v < vend) {
                    av = dYb[(size_t)v * co + o];
                    const int vn = nbr(g, cell(g, v), t);
                    if (vn >= 0) bv = Xb[(size_t)vn * ci + c];
                    if (t == 0) dbias += bf2f(av);
                }
                a[i] = (short)av;
                bb[i] = (short)bv;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
        }
        // D[4(l/16)+i][l%16] = dW[ob*16 + kq+i][t][cb*16 + l%16]
#pragma unroll
        for (int i = 0; i < 4; i++) P[((size_t)(ob * 16 + kq + i) * TAPS + t) * ci + c] = acc[i];
    }
    if (cb == 0) {   // lanes sharing r hold disjoint voxel quarters
        dbias += __shfl_xor(dbias, 16, 64);
        dbias += __shfl_xor(dbias, 32, 64);
        if (l < 16) P[(size_t)co * TAPS * ci + o] = dbias;
    }
}

__global__ void k_sum_slots(const float *__restrict__ part, long nslots, long slot, float *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= slot) return;
    float s = 0.f;
    for (long k = 0; k < nslots; k++) s += part[k * slot + i];
    out[i] = s;
}

// ---- grid transfers.  Coarse node c of a level lies on fine node 2c per
// axis (fine n = 2 (coarse n - 1) + 1).  Restriction (full weighting):
// coarse[c] = scale * sum_{d in {-1,0,1}^3} w(d) fine[2c+d], w = prod of
// 1/2 (d = 0) and 1/4 (d = +-1), neighbours outside the grid left out.
// Interpolation (trilinear): fine[f] = scale * sum_c p(f - 2c) coarse[c],
// p = prod of 1 (d = 0) and 1/2 (d = +-1).  The restriction's transpose is
// the interpolation / 8 and the interpolation's transpose 8 x the restriction.
__global__ void k_restrict(Vol fg, Vol cg, int ch, float scale, const uint16_t *__restrict__ F,
                           uint16_t *__restrict__ Cc, int nb) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nb * cg.nv * ch) return;
    const int c = (int)(idx % ch);
    const long bv = idx / ch;
    const int b = (int)(bv / cg.nv), v = (int)(bv % cg.nv);
    const int k = v % cg.nz, i = (v / cg.nz) % cg.nx, j = v / (cg.nz * cg.nx);
    const uint16_t *Fb = F + (size_t)b * fg.nv * ch;
    float s = 0.f;
    for (int dy = -1; dy <= 1; dy++) {
        const int jj = 2 * j + dy;
        if ((unsigned)jj >= (unsigned)fg.ny) continue;
        for (int dx = -1; dx <= 1; dx++) {
            const int ii = 2 * i + dx;
            if ((unsigned)ii >= (unsigned)fg.nx) continue;
            for (int dz = -1; dz <= 1; dz++) {
                const int kk = 2 * k + dz;
                if ((unsigned)kk >= (unsigned)fg.nz) continue;
                const float w = (dy ? 0.25f : 0.5f) * (dx ? 0.25f : 0.5f) * (dz ? 0.25f : 0.5f);
                s += w * bf2f(Fb[((size_t)kk + fg.nz * ((size_t)ii + fg.nx * jj)) * ch + c]);
            }
        }
    }
    Cc[idx] = f2bf(scale * s);
}

__global__ void k_interp(Vol cg, Vol fg, int ch, float scale, const uint16_t *__restrict__ Cc,
                         uint16_t *__restrict__ F, int nb) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nb * fg.nv * ch) return;
    const int c = (int)(idx % ch);
    const long bv = idx / ch;
    const int b = (int)(bv / fg.nv), v = (int)(bv % fg.nv);
    const int k = v % fg.nz, i = (v / fg.nz) % fg.nx, j = v / (fg.nz * fg.nx);
    const uint16_t *Cb = Cc + (size_t)b * cg.nv * ch;
    // per axis: even fine index -> one coarse node (weight 1), odd -> the two
    // coarse nodes either side (1/2 each)
    const int j0 = j >> 1, j1 = (j + 1) >> 1, i0 = i >> 1, i1 = (i + 1) >> 1, k0 = k >> 1, k1 = (k + 1) >> 1;
    const float wy = (j & 1) ? 0.5f : 1.f, wx = (i & 1) ? 0.5f : 1.f, wz = (k & 1) ? 0.5f : 1.f;
    float s = 0.f;
    for (int a = 0; a < ((j & 1) ? 2 : 1); a++) {
        const int jj = a ? j1 : j0;
        for (int e = 0; e < ((i & 1) ? 2 : 1); e++) {
            const int ii = e ? i1 : i0;
            for (int h = 0; h < ((k & 1) ? 2 : 1); h++) {
                const int kk = h ? k1 : k0;
                s += (wy * wx * wz) * bf2f(Cb[((size_t)kk + cg.nz * ((size_t)ii + cg.nx * jj)) * ch + c]);
            }
        }
    }
    F[idx] = f2bf(scale * s);
}

__global__ void k_add_bf16(uint16_t *__restrict__ a, const uint16_t *__restrict__ b, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

// weighted mean squared error over the first nout of ystride channels:
// loss = sum_v sum_c wch[c] (Y[v][c] - T[v][c])^2 / (nvox nout); dY (bf16,
// ystride channels, zero beyond nout) = d loss / d Y; part[block] = block sums
__global__ void k_wmse(const float *__restrict__ Y, int ystride, const float *__restrict__ T, long nvox, int nout,
                       const float *__restrict__ wch, uint16_t *__restrict__ dY, float *__restrict__ part) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.f;
    if (i < nvox * ystride) {
        const int c = (int)(i % ystride);
        const long v = i / ystride;
        float gr = 0.f;
        if (c < nout) {
            const float e = Y[i] - T[v * nout + c];
            e2 = wch[c] * e * e;
            gr = 2.f * wch[c] * e / (float)(nvox * nout);
        }
        dY[i] = f2bf(gr);
    }
    for (int w = 32; w >= 1; w >>= 1) e2 += __shfl_xor(e2, w, 64);
    __shared__ float sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

hipStream_t st(void *s) { return static_cast<hipStream_t>(s); }
int ret() { return hipGetLastError() == hipSuccess ? 0 : -1; }
unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }
bool chan_ok(int c) { return c >= 16 && c <= 256 && c % 16 == 0; }

// ---- the U-Net of the warm start (4 levels, 16/32/48/64 channels, 15
// convolutions; additive skip connections; 594,432 weights + 528 biases).
// Level l has (n - 1) / 2^l + 1 nodes per axis.  Encoder: two convolutions
// per level (ReLU), the second's output kept as the level's skip, then a
// full-weighting restriction to the next level; bottom: two convolutions;
// decoder: trilinear interpolation to the finer level, a convolution plus
// that level's skip (ReLU), a second convolution (ReLU); a last 16 -> 16
// convolution without activation gives the output (FP32, channels 0..3
// used).
struct UConv {
    int ci, co;
};
constexpr UConv UNET[15] = {{16, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 48}, {48, 48}, {48, 64}, {64, 64},
                            {64, 48}, {48, 48}, {48, 32}, {32, 32}, {32, 16}, {16, 16}, {16, 16}};
constexpr int UNET_LEVELS = 4;
// channels of the forward pass's buffers per level: A, B, S (skip)
constexpr int UA[UNET_LEVELS] = {32, 48, 64, 64}, UB[UNET_LEVELS] = {16, 32, 48, 64}, US[UNET_LEVELS] = {16, 32, 48, 0};

long unet_level_nv(int nx, int ny, int nz, int lev) {
    const int s = 1 << lev;
    return (long)((nx - 1) / s + 1) * ((ny - 1) / s + 1) * ((nz - 1) / s + 1);
}

}  // namespace

extern "C" {

int cl_conv3d_fwd_c(const uint16_t *x, int nb, int nx, int ny, int nz, int ci, int co, const uint16_t *w,
                    const float *bias, const uint16_t *res, int relu, int mirror, uint16_t *y, float *yf,
                    void *stream) {
    const Vol g{nx, ny, nz, nx * ny * nz};
    if (nb <= 0 || g.nv <= 0 || !chan_ok(ci) || !chan_ok(co)) return -1;
    const int nbo = (co % 64 == 0) ? 4 : (co % 48 == 0) ? 3 : (co % 32 == 0) ? 2 : 1;
    const dim3 grid((g.nv + 63) / 64, nb * (co / (16 * nbo)));
    switch (nbo) {
    case 4: hipLaunchKernelGGL(k_conv_fwd_c<4>, grid, dim3(256), 0, st(stream), g, ci, co, x, w, bias, res, relu, mirror, y, yf); break;
    case 3: hipLaunchKernelGGL(k_conv_fwd_c<3>, grid, dim3(256), 0, st(stream), g, ci, co, x, w, bias, res, relu, mirror, y, yf); break;
    case 2: hipLaunchKernelGGL(k_conv_fwd_c<2>, grid, dim3(256), 0, st(stream), g, ci, co, x, w, bias, res, relu, mirror, y, yf); break;
    default: hipLaunchKernelGGL(k_conv_fwd_c<1>, grid, dim3(256), 0, st(stream), g, ci, co, x, w, bias, res, relu, mirror, y, yf); break;
    }
    return ret();
}

int cl_conv3d_transpose_w(const uint16_t *w, int ci, int co, uint16_t *wt, void *stream) {
    if (!chan_ok(ci) || !chan_ok(co)) return -1;
    const long n = (long)co * TAPS * ci;
    hipLaunchKernelGGL(k_transpose_w, dim3(nblk(n)), dim3(256), 0, st(stream), w, ci, co, wt);
    return ret();
}

int cl_conv3d_bwd_data_c(const uint16_t *dy, int nb, int nx, int ny, int nz, int ci, int co, const uint16_t *wt,
                         uint16_t *dx, void *stream) {
    // dX[v][c] = sum_t sum_o dY[v - off_t][o] W[o][t][c]: the forward
    // convolution of dY (co channels in, ci out) with WT[c][t][o] and the
    // mirrored taps
    return cl_conv3d_fwd_c(dy, nb, nx, ny, nz, co, ci, wt, nullptr, nullptr, 0, 1, dx, nullptr, stream);
}

long cl_conv3d_bwd_weight_c_workspace(int nb, int nx, int ny, int nz, int ci, int co) {
    const long nv = (long)nx * ny * nz, chunks = (long)nb * ((nv + WCH - 1) / WCH);
    return chunks * ((long)co * TAPS * ci + co) * (long)sizeof(float);
}

int cl_conv3d_bwd_weight_c(const uint16_t *x, const uint16_t *dy, int nb, int nx, int ny, int nz, int ci, int co,
                           float *grad, float *workspace, void *stream) {
    const Vol g{nx, ny, nz, nx * ny * nz};
    if (nb <= 0 || g.nv <= 0 || !chan_ok(ci) || !chan_ok(co)) return -1;
    const long chunks = (long)nb * ((g.nv + WCH - 1) / WCH);
    const long waves = chunks * (co / 16) * (ci / 16);
    const long slot = (long)co * TAPS * ci + co;
    hipLaunchKernelGGL(k_conv_bwd_weight_c, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st(stream), g, nb, ci,
                       co, x, dy, workspace);
    hipLaunchKernelGGL(k_sum_slots, dim3(nblk(slot)), dim3(256), 0, st(stream), workspace, chunks, slot, grad);
    return ret();
}

int cl_restrict(const uint16_t *fine, int nb, int nx, int ny, int nz, int ch, float scale, uint16_t *coarse,
                void *stream) {
    if (nb <= 0 || nx < 3 || ny < 3 || nz < 3 || !(nx & ny & nz & 1) || ch <= 0) return -1;
    const Vol fg{nx, ny, nz, nx * ny * nz};
    const Vol cg{(nx - 1) / 2 + 1, (ny - 1) / 2 + 1, (nz - 1) / 2 + 1, ((nx - 1) / 2 + 1) * ((ny - 1) / 2 + 1) * ((nz - 1) / 2 + 1)};
    const long n = (long)nb * cg.nv * ch;
    hipLaunchKernelGGL(k_restrict, dim3(nblk(n)), dim3(256), 0, st(stream), fg, cg, ch, scale, fine, coarse, nb);
    return ret();
}

int cl_interp(const uint16_t *coarse, int nb, int nx, int ny, int nz, int ch, float scale, uint16_t *fine,
              void *stream) {
    if (nb <= 0 || nx < 3 || ny < 3 || nz < 3 || !(nx & ny & nz & 1) || ch <= 0) return -1;
    const Vol fg{nx, ny, nz, nx * ny * nz};
    const Vol cg{(nx - 1) / 2 + 1, (ny - 1) / 2 + 1, (nz - 1) / 2 + 1, ((nx - 1) / 2 + 1) * ((ny - 1) / 2 + 1) * ((nz - 1) / 2 + 1)};
    const long n = (long)nb * fg.nv * ch;
    hipLaunchKernelGGL(k_interp, dim3(nblk(n)), dim3(256), 0, st(stream), cg, fg, ch, scale, coarse, fine, nb);
    return ret();
}

int cl_add_bf16(uint16_t *a, const uint16_t *b, long n, void *stream) {
    hipLaunchKernelGGL(k_add_bf16, dim3(nblk(n)), dim3(256), 0, st(stream), a, b, n);
    return ret();
}

int cl_wmse(const float *y, int ystride, const float *target, long nvox, int nout, const float *wch, uint16_t *dy,
            float *partial, void *stream) {
    if (nout <= 0 || nout > ystride) return -1;
    hipLaunchKernelGGL(k_wmse, dim3(nblk(nvox * ystride)), dim3(256), 0, st(stream), y, ystride, target, nvox, nout, wch,
                       dy, partial);
    return ret();
}

}  // extern "C"

extern "C" {

int cl_unet_params(long *n_weights, long *n_biases) {
    long nw = 0, nb = 0;
    for (const UConv &c : UNET) {
        nw += (long)c.co * TAPS * c.ci;
        nb += c.co;
    }
    if (n_weights) *n_weights = nw;
    if (n_biases) *n_biases = nb;
    return 0;
}

long cl_unet_workspace(int nb, int nx, int ny, int nz) {
    long bytes = 0;
    for (int l = 0; l < UNET_LEVELS; l++) {
        const long nv = (long)nb * unet_level_nv(nx, ny, nz, l);
        // bf16 buffers A, B, S of the level, 256-byte aligned
        for (int ch : {UA[l], UB[l], US[l]}) bytes += (nv * ch * 2 + 255) / 256 * 256;
    }
    return bytes;
}

int cl_unet_forward(const uint16_t *x, int nb, int nx, int ny, int nz, const uint16_t *w, const float *bias,
                    float *out, void *workspace, void *stream) {
    for (int n : {nx, ny, nz})
        if (n < 17 || (n - 1) % 8) return -1;   // four levels of at least 3 nodes per axis
    if (nb <= 0 || !x || !w || !bias || !out || !workspace) return -1;
    // carve the buffers in the order cl_unet_workspace counts them
    uint16_t *A[UNET_LEVELS], *B[UNET_LEVELS], *Sk[UNET_LEVELS];
    char *p = static_cast<char *>(workspace);
    for (int l = 0; l < UNET_LEVELS; l++) {
        const long nv = (long)nb * unet_level_nv(nx, ny, nz, l);
        uint16_t **slots[3] = {&A[l], &B[l], &Sk[l]};
        const int chs[3] = {UA[l], UB[l], US[l]};
        for (int q = 0; q < 3; q++) {
            *slots[q] = reinterpret_cast<uint16_t *>(p);
            p += (nv * chs[q] * 2 + 255) / 256 * 256;
        }
    }
    int d[UNET_LEVELS][3];
    for (int l = 0; l < UNET_LEVELS; l++) {
        const int s = 1 << l;
        d[l][0] = (nx - 1) / s + 1;
        d[l][1] = (ny - 1) / s + 1;
        d[l][2] = (nz - 1) / s + 1;
    }
    const uint16_t *wl[15];
    const float *bl[15];
    {
        long ow = 0, ob = 0;
        for (int i = 0; i < 15; i++) {
            wl[i] = w + ow;
            bl[i] = bias + ob;
            ow += (long)UNET[i].co * TAPS * UNET[i].ci;
            ob += UNET[i].co;
        }
    }
    int li = 0;
    auto conv = [&](int lev, const uint16_t *in, const uint16_t *res, int relu, uint16_t *y, float *yf) {
        const UConv &c = UNET[li];
        const int r = cl_conv3d_fwd_c(in, nb, d[lev][0], d[lev][1], d[lev][2], c.ci, c.co, wl[li], bl[li], res, relu, 0,
                                      y, yf, stream);
        li++;
        return r;
    };
    int rc = 0;
    // encoder
    rc |= conv(0, x, nullptr, 1, B[0], nullptr);
    rc |= conv(0, B[0], nullptr, 1, Sk[0], nullptr);
    rc |= cl_restrict(Sk[0], nb, d[0][0], d[0][1], d[0][2], 16, 1.f, B[1], stream);
    rc |= conv(1, B[1], nullptr, 1, A[1], nullptr);
    rc |= conv(1, A[1], nullptr, 1, Sk[1], nullptr);
    rc |= cl_restrict(Sk[1], nb, d[1][0], d[1][1], d[1][2], 32, 1.f, B[2], stream);
    rc |= conv(2, B[2], nullptr, 1, A[2], nullptr);
    rc |= conv(2, A[2], nullptr, 1, Sk[2], nullptr);
    rc |= cl_restrict(Sk[2], nb, d[2][0], d[2][1], d[2][2], 48, 1.f, B[3], stream);
    // bottom
    rc |= conv(3, B[3], nullptr, 1, A[3], nullptr);
    rc |= conv(3, A[3], nullptr, 1, B[3], nullptr);
    // decoder
    rc |= cl_interp(B[3], nb, d[2][0], d[2][1], d[2][2], 64, 1.f, A[2], stream);
    rc |= conv(2, A[2], Sk[2], 1, B[2], nullptr);
    rc |= conv(2, B[2], nullptr, 1, A[2], nullptr);
    rc |= cl_interp(A[2], nb, d[1][0], d[1][1], d[1][2], 48, 1.f, A[1], stream);
    rc |= conv(1, A[1], Sk[1], 1, B[1], nullptr);
    rc |= conv(1, B[1], nullptr, 1, A[1], nullptr);
    rc |= cl_interp(A[1], nb, d[0][0], d[0][1], d[0][2], 32, 1.f, A[0], stream);
    rc |= conv(0, A[0], Sk[0], 1, B[0], nullptr);
    rc |= conv(0, B[0], nullptr, 1, A[0], nullptr);
    rc |= conv(0, A[0], nullptr, 0, B[0], out);
    return rc ? -1 : 0;
}

}  // extern "C"
