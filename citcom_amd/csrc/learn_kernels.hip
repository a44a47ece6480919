// BF16 MFMA kernels of the learned warm start: a 3-D convolutional network
// (3x3x3 kernels, 16 channels, zero padding) trained on the GPU by these
// forward and backward kernels.  No counterpart in the reference.
//
// Layouts: activations X[b][v][16] bf16, v = k + nz*(i + nx*j) (z fastest,
// the solver's node order); weights W[o][t][c] bf16 with FP32 master copies,
// tap t = (dz+1) + 3(dx+1) + 9(dy+1).  Every product runs on
// v_mfma_f32_16x16x16_bf16 with FP32 accumulation.  Operand layout of one
// 16x16x16 MFMA on a wavefront (checked on the device): lane l holds
// A[l%16][4(l/16)+i], B[4(l/16)+i][l%16] and D[4(l/16)+i][l%16], i = 0..3.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/citcom_learn.h"

namespace {

constexpr int CH = 16;   // channels (padded)
constexpr int TAPS = 27;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (finite values)
__device__ __forceinline__ uint16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct Vol {
    int nx, ny, nz, nv;
};

__device__ __forceinline__ int nbr(const Vol &g, int v, int t) {
    const int k = v % g.nz, i = (v / g.nz) % g.nx, j = v / (g.nz * g.nx);
    const int dz = t % 3 - 1, dx = (t / 3) % 3 - 1, dy = t / 9 - 1;
    const int kk = k + dz, ii = i + dx, jj = j + dy;
    if ((unsigned)kk >= (unsigned)g.nz || (unsigned)ii >= (unsigned)g.nx || (unsigned)jj >= (unsigned)g.ny) return -1;
    return kk + g.nz * (ii + g.nx * jj);
}

// the same with v's (k, i, j) formed once: in a loop over the taps, which
// is then unrolled, no division is left per tap
struct Cell {
    int v, k, i, j;
    bool in;
};
__device__ __forceinline__ Cell cell(const Vol &g, int v) {
    return {v, v % g.nz, (v / g.nz) % g.nx, v / (g.nz * g.nx), v < g.nv};
}
__device__ __forceinline__ int nbr(const Vol &g, const Cell &c, int t) {
    const int dz = t % 3 - 1, dx = (t / 3) % 3 - 1, dy = t / 9 - 1;
    const bool ok = c.in && (unsigned)(c.k + dz) < (unsigned)g.nz && (unsigned)(c.i + dx) < (unsigned)g.nx &&
                    (unsigned)(c.j + dy) < (unsigned)g.ny;
    return ok ? c.v + dz + g.nz * (dx + g.nx * dy) : -1;
}

__device__ __forceinline__ bf16x4 load4(const uint16_t *p) {
    const uint2 r = *reinterpret_cast<const uint2 *>(p);
    bf16x4 a;
    a[0] = (short)(r.x & 0xFFFF);
    a[1] = (short)(r.x >> 16);
    a[2] = (short)(r.y & 0xFFFF);
    a[3] = (short)(r.y >> 16);
    return a;
}

// ---- forward: Y[v][o] = act(sum_t sum_c X[v+off_t][c] W[o][t][c] + bias[o])
// one wavefront per 16 voxels; the 27 taps are the K loop (16 channels each)
__global__ void __launch_bounds__(256) k_conv_fwd(Vol g, const uint16_t *__restrict__ X, const uint16_t *__restrict__ W,
                                                  const float *__restrict__ bias, int relu, uint16_t *__restrict__ Y,
                                                  float *__restrict__ Yf) {
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, b = blockIdx.y;
    const int v0 = (blockIdx.x * 4 + wv) * 16;
    if (v0 >= g.nv) return;
    const uint16_t *Xb = X + (size_t)b * g.nv * CH;
    const int m = v0 + (l & 15), kq = 4 * (l >> 4);
    const Cell cm = cell(g, m);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < TAPS; t++) {
        const int vn = nbr(g, cm, t);
        bf16x4 a = {0, 0, 0, 0};
        if (vn >= 0) a = load4(Xb + (size_t)vn * CH + kq);
        const bf16x4 w = load4(W + ((size_t)(l & 15) * TAPS + t) * CH + kq);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, acc, 0, 0, 0);
    }
    const int o = l & 15;
    for (int i = 0; i < 4; i++) {
        const int v = v0 + kq + i;
        if (v >= g.nv) continue;
        float y = acc[i] + bias[o];
        if (relu) y = y > 0.f ? y : 0.f;
        Y[((size_t)b * g.nv + v) * CH + o] = f2bf(y);
        if (Yf) Yf[((size_t)b * g.nv + v) * CH + o] = y;
    }
}

// ---- backward through the activation: dYp = dY * (Y > 0)
__global__ void k_relu_bwd(const uint16_t *__restrict__ Y, uint16_t *__restrict__ dY, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(bf2f(Y[i]) > 0.f)) dY[i] = 0;
}

// ---- backward data: dX[v][c] = sum_t sum_o dY[v-off_t][o] W[o][t][c]
__global__ void __launch_bounds__(256) k_conv_bwd_data(Vol g, const uint16_t *__restrict__ dY,
                                                       const uint16_t *__restrict__ W, uint16_t *__restrict__ dX) {
    __shared__ __align__(16) uint16_t Wt[TAPS][CH][CH];   // [t][c][o]
    for (int idx = threadIdx.x; idx < TAPS * CH * CH; idx += blockDim.x) {
        const int o = idx % CH, c = (idx / CH) % CH, t = idx / (CH * CH);
        Wt[t][c][o] = W[((size_t)o * TAPS + t) * CH + c];
    }
    __syncthreads();
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, b = blockIdx.y;
    const int v0 = (blockIdx.x * 4 + wv) * 16;
    if (v0 >= g.nv) return;
    const uint16_t *dYb = dY + (size_t)b * g.nv * CH;
    const int m = v0 + (l & 15), kq = 4 * (l >> 4);
    const Cell cm = cell(g, m);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < TAPS; t++) {
        // the output voxel u with u + off_t = m is the neighbour of m at the
        // mirrored tap 26 - t
        const int u = nbr(g, cm, 26 - t);
        bf16x4 a = {0, 0, 0, 0};
        if (u >= 0) a = load4(dYb + (size_t)u * CH + kq);
        const bf16x4 w = load4(&Wt[t][l & 15][kq]);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, acc, 0, 0, 0);
    }
    const int c = l & 15;
    for (int i = 0; i < 4; i++) {
        const int v = v0 + kq + i;
        if (v < g.nv) dX[((size_t)b * g.nv + v) * CH + c] = f2bf(acc[i]);
    }
}

// ---- backward weights: dW[o][t][c] = sum_v dY[v][o] X[v+off_t][c], per
// wavefront over its voxel chunk (K = 16 voxels per MFMA), partial sums per
// wavefront; db[o] partials likewise
constexpr int WCHUNK = 256;   // voxels per wavefront
__global__ void __launch_bounds__(256) k_conv_bwd_weight(Vol g, int nb, const uint16_t *__restrict__ X,
                                                         const uint16_t *__restrict__ dY, float *__restrict__ part) {
    const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long chunk = (long)blockIdx.x * 4 + wv;   // over nb * ceil(nv / WCHUNK)
    const int per = (g.nv + WCHUNK - 1) / WCHUNK;
    float *P = part + (size_t)chunk * (TAPS * CH * CH + CH);
    if (chunk >= (long)nb * per) return;
    const int b = (int)(chunk / per), vbeg = (int)(chunk % per) * WCHUNK;
    const int vend = min(g.nv, vbeg + WCHUNK);
    const uint16_t *Xb = X + (size_t)b * g.nv * CH, *dYb = dY + (size_t)b * g.nv * CH;
    const int o = l & 15, kq = 4 * (l >> 4);
    float dbias = 0.f;
    for (int t = 0; t < TAPS; t++) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int v0 = vbeg; v0 < vend; v0 += 16) {
            bf16x4 a, bb;
            for (int i = 0; i < 4; i++) {
                const int v = v0 + kq + i;
                uint16_t av = 0, bv = 0;
                if (v < vend) {
                    av = dYb[(size_t)v * CH + o];
                    const int vn = nbr(g, v, t);
                    if (vn >= 0) bv = Xb[(size_t)vn * CH + o];
                    if (t == 0) dbias += bf2f(av);
                }
                a[i] = (short)av;
                bb[i] = (short)bv;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
        }
        // D[4(l/16)+i][l%16] = dW[o' = kq+i][t][c = l%16]
        for (int i = 0; i < 4; i++) P[((size_t)(kq + i) * TAPS + t) * CH + (l & 15)] = acc[i];
    }
    // bias: lanes sharing o = l%16 hold disjoint voxel quarters
    dbias += __shfl_xor(dbias, 16, 64);
    dbias += __shfl_xor(dbias, 32, 64);
    if (l < 16) P[TAPS * CH * CH + l] = dbias;
}

// dW (and db) = sum of the partials in chunk order
__global__ void k_sum_partials(const float *__restrict__ part, long nchunks, float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= TAPS * CH * CH + CH) return;
    float s = 0.f;
    for (long k = 0; k < nchunks; k++) s += part[(size_t)k * (TAPS * CH * CH + CH) + i];
    out[i] = s;
}

// mean squared error over the first nout channels; dY = d loss / d Y in bf16
__global__ void k_mse(const float *__restrict__ Y, const float *__restrict__ T, long nvox, int nout,
                      uint16_t *__restrict__ dY, float *__restrict__ part) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.f;
    if (i < nvox * CH) {
        const int c = (int)(i % CH);
        float g = 0.f;
        if (c < nout) {
            const float e = Y[i] - T[i];
            e2 = e * e;
            g = 2.f * e / (float)(nvox * nout);
        }
        dY[i] = f2bf(g);
    }
    for (int w = 32; w >= 1; w >>= 1) e2 += __shfl_xor(e2, w, 64);
    __shared__ float sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Adam on the FP32 master weights, then the BF16 copy
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ gr, float *__restrict__ m, float *__restrict__ v,
                       long n, float lr, float b1, float b2, float eps, float c1, float c2, uint16_t *__restrict__ wbf,
                       const unsigned char *__restrict__ mask) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mask && !mask[i]) return;
    const float g = gr[i];
    m[i] = b1 * m[i] + (1.f - b1) * g;
    v[i] = b2 * v[i] + (1.f - b2) * g * g;
    const float mh = m[i] / c1, vh = v[i] / c2;
    w[i] -= lr * mh / (sqrtf(vh) + eps);
    if (wbf) wbf[i] = f2bf(w[i]);
}

__global__ void k_to_bf16(const float *__restrict__ a, uint16_t *__restrict__ b, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

hipStream_t st(void *s) { return static_cast<hipStream_t>(s); }
int ret() { return hipGetLastError() == hipSuccess ? 0 : -1; }

// ---- the weight gradient's partial sums without a workspace argument: a
// file-scope device array, one copy per GPU, held until the process ends.
// Host code never passes `scratch` itself to a kernel (its host-side name is
// not a device address): the device address comes from hipGetSymbolAddress,
// once per device with that device current, and is kept in scratch_ptr.
constexpr long SCRATCH_FLOATS = 4L << 20;   // 16 MiB
__device__ __align__(256) float scratch[SCRATCH_FLOATS];
constexpr int MAX_DEVICES = 64;
float *scratch_ptr[MAX_DEVICES];
std::once_flag scratch_once[MAX_DEVICES];
hipError_t scratch_err[MAX_DEVICES];

// the device address of this device's scratch array (nullptr on failure,
// with the message in *why)
float *scratch_on_current_device(const char **why) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        *why = "no current HIP device (or more devices than the library tracks)";
        return nullptr;
    }
    std::call_once(scratch_once[dev], [dev] {
        void *p = nullptr;
        scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        scratch_ptr[dev] = scratch_err[dev] == hipSuccess ? static_cast<float *>(p) : nullptr;
    });
    if (!scratch_ptr[dev]) *why = "hipGetSymbolAddress of the scratch array failed";
    return scratch_ptr[dev];
}

thread_local std::string g_learn_err;

}  // namespace

extern "C" {

int cl_conv3d_fwd(const uint16_t *x, int nb, int nx, int ny, int nz, const uint16_t *w, const float *bias, int relu,
                  uint16_t *y, float *yf, void *stream) {
    const Vol g{nx, ny, nz, nx * ny * nz};
    if (nb <= 0 || g.nv <= 0) return -1;
    hipLaunchKernelGGL(k_conv_fwd, dim3((g.nv + 63) / 64, nb), dim3(256), 0, st(stream), g, x, w, bias, relu, y, yf);
    return ret();
}

int cl_relu_bwd(const uint16_t *y, uint16_t *dy, long n, void *stream) {
    hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st(stream), y, dy, n);
    return ret();
}

int cl_conv3d_bwd_data(const uint16_t *dy, int nb, int nx, int ny, int nz, const uint16_t *w, uint16_t *dx,
                       void *stream) {
    const Vol g{nx, ny, nz, nx * ny * nz};
    hipLaunchKernelGGL(k_conv_bwd_data, dim3((g.nv + 63) / 64, nb), dim3(256), 0, st(stream), g, dy, w, dx);
    return ret();
}

long cl_conv3d_bwd_weight_workspace(int nb, int nx, int ny, int nz) {
    const long nv = (long)nx * ny * nz, chunks = nb * ((nv + WCHUNK - 1) / WCHUNK);
    return ((chunks + 3) / 4 * 4) * (TAPS * CH * CH + CH) * (long)sizeof(float);
}

int cl_conv3d_bwd_weight(const uint16_t *x, const uint16_t *dy, int nb, int nx, int ny, int nz, float *grad,
                         float *workspace, void *stream) {
    const Vol g{nx, ny, nz, nx * ny * nz};
    const long chunks = (long)nb * ((g.nv + WCHUNK - 1) / WCHUNK);
    hipLaunchKernelGGL(k_conv_bwd_weight, dim3((unsigned)((chunks + 3) / 4)), dim3(256), 0, st(stream), g, nb, x, dy,
                       workspace);
    hipLaunchKernelGGL(k_sum_partials, dim3((TAPS * CH * CH + CH + 255) / 256), dim3(256), 0, st(stream), workspace,
                       chunks, grad);
    return ret();
}

long cl_scratch_bytes(void) { return SCRATCH_FLOATS * (long)sizeof(float); }

const char *cl_last_error(void) { return g_learn_err.c_str(); }

int cl_conv3d_bwd_weight_noalloc(const uint16_t *x, const uint16_t *dy, int nb, int nx, int ny, int nz, float *grad,
                                 void *stream) {
    const Vol g{nx, ny, nz, nx * ny * nz};
    if (nb <= 0 || nx <= 0 || ny <= 0 || nz <= 0) {
        g_learn_err = "cl_conv3d_bwd_weight_noalloc: empty volume";
        return -1;
    }
    const long need = cl_conv3d_bwd_weight_workspace(nb, nx, ny, nz);
    if (need > cl_scratch_bytes()) {
        g_learn_err = "cl_conv3d_bwd_weight_noalloc: " + std::to_string(nb) + " volumes of " + std::to_string(g.nv) +
                      " voxels need " + std::to_string(need) + " bytes of partial sums, more than the " +
                      std::to_string(cl_scratch_bytes()) + "-byte scratch array";
        return -1;
    }
    const char *why = "";
    float *part = scratch_on_current_device(&why);
    if (!part) {
        g_learn_err = std::string("cl_conv3d_bwd_weight_noalloc: ") + why;
        return -1;
    }
    // every chunk's slot of partial sums is written in full by
    // k_conv_bwd_weight before k_sum_partials reads it, so nothing is cleared
    const long chunks = (long)nb * ((g.nv + WCHUNK - 1) / WCHUNK);
    hipLaunchKernelGGL(k_conv_bwd_weight, dim3((unsigned)((chunks + 3) / 4)), dim3(256), 0, st(stream), g, nb, x, dy,
                       part);
    hipLaunchKernelGGL(k_sum_partials, dim3((TAPS * CH * CH + CH + 255) / 256), dim3(256), 0, st(stream), part, chunks,
                       grad);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_learn_err = std::string("cl_conv3d_bwd_weight_noalloc: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

int cl_mse(const float *y, const float *target, long nvox, int nout, uint16_t *dy, float *partial, void *stream) {
    hipLaunchKernelGGL(k_mse, dim3((unsigned)((nvox * CH + 255) / 256)), dim3(256), 0, st(stream), y, target, nvox, nout,
                       dy, partial);
    return ret();
}

int cl_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2, float eps, int step,
            uint16_t *wbf, const unsigned char *mask, void *stream) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st(stream), w, g, m, v, n, lr, b1, b2,
                       eps, c1, c2, wbf, mask);
    return ret();
}

int cl_to_bf16(const float *a, uint16_t *b, long n, void *stream) {
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st(stream), a, b, n);
    return ret();
}

}  // extern "C"
