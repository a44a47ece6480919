// Shared device helpers of the learned parts' kernels: BF16 conversion
// (round to nearest even), the 16x16x16 BF16 MFMA operand types, and the
// 3x3x3 neighbour arithmetic on node grids numbered v = k + nz*(i + nx*j)
// (z fastest, the solver's node order); tap t = (dz+1) + 3(dx+1) + 9(dy+1).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace learn {

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

// a voxel's grid position, formed once (the tap loops are unrolled and keep
// no division per tap)
struct Cell {
    int v, k, i, j;
    bool in;
};
__device__ __forceinline__ Cell cell(const Vol &g, int v) {
    return {v, v % g.nz, (v / g.nz) % g.nx, v / (g.nz * g.nx), v < g.nv};
}
// neighbour of c at tap t, -1 outside the grid (zero padding)
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

}  // namespace learn
