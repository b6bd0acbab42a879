// Launch-shape and block-reduction helpers shared by the kernel files.
#pragma once
#include <hip/hip_runtime.h>
#include "skt_internal.h"

namespace skt {

#define KL __global__ __launch_bounds__(256)

__device__ __forceinline__ int tid256() { return threadIdx.x + 64 * threadIdx.y; }

template <class T>
__device__ __forceinline__ T waveSum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <class T>
__device__ __forceinline__ T waveMax(T v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
// Sum over the 256 threads of a block (fixed order); result valid in thread 0.
__device__ __forceinline__ double blockSum(double v, double* sh) {
    v = waveSum(v);
    const int t = tid256();
    if ((t & 63) == 0) sh[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}
__device__ __forceinline__ double blockMax(double v, double* sh) {
    v = waveMax(v);
    const int t = tid256();
    if ((t & 63) == 0) sh[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0) r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    __syncthreads();
    return r;
}

// One thread per cell: 64 (x) x 4 (y) tiles, one z plane per block row.
#define CELL_SETUP                                   \
    const int x = blockIdx.x * 64 + threadIdx.x;     \
    const int y = blockIdx.y * 4 + threadIdx.y;      \
    const int z = blockIdx.z;                        \
    const bool inb = (x < G.X) && (y < G.Y);         \
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)

// One thread per cell from a tile index: 64 (x) x 4 (y) tiles, x tiles
// fastest, then y tiles, then z planes (the flattened three-dimensional grid
// of the per-cell launches).
#define CELL_SETUP_V                                                   \
    const unsigned gx_ = (unsigned)(G.X + 63) >> 6, gy_ = (unsigned)(G.Y + 3) >> 2; \
    const int x = (int)(vb.id % gx_) * 64 + (int)threadIdx.x;          \
    const int y = (int)((vb.id / gx_) % gy_) * 4 + (int)threadIdx.y;   \
    const int z = (int)(vb.id / (gx_ * gy_));                          \
    const bool inb = (x < G.X) && (y < G.Y);                           \
    const int bid = (int)vb.id

// Workgroups are dispatched to the 8 XCDs round-robin by index. Remapping
// gives each XCD a contiguous range of tiles, so the rows a tile reads from
// its y neighbours were mostly fetched into the same XCD's L2.
__device__ __forceinline__ int xcdRemap(int b, int nb) {
    return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// z-marching: the block's 64x4 tile walks planes [z0, z1); 1-D grid of
// gx * gy * ceil(Z / zc) blocks.
#define ZM_SETUP                                                   \
    const int gx_ = (G.X + 63) >> 6, gy_ = (G.Y + 3) >> 2;         \
    const int bid = xcdRemap(blockIdx.x, gridDim.x);               \
    const int x = (bid % gx_) * 64 + threadIdx.x;                  \
    const int y = ((bid / gx_) % gy_) * 4 + threadIdx.y;           \
    const int z0 = (bid / (gx_ * gy_)) * zc;                       \
    const int z1 = min(G.Z, z0 + zc);                              \
    const bool inb = (x < G.X) && (y < G.Y)

// The same from the launch's tile index.
#define ZM_SETUP_V                                                 \
    const int gx_ = (G.X + 63) >> 6, gy_ = (G.Y + 3) >> 2;         \
    const int bid = xcdRemap((int)vb.id, (int)vb.n);               \
    const int x = (bid % gx_) * 64 + threadIdx.x;                  \
    const int y = ((bid / gx_) % gy_) * 4 + threadIdx.y;           \
    const int z0 = (bid / (gx_ * gy_)) * zc;                       \
    const int z1 = min(G.Z, z0 + zc);                              \
    const bool inb = (x < G.X) && (y < G.Y)

// The same for a block whose tile index (already remapped) is given.
#define ZM_SETUP_B(blin)                                           \
    const int gx_ = (G.X + 63) >> 6, gy_ = (G.Y + 3) >> 2;         \
    const int bid = (int)(blin);                                   \
    const int x = (bid % gx_) * 64 + threadIdx.x;                  \
    const int y = ((bid / gx_) % gy_) * 4 + threadIdx.y;           \
    const int z0 = (bid / (gx_ * gy_)) * zc;                       \
    const int z1 = min(G.Z, z0 + zc);                              \
    const bool inb = (x < G.X) && (y < G.Y)

static inline dim3 grid3(const EnvState& e) { return dim3((e.X + 63) / 64, (e.Y + 3) / 4, e.Z); }
// Per-cell passes that give each thread ZPT cells along z (ZM_SETUP_V with
// zc = ZPT): the loads of the thread's cells are issued together.
constexpr int ZPT = 4;
static inline unsigned gridZpt(const EnvState& e) {
    return (unsigned)((e.X + 63) / 64) * (unsigned)((e.Y + 3) / 4) * (unsigned)((e.Z + ZPT - 1) / ZPT);
}
static inline dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }
static const dim3 B3(64, 4, 1);

}  // namespace skt
