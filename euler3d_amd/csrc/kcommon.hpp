// Small device helpers shared by the granular and fused kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace e3d {

// Tile of kWG cells handled by this workgroup, under one of three orders
// (e3d_set_tile_order). Workgroups are dispatched round-robin over the 8 XCDs
// (each with its own L2), so workgroup b runs on XCD b & 7 as that XCD's
// (b >> 3)-th workgroup.
//   0 linear:          tile = b; neighbouring tiles run on different XCDs.
//   1 XCD-contiguous:  XCD x walks tiles [x * per, (x + 1) * per).
//   2 XCD chunks:      runs of kTileChunk consecutive tiles dealt to the XCDs
//                      in turn: gathers within a run stay in one L2 while the
//                      whole chip moves through one window of memory.
// The grid (xcd_grid) is a multiple of 8 * kTileChunk; every order maps the
// grid one-to-one onto [0, grid), and workgroups whose tile is >= ntiles do
// nothing.
constexpr int kTileChunk = 16;
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int b = blockIdx.x;
    const int per = (ntiles + 7) >> 3;
    if ((b >> 3) >= per) return ntiles;  // grid padding beyond 8 * per
    return (b & 7) * per + (b >> 3);
}
__device__ __forceinline__ int chunk_tile() {
    const int b = blockIdx.x;
    const int x = b & 7, w = b >> 3;
    return ((w / kTileChunk) * 8 + x) * kTileChunk + (w % kTileChunk);
}
inline int xcd_grid(int ntiles) {
    const int g = 8 * kTileChunk;
    return g * ((ntiles + g - 1) / g);
}

// Tile of this workgroup under the mesh's tile order.
__device__ __forceinline__ int tile_of(const DMesh& m, int ntiles) {
    return m.tile_map == 2 ? chunk_tile() : m.tile_map == 1 ? xcd_tile(ntiles) : (int)blockIdx.x;
}

// Instance of a tile, and whether it still iterates.
__device__ __forceinline__ int tile_instance(const DMesh& m, int tile) { return m.blk_inst[m.wg_block[tile]]; }
__device__ __forceinline__ bool instance_active(const DMesh& m, int inst) { return !m.active || m.active[inst]; }

struct Free {  // freestream primitive state, by value in kernel arguments
    double v[5];
};

struct Slot {
    int nb, f;
    bool right;
};
__device__ __forceinline__ Slot slot(const DMesh& m, int k, int c) {
    Slot s;
    s.nb = m.nbr[(size_t)k * m.NRp + c];
    const int fw = m.fid[(size_t)k * m.NRp + c];
    s.right = fw < 0;
    s.f = fw & 0x7fffffff;
    return s;
}

__device__ __forceinline__ void load_w(const double* W, int NA, int x, double w[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) w[v] = W[(size_t)v * NA + x];
}

__device__ __forceinline__ void load3(const double* a, int N, int i, double r[3]) {
    r[0] = a[i];
    r[1] = a[(size_t)N + i];
    r[2] = a[2 * (size_t)N + i];
}

}  // namespace e3d
