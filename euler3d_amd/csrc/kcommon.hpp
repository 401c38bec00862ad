// Small device helpers shared by the granular and fused kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace e3d {

// Tile of WG cells (64, 128 or 256: a divisor of the kWG-cell padding of
// every block, so a tile never straddles two blocks) handled by this
// workgroup, under one of three orders (e3d_set_tile_order). Workgroups are
// dispatched round-robin over the 8 XCDs (each with its own L2), so workgroup
// b runs on XCD b & 7 as that XCD's (b >> 3)-th workgroup.
//   0 linear:          tile = b; neighbouring tiles run on different XCDs.
//   1 XCD-contiguous:  XCD x walks tiles [x * per, (x + 1) * per).
//   2 XCD chunks:      runs of kChunkCells consecutive cells (kChunkCells / WG
//                      tiles) dealt to the XCDs in turn: gathers within a run
//                      stay in one L2 while the whole chip moves through one
//                      window of memory.
// The grid (xcd_grid) is a multiple of 8 runs; every order maps the grid
// one-to-one onto [0, grid), and workgroups whose tile is >= ntiles do
// nothing.
constexpr int kChunkCells = 4096;
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int b = blockIdx.x;
    const int per = (ntiles + 7) >> 3;
    if ((b >> 3) >= per) return ntiles;  // grid padding beyond 8 * per
    return (b & 7) * per + (b >> 3);
}
template <int WG>
__device__ __forceinline__ int chunk_tile() {
    constexpr int CH = kChunkCells / WG;
    const int b = blockIdx.x;
    const int x = b & 7, w = b >> 3;
    return ((w / CH) * 8 + x) * CH + (w % CH);
}
inline int xcd_grid(int ntiles, int wg) {
    const int g = 8 * (kChunkCells / wg);
    return g * ((ntiles + g - 1) / g);
}

// Tile of this workgroup under the mesh's tile order.
template <int WG>
__device__ __forceinline__ int tile_of(const DMesh& m, int ntiles) {
    return m.tile_map == 2 ? chunk_tile<WG>() : m.tile_map == 1 ? xcd_tile(ntiles) : (int)blockIdx.x;
}

// Instance of the block that real cell c belongs to, and whether it still iterates.
__device__ __forceinline__ int cell_instance(const DMesh& m, int c) { return m.blk_inst[m.wg_block[c / kWG]]; }
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
