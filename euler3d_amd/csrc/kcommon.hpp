// Small device helpers shared by the granular and fused kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace e3d {

// Tile of kWG cells handled by this workgroup. Workgroups are dispatched
// round-robin over the 8 XCDs; give each XCD a contiguous run of tiles so that
// neighbour gathers of adjacent tiles hit the same L2.
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int b = blockIdx.x;
    const int per = (ntiles + 7) >> 3;
    return (b & 7) * per + (b >> 3);
}
inline int xcd_grid(int ntiles) { return 8 * ((ntiles + 7) / 8); }

// Tile of this workgroup under the mesh's tile order.
__device__ __forceinline__ int tile_of(const DMesh& m, int ntiles) {
    return m.tile_map ? xcd_tile(ntiles) : (int)blockIdx.x;
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
