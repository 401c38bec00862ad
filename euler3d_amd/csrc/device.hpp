// Device-side mesh description shared by the kernels and the host context.
#pragma once
#include <hip/hip_runtime.h>

namespace e3d {

constexpr int kWG = 256;          // threads per workgroup of the granular kernels; block real ranges are padded to this
constexpr int kMaxSlots = 6;      // faces per cell: tet 4, pyramid 5, hexahedron 6

// Ghost kind bits.
constexpr int GK_BC_MASK = 3;     // 0 none, 1 slip (wall/symmetry), 2 farfield
constexpr int GK_SLIP = 1;
constexpr int GK_FARFIELD = 2;
constexpr int GK_REFLECT = 4;     // gradient reflected from the interior cell
constexpr int GK_WALL = 8;        // marker is WALL (aerodynamic coefficients)

// Global cell space: real cells of every block in [0, NRp) (block b from roff[b],
// padded to kWG), then ghosts of every block in [NRp, NA). Arrays of real-cell
// quantities have NRp entries, of all-cell quantities NA, of faces NF.
// Per-cell face slots are slot-major: x[k*NRp + c], k < nslot, in the
// reference's ascending face order.
struct DMesh {
    int NRp, NG, NA, NF, nslot;
    const int* nf;        // [NRp] faces of the cell (0 = padding)
    const int* nbr;       // [nslot][NRp] neighbour (global cell)
    const int* fid;       // [nslot][NRp] global face, bit 31 set when the cell is the face's right cell
    const double* S;      // [3][NF] area vectors
    const double* area;   // [NF]
    const double* rL;     // [3][NF]
    const double* rR;     // [3][NF]
    const double* vol;    // [NA]
    const double* lsqw;   // [nslot*3][NRp] least-squares weights, (k*3+d)*NRp + c
    const int* g_in;      // [NG] interior (global) of each ghost
    const int* g_face;    // [NG] face of each ghost
    const int* g_kind;    // [NG] GK_* bits
    const int* wg_block;  // [NRp/kWG] block of each kWG-cell range
    // Per-slot copies of face geometry for the fused path, slot-major so a
    // workgroup streams them coalesced: entry of slot k of cell c at k*NRp + c,
    // each vector as an (x, y) pair plane and a z plane.
    const double2* sSxy;  // [nslot][NRp] area vector S of the face, negated when this cell is its right cell
    const double* sSz;
    const double2* sRxy;  // [nslot][NRp] the face's r vector on this cell's side
    const double* sRz;
    const int* mate;      // [nslot][NRp] face-state record of the other cell of the face; bit 31: this cell is right
    // Tet meshes: each real cell's four node coordinates in its node order, as
    // six (x, y)-style pair planes (x0,y0) (z0,x1) (y1,z1) (x2,y2) (z2,x3)
    // (y3,z3), and its slot code (tet_geometry.hpp), from which the Green-Gauss
    // kernel rebuilds S and r of every slot; null where the mesh has other cells
    // or a slot's rebuilt values would differ from the stored ones.
    const double2* cnode; // [6][NRp]
    const unsigned* cgeo; // [NRp], 0 on padding cells
    const double* g_r;    // [3][NG] r vector of each ghost on its face (the reflected right vector)
    const double* g_n;    // [3][NG] unit normal S/|S| of each ghost's face (S as stored: out of the interior cell)
    // Instances: a context may hold a batch of independent meshes, each with its
    // own flow conditions; a single mesh is instance 0.
    const int* blk_inst;  // [nb] instance of each block
    const double* inst_fs;  // [n_inst][5] freestream primitive state of each instance
    const int* active;    // [n_inst] nonzero while the instance iterates; null: every instance iterates
    int tile_map;         // tile order (tile_of): 0 linear, 1 XCD-contiguous, 2 XCD chunks
};

// Face-state records of the fused path: record k*NRp + c holds the
// reconstructed primitive state of real cell c on its k-th face, record
// nslot*NRp + i that of ghost i on its face; five variables per record, stored
// in three planes: (rho, u) pairs, (v, w) pairs, p.
__device__ __host__ inline size_t ghost_record(int nslot, int NRp, int i) { return (size_t)nslot * NRp + i; }

// Cell range of a granular launch: real cells [c0, c1) (global), ghosts
// [g0, g1) (ghost index, global cell NRp + i). One block or all blocks.
struct Range {
    int c0, c1, g0, g1;
};

struct DState {
    double* W;            // [5][NA] primitive state (current)
    double* Wn;           // [5][NA] next-stage buffer of the fused path
    double* W0;           // [5][NRp]
    double* conv;         // [5][NRp]
    double* diss;         // [5][NRp]
    double* grad;         // [15][NA]
    double* lim;          // [5][NA]
    double* Wf;           // [5 * (nslot*NRp + NG)] reconstructed face states (fused path), three planes
    double* dt;           // [NRp]
    double* specrad;      // [NRp]
    double* part;         // [NRp/64][8] per-wavefront partial sums of the residual monitor
    double* jac;          // [10][NRp] Jacobi values of the residual smoother (allocated on first use)
};

// Per-instance constants of a batched run's monitors (PostProcessing/src/PostProcessing.cpp:222-250,
// AerodynamicParameters.cpp:120-132): rotation of the force sums by the inlet angle, stop
// criterion and iteration limit.
struct InstParam {
    double cos_a, sin_a, criterion;
    int max_iter, has_ref;  // has_ref: the drop's reference norms were given, not taken at iteration 0
};

// One monitor row, laid out as e3d_monitor (include/euler3d.h).
struct MonRow {
    int iter, stop;
    double rms[5], drop[5], cl, cd, cmx, cmy, cmz;
};

}  // namespace e3d
