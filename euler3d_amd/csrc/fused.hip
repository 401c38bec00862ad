// Fused path: each RK stage is two kernels over all blocks at once.
//
//   k_interp : gradient (Green-Gauss or least squares) + limiter of every real
//              cell, reflected gradient and limiter of the ghosts hanging off
//              it, and the reconstructed state W + (lim*grad).r of the cell on
//              each of its faces, written as one 5-double record per
//              (cell, face).
//   k_flux   : per cell and face, the flux from the cell's own face-state
//              record and the other cell's record (one gathered record),
//              accumulated in the reference's face order; then the RK update
//              and the ghost BCs, into the second state buffer. On stage 0
//              also saveW0, the spectral radius and the local time step.
//
// Face geometry is stored per (cell, face) slot so it streams coalesced: the
// face's area vector S, oriented away from the cell, and the r vector on the
// cell's side. k_flux rebuilds the unit normal and |S| from S with the
// reference's own expressions (FluxScheme.cpp:105-112), which give the same
// doubles as the stored face normal and area. The only gathers are the
// neighbour states in k_interp and one face-state record per face in k_flux.
// Every value equals the granular path's bit for bit: the per-slot copies hold
// the same doubles, and the reconstruction, gradient, limiter and flux
// expressions are evaluated in the reference's order.
//
// Both kernels are instantiated for tiles of 64, 128 or 256 cells per
// workgroup (e3d_set_tile_cells); a thread owns one cell either way, so the
// tile size moves no result bit. The last stage's residual-norm partial sums
// are formed per wavefront (64 cells) for the same reason.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kcommon.hpp"
#include "physics.hpp"
#include "tet_geometry.hpp"

namespace e3d {

__device__ __forceinline__ double slotv(const double* a, int q, int nslot, int NRp, int k, int c) {
    return a[((size_t)q * nslot + k) * NRp + c];
}

// Element i >= 0 of an array whose base is the same for the whole wavefront,
// addressed as base + 32-bit byte offset (i * sizeof(T) < 2^32 for T up to 16
// bytes, which the upload's limit of 2^28 cells guarantees): one address
// register per lane instead of two, and no 64-bit address arithmetic.
template <class T>
__device__ __forceinline__ T ld32(const T* base, int i) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (unsigned)i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void st32(T* base, int i, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (unsigned)i * (unsigned)sizeof(T)) = v;
}
// Streaming (nontemporal) forms of ld32/st32, for data a launch touches once
// in a coalesced stream: per-cell and per-slot inputs (face count, volume,
// neighbour and record ids, slot geometry, W0, dt, dissipation) and the
// face-state records, residuals, W0, dt and new state the kernels write.
// They leave the caches to the gathered neighbour states and records.
// Measured on the 4.7M-cell channel in 8 blocks: 5.14 -> 4.94 ms per
// iteration. Streaming loads of the records in k_flux (each is read again as
// the other side's record) and streaming stores of the scattered ghost values
// measured slower and are not used.
typedef double double2_ev __attribute__((ext_vector_type(2)));
template <class T>
__device__ __forceinline__ void st32_stream(T* base, int i, T v) {
    __builtin_nontemporal_store(v, reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (unsigned)i * (unsigned)sizeof(T)));
}
__device__ __forceinline__ void st32_stream(double2* base, int i, double2 v) {
    double2_ev w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<double2_ev*>(reinterpret_cast<char*>(base) + (unsigned)i * 16u));
}
template <class T>
__device__ __forceinline__ T ld32_stream(const T* base, int i) {
    return __builtin_nontemporal_load(
        reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (unsigned)i * (unsigned)sizeof(T)));
}
__device__ __forceinline__ double2 ld32_stream(const double2* base, int i) {
    const double2_ev v = __builtin_nontemporal_load(
        reinterpret_cast<const double2_ev*>(reinterpret_cast<const char*>(base) + (unsigned)i * 16u));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void store_w32(double* W, int NA, int x, const double w[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) st32_stream(W + (size_t)v * NA, x, w[v]);
}
// S (oriented away from this cell) and r vector of slot k of cell c.
__device__ __forceinline__ void slot_S(const DMesh& m, int k, int c, double S[3]) {
    const size_t o = (size_t)k * m.NRp;
    const double2 xy = ld32_stream(m.sSxy + o, c);
    S[0] = xy.x;
    S[1] = xy.y;
    S[2] = ld32_stream(m.sSz + o, c);
}
__device__ __forceinline__ void slot_r(const DMesh& m, int k, int c, double r[3]) {
    const size_t o = (size_t)k * m.NRp;
    const double2 xy = ld32_stream(m.sRxy + o, c);
    r[0] = xy.x;
    r[1] = xy.y;
    r[2] = ld32_stream(m.sRz + o, c);
}
__device__ __forceinline__ void load_w32(const double* W, int NA, int x, double w[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) w[v] = ld32(W + (size_t)v * NA, x);
}

// Face-state records in three planes so that a record is three loads: (rho, u)
// and (v, w) as 16-byte pairs, p alone. Record r: pair q (q = 0, 1) at
// Wf[2*q*nrec + 2*r], p at Wf[4*nrec + r]; the records of a wavefront's cells
// on one face slot are contiguous in each plane.
__device__ __forceinline__ size_t rec_count(const DMesh& m) { return (size_t)m.nslot * m.NRp + m.NG; }

__device__ __forceinline__ void load_rec(const double* Wf, size_t nrec, size_t rec, double w[5]) {
    const double2 a = reinterpret_cast<const double2*>(Wf)[rec];
    const double2 b = reinterpret_cast<const double2*>(Wf + 2 * nrec)[rec];
    w[0] = a.x;
    w[1] = a.y;
    w[2] = b.x;
    w[3] = b.y;
    w[4] = Wf[4 * nrec + rec];
}

// Address of variable v of record r.
__device__ __forceinline__ size_t rec_index(size_t nrec, size_t rec, int v) {
    return v < 4 ? (v >> 1) * 2 * nrec + 2 * rec + (v & 1) : 4 * nrec + rec;
}

__device__ __forceinline__ void store_rec(double* Wf, size_t nrec, size_t rec, const double w[5]) {
    reinterpret_cast<double2*>(Wf)[rec] = make_double2(w[0], w[1]);
    reinterpret_cast<double2*>(Wf + 2 * nrec)[rec] = make_double2(w[2], w[3]);
    Wf[4 * nrec + rec] = w[4];
}

// Record of cell c on its face slot k (record k*NRp + c), with the slot's
// plane base uniform and a 32-bit offset per lane.
__device__ __forceinline__ void load_own_rec(const double* Wf, size_t nrec, int NRp, int k, int c, double w[5]) {
    const size_t r0 = (size_t)k * NRp;
    const double2 a = ld32(reinterpret_cast<const double2*>(Wf) + r0, c);
    const double2 b = ld32(reinterpret_cast<const double2*>(Wf + 2 * nrec) + r0, c);
    w[0] = a.x;
    w[1] = a.y;
    w[2] = b.x;
    w[3] = b.y;
    w[4] = ld32(Wf + 4 * nrec + r0, c);
}
__device__ __forceinline__ void store_own_rec(double* Wf, size_t nrec, int NRp, int k, int c, const double w[5]) {
    const size_t r0 = (size_t)k * NRp;
    st32_stream(reinterpret_cast<double2*>(Wf) + r0, c, make_double2(w[0], w[1]));
    st32_stream(reinterpret_cast<double2*>(Wf + 2 * nrec) + r0, c, make_double2(w[2], w[3]));
    st32_stream(Wf + 4 * nrec + r0, c, w[4]);
}

// W + (lim*grad).r, accumulated per component as FluxScheme.cpp:127-134 does.
__device__ __forceinline__ void face_state(const double w[5], const double lim[5], const double g[15], const double r[3],
                                           double out[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = w[v];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < 5; ++v) out[v] += lim[v] * g[v * 3 + d] * r[d];
}

// Limiter, reconstructed state and record of the ghost hanging off face slot k
// of real cell c, one variable at a time: reflected (or zero) gradient
// (GreenGauss.cpp:175-234), the ghost's limiter over its single neighbour c
// (Venkatakrishnan.cpp:87-243 run over a ghost), and W + (lim*grad).r with the
// ghost's reflected r vector.
template <bool VENKAT>
__device__ __forceinline__ void ghost_face_state(const DMesh& m, const DState& s, int k, int c, int nb,
                                                 const double g[15], const double wc[5], double eps2, double omega) {
    const int NRp = m.NRp, ns = m.nslot;
    const int gi = nb - NRp;
    const bool reflect = m.g_kind[gi] & GK_REFLECT;
    double n[3], rg[3];
    load3(m.g_n, m.NG, gi, n);  // S / area
    load3(m.g_r, m.NG, gi, rg);
    const size_t nrec = rec_count(m), rec = ghost_record(ns, NRp, gi);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        double gg[3];
        if (reflect) {
            double dot = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) dot += g[v * 3 + d] * n[d];
#pragma unroll
            for (int d = 0; d < 3; ++d) gg[d] = g[v * 3 + d] - 2.0 * dot * n[d];
        } else {
            gg[0] = gg[1] = gg[2] = 0.0;
        }
        const double wg = s.W[(size_t)v * m.NA + nb];
        double mx = wg, mn = wg;
        if (wc[v] > mx) mx = wc[v];
        if (wc[v] < mn) mn = wc[v];
        double d2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) d2 += gg[d] * rg[d];
        const double phi = face_limiter<VENKAT>(d2, mx - wg, mn - wg, eps2, omega);
        double lim = __builtin_inf();
        if (phi < lim) lim = phi;
        double out = wg;
#pragma unroll
        for (int d = 0; d < 3; ++d) out += lim * gg[d] * rg[d];
        s.Wf[rec_index(nrec, rec, v)] = out;
    }
}

// Position of the j-th set bit (j from 0) of a 64-bit lane mask, j < popcount(mask).
__device__ __forceinline__ int nth_set_bit(unsigned long long mask, int j) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const unsigned long long low = (1ull << w) - 1;
        const int n = __popcll(mask & low);
        if (j >= n) {
            j -= n;
            mask >>= w;
            pos += w;
        } else {
            mask &= low;
        }
    }
    return pos;
}

// The ghost slots of a wavefront dealt to its active lanes, for code that
// only boundary cells run: slot k of a lane is a task where bit k of the
// lane's slot set is set. Tasks are numbered slot by slot, lanes in order
// within a slot, and handed to the active lanes in rank order, nact at a
// time, so one pass of the code serves up to nact ghosts.
template <int MAXS>
struct SlotTasks {
    unsigned long long mask[MAXS];
    int first[MAXS + 1];
    int rank, nact;
    __device__ __forceinline__ explicit SlotTasks(unsigned slots) {
        const int lane = __lane_id();
        const unsigned long long act = __ballot(1);
        rank = __popcll(act & ((1ull << lane) - 1));
        nact = __popcll(act);
        first[0] = 0;
#pragma unroll
        for (int k = 0; k < MAXS; ++k) {
            mask[k] = __ballot((slots >> k) & 1u);
            first[k + 1] = first[k] + __popcll(mask[k]);
        }
    }
    __device__ __forceinline__ int count() const { return first[MAXS]; }  // the same on every lane
    // slot and owning lane of task t < count()
    // (selects with constant indices only: an array indexed at run time
    // would be placed in scratch memory)
    __device__ __forceinline__ void at(int t, int& k, int& src) const {
        k = 0;
        int base = 0;
        unsigned long long msk = mask[0];
#pragma unroll
        for (int kk = 1; kk < MAXS; ++kk)
            if (t >= first[kk]) {
                k = kk;
                base = first[kk];
                msk = mask[kk];
            }
        src = nth_set_bit(msk, t - base);
    }
};

// Lane src's value of v[k] for this lane's k (every lane takes part).
template <int MAXS>
__device__ __forceinline__ int shfl_slot(const int (&v)[MAXS], int k, int src) {
    int r = __shfl(v[0], src, 64);
#pragma unroll
    for (int kk = 1; kk < MAXS; ++kk) {
        const int x = __shfl(v[kk], src, 64);
        if (k == kk) r = x;
    }
    return r;
}

// One cell of k_interp. FULL: every cell of the wavefront has MAXS faces (a
// tet mesh with MAXS = 4 always does), so the per-slot tests drop out and the
// face slots are straight-line code whose loads the compiler can overlap;
// otherwise slot k is skipped where k >= nf. Both forms evaluate the same
// expressions in the same order.
//
// NODE (tet meshes, Green-Gauss): S and r of each slot are rebuilt from the
// cell's four node coordinates and its slot code (tet_geometry.hpp), 100
// bytes per cell, instead of streaming the stored per-slot copies, 192 bytes;
// the doubles are the same (the upload checks every slot before it enables
// this form). The nodes are staged in the lane's column of the r-vector LDS
// array (12 doubles, the size of the four r vectors), S is formed from them
// per slot in pass 1, and the r vectors then replace them for passes 2 and 3.
// Each lane touches only its own column, so no barrier is needed.
template <int GRAD, bool VENKAT, int MAXS, int WG, bool FULL, bool NODE>
__device__ __forceinline__ void interp_cell(const DMesh& m, const DState& s, double k3, double omega, int c, int nf,
                                            unsigned code, double (*r_lds)[WG]) {
    static_assert(!NODE || (GRAD == 0 && MAXS == 4), "node-rebuilt geometry: Green-Gauss on tets");
    const int NRp = m.NRp, ns = m.nslot;
    const int t = threadIdx.x;
    double wc[5];
    load_w32(s.W, m.NA, c, wc);
    const double vol = ld32_stream(m.vol, c);
    // pass 1: gradient and the neighbour extrema, faces in the reference's order.
    // The neighbour ids of all slots are loaded first (slots k >= nf hold
    // in-bounds padding and are not used), so each face then waits for one
    // memory round trip (its neighbour's state and its geometry), not two.
    double g[15], mx[5], mn[5];
    unsigned ghost_slots = 0;  // bit k: the neighbour across face slot k is a ghost
    int nbv[MAXS];
#pragma unroll
    for (int k = 0; k < MAXS; ++k) nbv[k] = k < ns ? ld32_stream(m.nbr + (size_t)k * NRp, c) : c;
    if (NODE) {  // node j's x, y, z in rows 3j, 3j+1, 3j+2; planes (x0,y0) (z0,x1) (y1,z1) (x2,y2) (z2,x3) (y3,z3)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const double2 p = ld32_stream(m.cnode + (size_t)j * NRp, c);
            r_lds[2 * j][t] = p.x;
            r_lds[2 * j + 1][t] = p.y;
        }
        asm volatile("" ::: "memory");  // read back per slot, not forwarded in registers (as below)
    }
#pragma unroll
    for (int i = 0; i < 15; ++i) g[i] = 0.0;
#pragma unroll
    for (int v = 0; v < 5; ++v) mx[v] = mn[v] = wc[v];
    struct Slot {
        double wn[5], geo[3], r[3];
    };
    // the neighbour's state, S oriented away from this cell (Green-Gauss) or the
    // least-squares weights, and the r vector of face slot k
    // (NODE: the neighbour's state only; S is formed when the slot is used)
    auto load_slot = [&](int k, Slot& q) {
        load_w32(s.W, m.NA, nbv[k], q.wn);
        if (NODE) return;
        if (GRAD == 0) {
            slot_S(m, k, c, q.geo);
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) q.geo[d] = ld32_stream(m.lsqw + (size_t)(k * 3 + d) * NRp, c);
        }
        slot_r(m, k, c, q.r);
    };
    // node i (0-2) of slot k's face, from the staged node coordinates
    auto face_node = [&](int k, int i, double& x, double& y, double& z) {
        const int l = (code >> (8 * k + 2 * i)) & 3;
        x = r_lds[3 * l][t];
        y = r_lds[3 * l + 1][t];
        z = r_lds[3 * l + 2][t];
    };
    // its terms, in the reference's order; the r vector goes to LDS for the
    // limiter and the reconstruction
    auto use_slot = [&](int k, Slot& q) {
        if (NODE) {
            double x1, y1, z1, x2, y2, z2, x3, y3, z3;
            face_node(k, 0, x1, y1, z1);
            face_node(k, 1, x2, y2, z2);
            face_node(k, 2, x3, y3, z3);
            tri_area_vector(x1, y1, z1, x2, y2, z2, x3, y3, z3, (code >> (8 * k + 6)) & 1, q.geo);
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d) r_lds[k * 3 + d][t] = q.r[d];
        }
        if (nbv[k] >= NRp) ghost_slots |= 1u << k;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            if (q.wn[v] > mx[v]) mx[v] = q.wn[v];
            if (q.wn[v] < mn[v]) mn[v] = q.wn[v];
        }
        if (GRAD == 0) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 0; v < 5; ++v) g[v * 3 + d] += 0.5 * (wc[v] + q.wn[v]) * q.geo[d];
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v)
#pragma unroll
                for (int d = 0; d < 3; ++d) g[v * 3 + d] += q.geo[d] * (q.wn[v] - wc[v]);
        }
    };
    if (FULL) {
        // one slot ahead: slot k + 1's loads are in flight while slot k's terms
        // are formed; the scheduling barriers keep the compiler from hoisting
        // further loads (more registers, fewer waves)
        Slot q[2];
        load_slot(0, q[0]);
#pragma unroll
        for (int k = 0; k < MAXS; ++k) {
            if (k + 1 < MAXS) load_slot(k + 1, q[(k + 1) & 1]);
            __builtin_amdgcn_sched_barrier(0);
            use_slot(k, q[k & 1]);
            __builtin_amdgcn_sched_barrier(0);
        }
    } else {
#pragma unroll
        for (int k = 0; k < MAXS; ++k)
            if (k < nf) {
                Slot q;
                load_slot(k, q);
                use_slot(k, q);
            }
    }
    if (NODE) {  // the r vectors of every slot replace the staged nodes (held in registers meanwhile)
        double x[4], y[4], z[4], cc[3];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            x[j] = r_lds[3 * j][t];
            y[j] = r_lds[3 * j + 1][t];
            z[j] = r_lds[3 * j + 2][t];
        }
        tet_centre(x, y, z, cc);
        double r[3 * MAXS];
#pragma unroll
        for (int k = 0; k < MAXS; ++k) {
            double x1, y1, z1, x2, y2, z2, x3, y3, z3;
            face_node(k, 0, x1, y1, z1);
            face_node(k, 1, x2, y2, z2);
            face_node(k, 2, x3, y3, z3);
            tri_r_vector(x1, y1, z1, x2, y2, z2, x3, y3, z3, cc, r + 3 * k);
        }
#pragma unroll
        for (int i = 0; i < 3 * MAXS; ++i) r_lds[i][t] = r[i];
        // passes 2 and 3 read the r vectors back from LDS: without this the
        // compiler forwards the stored values and keeps all twelve in
        // registers through both passes (spilling to scratch)
        asm volatile("" ::: "memory");
    }
    if (GRAD == 0) {  // g / vol (GreenGauss.cpp:153-163); g is a sum started from +0.0
        const double rv = 1.0 / vol;
#pragma unroll
        for (int i = 0; i < 15; ++i) g[i] = div_by(g[i], vol, rv);
    }
    // pass 2: limiter over the faces, r vectors from LDS
    const double eps2 = k3 * (vol);
    double lim[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        mx[v] -= wc[v];
        mn[v] -= wc[v];
        lim[v] = __builtin_inf();
    }
    // (the node-geometry form one slot at a time, for its register budget)
#define E3D_LIMIT_SLOT(k)                                                                   \
    {                                                                                       \
        double r[3];                                                                        \
        _Pragma("unroll") for (int d = 0; d < 3; ++d) r[d] = r_lds[(k) * 3 + d][threadIdx.x]; \
        _Pragma("unroll") for (int v = 0; v < 5; ++v) {                                     \
            double d2 = 0.0;                                                                \
            _Pragma("unroll") for (int d = 0; d < 3; ++d) d2 += g[v * 3 + d] * r[d];        \
            const double phi = face_limiter<VENKAT>(d2, mx[v], mn[v], eps2, omega);         \
            if (phi < lim[v]) lim[v] = phi;                                                 \
        }                                                                                   \
    }
    if (NODE) {
#pragma unroll 1
        for (int k = 0; k < MAXS; ++k) E3D_LIMIT_SLOT(k)
    } else {
#pragma unroll
        for (int k = 0; k < MAXS; ++k)
            if (k < nf) E3D_LIMIT_SLOT(k)
    }
#undef E3D_LIMIT_SLOT
    // pass 3: face states of the cell
#pragma unroll
    for (int k = 0; k < MAXS; ++k)
        if (k < nf) {
            double r[3], out[5];
#pragma unroll
            for (int d = 0; d < 3; ++d) r[d] = r_lds[k * 3 + d][threadIdx.x];
            face_state(wc, lim, g, r, out);
            store_own_rec(s.Wf, rec_count(m), NRp, k, c, out);
        }
    // Ghosts hanging off the cells of this wavefront: their face states, by
    // ghost_face_state with the interior cell's gradient, state and eps2 (a
    // ghost's volume is its interior cell's). The wavefront's ghost slots are
    // dealt to its active lanes in turn (slot 0's ghosts in lane order, then
    // slot 1's, ...), each lane fetching its task's operands from the interior
    // cell's lane (the gradient through that lane's LDS column, free after
    // pass 3, the rest by lane shuffles), so one pass serves every ghost of
    // the wavefront instead of one divergent pass per face slot. Same
    // operands, same records.
    {
        const SlotTasks<MAXS> tasks(ghost_slots);
        const int ntask = tasks.count(), lane = __lane_id();
        if (ntask > 0) {
#pragma unroll
            for (int i = 0; i < 15; ++i) r_lds[i][t] = g[i];
            // the stores above before the other lanes' loads below (LDS
            // operations of one wavefront complete in order)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
#pragma unroll 1
        for (int t0 = 0; t0 < ntask; t0 += tasks.nact) {
            const int task = t0 + tasks.rank;
            const bool mine = task < ntask;
            int k = 0, src = lane;
            if (mine) tasks.at(task, k, src);
            double gs[15], ws[5];
#pragma unroll
            for (int i = 0; i < 15; ++i) gs[i] = r_lds[i][t - lane + src];
#pragma unroll
            for (int v = 0; v < 5; ++v) ws[v] = __shfl(wc[v], src, 64);
            const double es = __shfl(eps2, src, 64);
            const int nb = shfl_slot<MAXS>(nbv, k, src);
            if (mine) ghost_face_state<VENKAT>(m, s, k, c - lane + src, nb, gs, ws, es, omega);
        }
    }
}

// Streamed geometry: four waves per SIMD (at most 128 VGPRs; 120-128 used, no
// scratch): with the slot pipeline the compiler would otherwise take more and
// run three, which measured no faster than the serial slots. Node-rebuilt
// geometry: three waves (140-154 VGPRs, no scratch). Measured on the 4.7M-cell
// channel: at four the Venkatakrishnan form keeps 20 bytes per lane in
// scratch and runs 0.392 against 0.388 ms over the first 100 iterations
// (0.364 against 0.368 ms while its limiter still branched on the sign),
// Barth-Jespersen keeps 28 bytes and runs 0.366 against 0.353 ms.
template <int GRAD, bool VENKAT, int MAXS, int WG, bool NODE>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(NODE ? 3 : 4)))
k_interp(DMesh m, DState s, double k3, double omega, int ntiles) {
    const int tile = tile_of<WG>(m, ntiles);
    if (tile >= ntiles || !instance_active(m, cell_instance(m, tile * WG))) return;
    const int c = tile * WG + threadIdx.x;
    // NODE: the slot code stands in for the face count (zero on padding cells, a tet's 4 otherwise)
    const unsigned code = NODE ? ld32_stream(m.cgeo, c) : 0u;
    const int nf = NODE ? (code ? MAXS : 0) : ld32_stream(m.nf, c);
    if (nf == 0) return;
    // r vectors of the cell's faces for passes 2 and 3, then its gradient for the ghost pass
    __shared__ double r_lds[MAXS * 3 > 15 ? MAXS * 3 : 15][WG];
    // (NODE: every cell left is a tet, nf == MAXS)
    if (NODE || __all(nf == MAXS)) interp_cell<GRAD, VENKAT, MAXS, WG, true, NODE>(m, s, k3, omega, c, nf, code, r_lds);
    else interp_cell<GRAD, VENKAT, MAXS, WG, false, NODE>(m, s, k3, omega, c, nf, code, r_lds);
}

// FIRST (stage 0) also does what the reference runs once per iteration before
// the stages: saveW0 (Solver.cpp:161-210) and the spectral radius and local
// time step (Timestep.cpp:67-146), all from the cell's own state at the start
// of the iteration, with the spectral radius summed in the same face loop.
template <int FLUX, bool ORDER2, bool DISS, bool LAST, bool FIRST, int MAXS, int WG>
__global__ void __launch_bounds__(WG) k_flux(DMesh m, DState s, Gas gas, double beta, double alpha, double cfl,
                                             double inv_gm1, int ntiles) {
    const int tile = tile_of<WG>(m, ntiles);
    if (tile >= ntiles) return;
    const int inst = cell_instance(m, tile * WG);
    if (!instance_active(m, inst)) return;
    const double* fs = m.inst_fs + 5 * inst;
    const int c = tile * WG + threadIdx.x;
    const int nf = ld32_stream(m.nf, c);
    const int NRp = m.NRp, ns = m.nslot;
    const size_t ghost0 = (size_t)ns * NRp;
    double res2[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (nf != 0) {
        double cv[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, ds[5], wc[5];
        if (!ORDER2 || FIRST) load_w32(s.W, m.NA, c, wc);
        double c_center = 0.0, sr = 0.0;
        if (FIRST) c_center = sqrt(gas.gamma * wc[4] / wc[0]);
#pragma unroll
        for (int v = 0; v < 5; ++v)
            ds[v] = DISS ? ld32_stream(s.diss + (size_t)v * NRp, c) * (1.0 - beta) : ld32_stream(s.diss + (size_t)v * NRp, c);
        // the other side's record index of every slot first (padding slots
        // hold 0, in bounds), so each face then waits for one round trip
        int mtv[MAXS];
#pragma unroll
        for (int k = 0; k < MAXS; ++k) mtv[k] = k < ns ? ld32_stream(m.mate + (size_t)k * NRp, c) : 0;
#pragma unroll
        for (int k = 0; k < MAXS; ++k)
            if (k < nf) {
                const int mt = mtv[k];
                const bool right = mt < 0;
                const size_t rec = (size_t)(mt & 0x7fffffff);
                // the face's S as stored (the slot holds it negated on the right
                // cell: negation is exact), its unit normal and |S|
                double Sv[3], n[3], norm;
                slot_S(m, k, c, Sv);
                const double sgn = right ? -1.0 : 1.0;  // exact: S * -1 is S negated
                unit_normal(Sv[0] * sgn, Sv[1] * sgn, Sv[2] * sgn, n, norm);
                double own[5], oth[5];
                if (FIRST) sr += (fabs((wc[1]) * n[0] + (wc[2]) * n[1] + (wc[3]) * n[2]) + (c_center)) * norm;
                if (ORDER2) {
                    load_own_rec(s.Wf, rec_count(m), NRp, k, c, own);
                    load_rec(s.Wf, rec_count(m), rec, oth);
                } else {
#pragma unroll
                    for (int v = 0; v < 5; ++v) own[v] = wc[v];
                    load_w(s.W, m.NA, m.nbr[(size_t)k * NRp + c], oth);
                }
                const double* wl = right ? oth : own;
                const double* wr = right ? own : oth;
                double F[5];
                if (FLUX == 0) central_flux(gas, wl, wr, n, norm, F);
                else ausm_conv(gas, wl, wr, n, norm, F);
                if (!right) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) cv[v] += F[v];
                } else {
#pragma unroll
                    for (int v = 0; v < 5; ++v) cv[v] -= F[v];
                }
                if (DISS) {
                    double D[5];
                    if (FLUX == 0) roe_dissipation(gas, wl, wr, n, norm, D);
                    else ausm_diss(gas, wl, wr, n, norm, D);
                    if (!right) {
#pragma unroll
                        for (int v = 0; v < 5; ++v) ds[v] += D[v] * beta;
                    } else {
#pragma unroll
                        for (int v = 0; v < 5; ++v) ds[v] -= D[v] * beta;
                    }
                }
            }
        if (DISS) {
#pragma unroll
            for (int v = 0; v < 5; ++v) st32_stream(s.diss + (size_t)v * NRp, c, ds[v]);
        }
        if (LAST) {  // the last stage's convective residual feeds the monitors and the array getters
#pragma unroll
            for (int v = 0; v < 5; ++v) st32_stream(s.conv + (size_t)v * NRp, c, cv[v]);
        }
        double w0[5], wnew[5], dt;
        const double vol = ld32_stream(m.vol, c);
        if (FIRST) {
            w0[0] = wc[0];
            w0[1] = wc[0] * wc[1];
            w0[2] = wc[0] * wc[2];
            w0[3] = wc[0] * wc[3];
            w0[4] = 0.5 * wc[0] * (wc[1] * wc[1] + wc[2] * wc[2] + wc[3] * wc[3]) + inv_gm1 * wc[4];
            dt = cfl * vol / sr;
#pragma unroll
            for (int v = 0; v < 5; ++v) st32_stream(s.W0 + (size_t)v * NRp, c, w0[v]);
            st32_stream(s.dt, c, dt);
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) w0[v] = ld32_stream(s.W0 + (size_t)v * NRp, c);
            dt = ld32_stream(s.dt, c);
        }
        rk_update(gas.gm1, alpha * dt / vol, w0, cv, ds, wnew);
        store_w32(s.Wn, m.NA, c, wnew);
        if (LAST) {
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const double r = cv[v] - ds[v];
                res2[v] = r * r;
            }
        }
        // ghosts of this cell: BC from the new interior state, or carried over
        auto ghost_bc = [&](int gi, const double wi[5]) {
            const int gcell = NRp + gi;
            const int kind = m.g_kind[gi];
            double gw[5];
            if (kind & GK_BC_MASK) {
                double n[3];
                load3(m.g_n, m.NG, gi, n);  // S / area
                if ((kind & GK_BC_MASK) == GK_SLIP) ghost_slip(wi, n, gw);
                else ghost_farfield(gas.gamma, wi, fs, n, gw);
            } else {
                load_w(s.W, m.NA, gcell, gw);
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) s.Wn[(size_t)v * m.NA + gcell] = gw[v];
        };
        if (FIRST) {  // slot by slot (measured 2 % faster than one pass for this variant)
#pragma unroll
            for (int k = 0; k < MAXS; ++k)
                if (k < nf) {
                    const size_t rec = (size_t)(mtv[k] & 0x7fffffff);
                    if (rec >= ghost0) ghost_bc((int)(rec - ghost0), wnew);
                }
        } else {  // one pass, the ghost slots dealt to the active lanes (1-1.5 % faster)
            unsigned ghost_slots = 0;
#pragma unroll
            for (int k = 0; k < MAXS; ++k)
                if (k < nf && (size_t)(mtv[k] & 0x7fffffff) >= ghost0) ghost_slots |= 1u << k;
            const SlotTasks<MAXS> tasks(ghost_slots);
            const int ntask = tasks.count(), lane = __lane_id();
#pragma unroll 1
            for (int t0 = 0; t0 < ntask; t0 += tasks.nact) {
                const int task = t0 + tasks.rank;
                const bool mine = task < ntask;
                int k = 0, src = lane;
                if (mine) tasks.at(task, k, src);
                double wi[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) wi[v] = __shfl(wnew[v], src, 64);  // the interior cell's new state
                const int mt = shfl_slot<MAXS>(mtv, k, src);
                if (mine) ghost_bc((int)((size_t)(mt & 0x7fffffff) - ghost0), wi);
            }
        }
    }
    if (LAST) {  // per-wavefront partial sums of the squared residual (fixed tree: deterministic)
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            double x = res2[v];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            if ((threadIdx.x & 63) == 0) s.part[(size_t)(c >> 6) * 8 + v] = x;
        }
    }
}

// ------------------------------------------------------------ host launchers

// nodes: rebuild the Green-Gauss slot geometry from the per-cell node copies
// (m.cnode, m.cgeo; tet meshes only), else stream the stored per-slot copies.
void launch_interp_fused(const DMesh& m, const DState& s, int grad, int venkat, double k3, double omega, int wg,
                         bool nodes, hipStream_t st) {
    const int nt = m.NRp / wg;
    const dim3 gr(xcd_grid(nt, wg)), bl(wg);
#define IL(G, V, S, W, N) hipLaunchKernelGGL((k_interp<G, V, S, W, N>), gr, bl, 0, st, m, s, k3, omega, nt)
#define IL_W(G, V, S, N) { if (wg == 64) IL(G, V, S, 64, N); else if (wg == 128) IL(G, V, S, 128, N); else IL(G, V, S, 256, N); }
#define IL_S(G, V) { if (m.nslot <= 4) IL_W(G, V, 4, false) else IL_W(G, V, 6, false) }
#define IL_V(G) { if (venkat) IL_S(G, true) else IL_S(G, false) }
    const bool node_form = grad == 0 && nodes && m.nslot == 4 && m.cnode && m.cgeo;
    if (node_form) {
        if (venkat) IL_W(0, true, 4, true) else IL_W(0, false, 4, true)
    } else if (grad == 0) IL_V(0) else IL_V(1)
#undef IL_V
#undef IL_S
#undef IL_W
#undef IL
}

void launch_flux_update(const DMesh& m, const DState& s, int flux, bool order2, bool diss, bool last, bool first,
                        const Gas& g, double beta, double alpha, double cfl, double inv_gm1, int wg, hipStream_t st) {
    const int nt = m.NRp / wg;
    const dim3 gr(xcd_grid(nt, wg)), bl(wg);
    // Stage 0 always has dissipation (beta_0 = 1 in every RK table), so FIRST implies DISS.
#define FL(FX, O2, DI, LA, FI, S, W) \
    hipLaunchKernelGGL((k_flux<FX, O2, DI, LA, FI, S, W>), gr, bl, 0, st, m, s, g, beta, alpha, cfl, inv_gm1, nt)
#define FL_W(FX, O2, DI, LA, FI, S) \
    { if (wg == 64) FL(FX, O2, DI, LA, FI, S, 64); else if (wg == 128) FL(FX, O2, DI, LA, FI, S, 128); else FL(FX, O2, DI, LA, FI, S, 256); }
#define FL_S(FX, O2, DI, LA, FI) { if (m.nslot <= 4) FL_W(FX, O2, DI, LA, FI, 4) else FL_W(FX, O2, DI, LA, FI, 6) }
#define FL_L(FX, O2, DI, FI) { if (last) FL_S(FX, O2, DI, true, FI) else FL_S(FX, O2, DI, false, FI) }
#define FL_D(FX, O2) { if (first) FL_L(FX, O2, true, true) else if (diss) FL_L(FX, O2, true, false) else FL_L(FX, O2, false, false) }
#define FL_O(FX) { if (order2) FL_D(FX, true) else FL_D(FX, false) }
    if (flux == 0) FL_O(0) else FL_O(1)
#undef FL_O
#undef FL_D
#undef FL_L
#undef FL_S
#undef FL_W
#undef FL
}

}  // namespace e3d
