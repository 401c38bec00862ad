// Fused path: each RK stage is two kernels over all blocks at once.
//
//   k_interp : gradient (Green-Gauss or least squares) + limiter of every real
//              cell, reflected gradient and limiter of the ghosts hanging off
//              it, and the reconstructed state W + (lim*grad).r of the cell on
//              each of its faces, written as one 5-double record per
//              (cell, face). On stage 0 also saveW0 and the local time step.
//   k_flux   : per cell and face, the flux from the cell's own face-state
//              record and the other cell's record (one 40-byte gather),
//              accumulated in the reference's face order; then the RK update
//              and the ghost BCs, into the second state buffer.
//
// Face geometry is stored per (cell, face) slot so it streams coalesced; the
// only gathers are the neighbour states in k_interp and one face-state record
// per face in k_flux. Every value equals the granular path's bit for bit: the
// per-slot copies hold the same doubles, and the reconstruction, gradient,
// limiter and flux expressions are evaluated in the reference's order.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kcommon.hpp"
#include "physics.hpp"

namespace e3d {

__device__ __forceinline__ double slotv(const double* a, int q, int nslot, int NRp, int k, int c) {
    return a[((size_t)q * nslot + k) * NRp + c];
}

__device__ __forceinline__ void load_rec(const double* Wf, size_t rec, double w[5]) {
    const double* p = Wf + rec * 5;
#pragma unroll
    for (int v = 0; v < 5; ++v) w[v] = p[v];
}

__device__ __forceinline__ void store_rec(double* Wf, size_t rec, const double w[5]) {
    double* p = Wf + rec * 5;
#pragma unroll
    for (int v = 0; v < 5; ++v) p[v] = w[v];
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

template <bool VENKAT>
struct Lim {
    double mx[5], mn[5], wc[5], lim[5];
    __device__ __forceinline__ void start(const double w[5]) {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            wc[v] = w[v];
            mx[v] = w[v];
            mn[v] = w[v];
            lim[v] = __builtin_inf();
        }
    }
    __device__ __forceinline__ void neighbour(const double wn[5]) {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            if (wn[v] > mx[v]) mx[v] = wn[v];
            if (wn[v] < mn[v]) mn[v] = wn[v];
        }
    }
    __device__ __forceinline__ void face(const double g[15], const double r[3], double eps2, double omega) {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            double d2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) d2 += g[v * 3 + d] * r[d];
            const double phi = face_limiter<VENKAT>(d2, mx[v] - wc[v], mn[v] - wc[v], eps2, omega);
            if (phi < lim[v]) lim[v] = phi;
        }
    }
};

__device__ __forceinline__ void save_w0_dt(const DMesh& m, const DState& s, int c, int nf, const double wc[5],
                                           double vol, double gamma, double inv_gm1, double cfl) {
    const size_t N = m.NRp;
    s.W0[c] = wc[0];
    s.W0[N + c] = wc[0] * wc[1];
    s.W0[2 * N + c] = wc[0] * wc[2];
    s.W0[3 * N + c] = wc[0] * wc[3];
    s.W0[4 * N + c] = 0.5 * wc[0] * (wc[1] * wc[1] + wc[2] * wc[2] + wc[3] * wc[3]) + inv_gm1 * wc[4];
    // Timestep::computeSpectralRadius (Timestep.cpp:67-119) and computeTimestep (:139-146)
    const double c_center = sqrt(gamma * wc[4] / wc[0]);
    double sr = 0.0;
    for (int k = 0; k < nf; ++k) {
        const double nx = slotv(m.sN, 0, m.nslot, m.NRp, k, c), ny = slotv(m.sN, 1, m.nslot, m.NRp, k, c),
                     nz = slotv(m.sN, 2, m.nslot, m.NRp, k, c);
        sr += (fabs((wc[1]) * nx + (wc[2]) * ny + (wc[3]) * nz) + (c_center)) * m.sA[(size_t)k * m.NRp + c];
    }
    s.dt[c] = cfl * vol / sr;
}

template <int GRAD, bool VENKAT, bool FIRST, int MAXS>
__global__ void __launch_bounds__(kWG) k_interp(DMesh m, DState s, double k3, double omega, double gamma,
                                                double inv_gm1, double cfl, int ntiles) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const int c = tile * kWG + threadIdx.x;
    const int nf = m.nf[c];
    if (nf == 0) return;
    const int NRp = m.NRp, ns = m.nslot;
    double wc[5];
    load_w(s.W, m.NA, c, wc);
    // pass 1: gradient and the neighbour extrema, faces in the reference's order
    double g[15];
#pragma unroll
    for (int i = 0; i < 15; ++i) g[i] = 0.0;
    Lim<VENKAT> L;
    L.start(wc);
#pragma unroll
    for (int k = 0; k < MAXS; ++k)
        if (k < nf) {
            double wn[5];
            load_w(s.W, m.NA, m.nbr[(size_t)k * NRp + c], wn);
            L.neighbour(wn);
            if (GRAD == 0) {
                double n[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) n[d] = slotv(m.sS, d, ns, NRp, k, c);
                if (m.mate[(size_t)k * NRp + c] < 0) {  // this cell is the face's right cell
                    n[0] *= -1.0;
                    n[1] *= -1.0;
                    n[2] *= -1.0;
                }
#pragma unroll
                for (int d = 0; d < 3; ++d)
#pragma unroll
                    for (int v = 0; v < 5; ++v) g[v * 3 + d] += 0.5 * (wc[v] + wn[v]) * n[d];
            } else {
                double wt[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) wt[d] = m.lsqw[(size_t)(k * 3 + d) * NRp + c];
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d) g[v * 3 + d] += wt[d] * (wn[v] - wc[v]);
            }
        }
    const double vol = m.vol[c];
    if (GRAD == 0) {
#pragma unroll
        for (int i = 0; i < 15; ++i) g[i] /= vol;
    }
    // pass 2: limiter over the faces
    const double eps2 = k3 * (vol);
#pragma unroll
    for (int k = 0; k < MAXS; ++k)
        if (k < nf) {
            double r[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) r[d] = slotv(m.sR, d, ns, NRp, k, c);
            L.face(g, r, eps2, omega);
        }
    // pass 3: face states of the cell; ghosts hanging off it
#pragma unroll
    for (int k = 0; k < MAXS; ++k)
        if (k < nf) {
            double r[3], out[5];
#pragma unroll
            for (int d = 0; d < 3; ++d) r[d] = slotv(m.sR, d, ns, NRp, k, c);
            face_state(wc, L.lim, g, r, out);
            store_rec(s.Wf, (size_t)k * NRp + c, out);
            const int nb = m.nbr[(size_t)k * NRp + c];
            if (nb < NRp) continue;
            // ghost: reflected (or zero) gradient, its limiter, its face state
            const int gi = nb - NRp;
            const int kind = m.g_kind[gi];
            double gg[15];
            if (kind & GK_REFLECT) {
                double n[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) n[d] = slotv(m.sN, d, ns, NRp, k, c);  // S / area
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    double dot = 0.0;
#pragma unroll
                    for (int d = 0; d < 3; ++d) dot += g[v * 3 + d] * n[d];
#pragma unroll
                    for (int d = 0; d < 3; ++d) gg[v * 3 + d] = g[v * 3 + d] - 2.0 * dot * n[d];
                }
            } else {
#pragma unroll
                for (int i = 0; i < 15; ++i) gg[i] = 0.0;
            }
            double rg[3], wg[5];
            load3(m.g_r, m.NG, gi, rg);
            load_w(s.W, m.NA, nb, wg);
            Lim<VENKAT> LG;
            LG.start(wg);
            LG.neighbour(wc);
            LG.face(gg, rg, eps2, omega);  // a ghost's volume is its interior cell's
            face_state(wg, LG.lim, gg, rg, out);
            store_rec(s.Wf, ghost_record(ns, NRp, gi), out);
        }
    if (FIRST) save_w0_dt(m, s, c, nf, wc, vol, gamma, inv_gm1, cfl);
}

// Stage 0 of a first-order run: saveW0 + spectral radius + time step.
__global__ void __launch_bounds__(kWG) k_prep(DMesh m, DState s, double gamma, double inv_gm1, double cfl,
                                              int ntiles) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const int c = tile * kWG + threadIdx.x;
    const int nf = m.nf[c];
    if (nf == 0) return;
    double wc[5];
    load_w(s.W, m.NA, c, wc);
    save_w0_dt(m, s, c, nf, wc, m.vol[c], gamma, inv_gm1, cfl);
}

template <int FLUX, bool ORDER2, bool DISS, bool LAST, int MAXS>
__global__ void __launch_bounds__(kWG) k_flux(DMesh m, DState s, Gas gas, double beta, double alpha, Free fs,
                                              int ntiles) {
    const int tile = xcd_tile(ntiles);
    if (tile >= ntiles) return;
    const int c = tile * kWG + threadIdx.x;
    const int nf = m.nf[c];
    const int NRp = m.NRp, ns = m.nslot;
    const size_t ghost0 = (size_t)ns * NRp;
    double res2[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (nf != 0) {
        double cv[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, ds[5], wc[5];
        if (!ORDER2) load_w(s.W, m.NA, c, wc);
#pragma unroll
        for (int v = 0; v < 5; ++v)
            ds[v] = DISS ? s.diss[(size_t)v * NRp + c] * (1.0 - beta) : s.diss[(size_t)v * NRp + c];
#pragma unroll
        for (int k = 0; k < MAXS; ++k)
            if (k < nf) {
                const int mt = m.mate[(size_t)k * NRp + c];
                const bool right = mt < 0;
                const size_t rec = (size_t)(mt & 0x7fffffff);
                double n[3], own[5], oth[5];
#pragma unroll
                for (int d = 0; d < 3; ++d) n[d] = slotv(m.sN, d, ns, NRp, k, c);
                const double norm = m.sA[(size_t)k * NRp + c];
                if (ORDER2) {
                    load_rec(s.Wf, (size_t)k * NRp + c, own);
                    load_rec(s.Wf, rec, oth);
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
            for (int v = 0; v < 5; ++v) s.diss[(size_t)v * NRp + c] = ds[v];
        }
        if (LAST) {  // the last stage's convective residual feeds the monitors and the array getters
#pragma unroll
            for (int v = 0; v < 5; ++v) s.conv[(size_t)v * NRp + c] = cv[v];
        }
        double w0[5], wnew[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) w0[v] = s.W0[(size_t)v * NRp + c];
        rk_update(gas.gm1, alpha * s.dt[c] / m.vol[c], w0, cv, ds, wnew);
#pragma unroll
        for (int v = 0; v < 5; ++v) s.Wn[(size_t)v * m.NA + c] = wnew[v];
        if (LAST) {
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const double r = cv[v] - ds[v];
                res2[v] = r * r;
            }
        }
        // ghosts of this cell: BC from the new interior state, or carried over
#pragma unroll
        for (int k = 0; k < MAXS; ++k)
            if (k < nf) {
                const size_t rec = (size_t)(m.mate[(size_t)k * NRp + c] & 0x7fffffff);
                if (rec < ghost0) continue;
                const int gi = (int)(rec - ghost0);
                const int gcell = NRp + gi;
                const int kind = m.g_kind[gi];
                double gw[5];
                if (kind & GK_BC_MASK) {
                    double n[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) n[d] = slotv(m.sN, d, ns, NRp, k, c);  // S / area
                    if ((kind & GK_BC_MASK) == GK_SLIP) ghost_slip(wnew, n, gw);
                    else ghost_farfield(gas.gamma, wnew, fs.v, n, gw);
                } else {
                    load_w(s.W, m.NA, gcell, gw);
                }
#pragma unroll
                for (int v = 0; v < 5; ++v) s.Wn[(size_t)v * m.NA + gcell] = gw[v];
            }
    }
    if (LAST) {  // per-tile partial sums of the squared residual (fixed tree: deterministic)
        __shared__ double red[kWG / 64][5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            double x = res2[v];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][v] = x;
        }
        __syncthreads();
        if (threadIdx.x < 5) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kWG / 64; ++w) t += red[w][threadIdx.x];
            s.part[(size_t)tile * 8 + threadIdx.x] = t;
        }
    }
}

// ------------------------------------------------------------ host launchers

void launch_interp_fused(const DMesh& m, const DState& s, int grad, int venkat, bool first, double k3, double omega,
                         double gamma, double inv_gm1, double cfl, hipStream_t st) {
    const int nt = m.NRp / kWG;
    const dim3 gr(xcd_grid(nt)), bl(kWG);
#define IL(G, V, F, S) hipLaunchKernelGGL((k_interp<G, V, F, S>), gr, bl, 0, st, m, s, k3, omega, gamma, inv_gm1, cfl, nt)
#define IL_S(G, V, F) { if (m.nslot <= 4) IL(G, V, F, 4); else IL(G, V, F, 6); }
#define IL_F(G, V) { if (first) IL_S(G, V, true) else IL_S(G, V, false) }
#define IL_V(G) { if (venkat) IL_F(G, true) else IL_F(G, false) }
    if (grad == 0) IL_V(0) else IL_V(1)
#undef IL_V
#undef IL_F
#undef IL_S
#undef IL
}

void launch_prep_fused(const DMesh& m, const DState& s, double gamma, double inv_gm1, double cfl, hipStream_t st) {
    const int nt = m.NRp / kWG;
    hipLaunchKernelGGL(k_prep, dim3(xcd_grid(nt)), dim3(kWG), 0, st, m, s, gamma, inv_gm1, cfl, nt);
}

void launch_flux_update(const DMesh& m, const DState& s, int flux, bool order2, bool diss, bool last, const Gas& g,
                        double beta, double alpha, const double fs[5], hipStream_t st) {
    const int nt = m.NRp / kWG;
    const dim3 gr(xcd_grid(nt)), bl(kWG);
    Free f;
    for (int v = 0; v < 5; ++v) f.v[v] = fs[v];
#define FL(FX, O2, DI, LA, S) hipLaunchKernelGGL((k_flux<FX, O2, DI, LA, S>), gr, bl, 0, st, m, s, g, beta, alpha, f, nt)
#define FL_S(FX, O2, DI, LA) { if (m.nslot <= 4) FL(FX, O2, DI, LA, 4); else FL(FX, O2, DI, LA, 6); }
#define FL_L(FX, O2, DI) { if (last) FL_S(FX, O2, DI, true) else FL_S(FX, O2, DI, false) }
#define FL_D(FX, O2) { if (diss) FL_L(FX, O2, true) else FL_L(FX, O2, false) }
#define FL_O(FX) { if (order2) FL_D(FX, true) else FL_D(FX, false) }
    if (flux == 0) FL_O(0) else FL_O(1)
#undef FL_O
#undef FL_D
#undef FL_L
#undef FL_S
#undef FL
}

}  // namespace e3d
