// HIP kernels of the residual-and-update path for gfx950, FP64 throughout.
//
// Cell-gather formulation: one thread per real cell walks the cell's faces in
// the reference's ascending face order, evaluates each face flux from the
// face's left/right states exactly as the reference's global face loop does,
// and adds (+ left cell, - right cell) into registers. A cell's residual is
// therefore the same sequence of IEEE operations as in the reference
// (Solver/src/FluxScheme.cpp:100-207), with no atomics and no colouring, and
// the result is deterministic and bit-identical for any cell ordering in
// memory. Each interior face is evaluated twice (once per side); the extra
// flops are the price of an atomic-free, order-exact gather.
//
// Granular kernels mirror the reference's operators one to one (and back the
// per-operator C ABI); the fused path runs each RK stage as two kernels:
//   interp_fused : gradient + limiter (+ ghost reflection, ghost limiters,
//                  and on stage 0 saveW0 + spectral radius + time step)
//   flux_update  : convective + dissipative residual + RK update + ghost BCs
//                  (+ residual-norm partial sums on the last stage)
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kcommon.hpp"
#include "physics.hpp"

namespace e3d {

// Reconstructed face state W + (lim*grad).r (FluxScheme.cpp:127-134), from the
// separate limiter and gradient arrays (granular path).
__device__ __forceinline__ void recon_lim_grad(const DState& s, int NA, int x, const double r[3], double w[5]) {
    load_w(s.W, NA, x, w);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < 5; ++v)
            w[v] += s.lim[(size_t)v * NA + x] * s.grad[(size_t)(v * 3 + d) * NA + x] * r[d];
}

// ---------------------------------------------------------------- granular

__global__ void __launch_bounds__(kWG) k_save_w0(Range rg, DMesh m, DState s, double inv_gm1) {
    const int c = rg.c0 + blockIdx.x * kWG + threadIdx.x;
    if (c >= rg.c1 || m.nf[c] == 0) return;
    double w[5];
    load_w(s.W, m.NA, c, w);
    const size_t N = m.NRp;
    s.W0[c] = w[0];
    s.W0[N + c] = w[0] * w[1];
    s.W0[2 * N + c] = w[0] * w[2];
    s.W0[3 * N + c] = w[0] * w[3];
    s.W0[4 * N + c] = 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) + inv_gm1 * w[4];
}

// Timestep::computeSpectralRadius (Timestep.cpp:67-119): own-cell state only.
__device__ __forceinline__ double spectral_radius(const DMesh& m, int c, int nf, const double w[5], double gamma) {
    const double c_center = sqrt(gamma * w[4] / w[0]);
    double sr = 0.0;
    for (int k = 0; k < nf; ++k) {
        const Slot sl = slot(m, k, c);
        double Sv[3], n[3], ds;
        load3(m.S, m.NF, sl.f, Sv);
        unit_normal(Sv[0], Sv[1], Sv[2], n, ds);
        sr += (fabs((w[1]) * n[0] + (w[2]) * n[1] + (w[3]) * n[2]) + (c_center)) * ds;
    }
    return sr;
}

__global__ void __launch_bounds__(kWG) k_spectral_radius(Range rg, DMesh m, DState s, double gamma) {
    const int c = rg.c0 + blockIdx.x * kWG + threadIdx.x;
    if (c >= rg.c1) return;
    const int nf = m.nf[c];
    if (nf == 0) return;
    double w[5];
    load_w(s.W, m.NA, c, w);
    s.specrad[c] = spectral_radius(m, c, nf, w, gamma);
}

__global__ void __launch_bounds__(kWG) k_timestep(Range rg, DMesh m, DState s, double cfl) {
    const int c = rg.c0 + blockIdx.x * kWG + threadIdx.x;
    if (c >= rg.c1 || m.nf[c] == 0) return;
    s.dt[c] = cfl * m.vol[c] / s.specrad[c];
}

// Green-Gauss (GreenGauss.cpp:86-166) or least-squares (LeastSquares.cpp:80-119)
// gradient of a real cell, into g[15] (v*3+d).
template <int GRAD>
__device__ __forceinline__ void cell_gradient(const DMesh& m, const double* W, int c, int nf, const double wc[5],
                                              double g[15]) {
#pragma unroll
    for (int i = 0; i < 15; ++i) g[i] = 0.0;
    for (int k = 0; k < nf; ++k) {
        const Slot sl = slot(m, k, c);
        double wn[5];
        load_w(W, m.NA, sl.nb, wn);
        if (GRAD == 0) {
            double n[3];
            load3(m.S, m.NF, sl.f, n);
            if (sl.right) {
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
            for (int d = 0; d < 3; ++d) wt[d] = m.lsqw[(size_t)(k * 3 + d) * m.NRp + c];
#pragma unroll
            for (int v = 0; v < 5; ++v)
#pragma unroll
                for (int d = 0; d < 3; ++d) g[v * 3 + d] += wt[d] * (wn[v] - wc[v]);
        }
    }
    if (GRAD == 0) {
        const double vol = m.vol[c];
#pragma unroll
        for (int i = 0; i < 15; ++i) g[i] /= vol;
    }
}

template <int GRAD>
__global__ void __launch_bounds__(kWG) k_gradient(Range rg, DMesh m, DState s) {
    const int c = rg.c0 + blockIdx.x * kWG + threadIdx.x;
    if (c >= rg.c1) return;
    const int nf = m.nf[c];
    if (nf == 0) return;
    double wc[5], g[15];
    load_w(s.W, m.NA, c, wc);
    cell_gradient<GRAD>(m, s.W, c, nf, wc, g);
#pragma unroll
    for (int i = 0; i < 15; ++i) s.grad[(size_t)i * m.NA + c] = g[i];
}

// Ghost gradient: reflection r = g - 2 (g.n) n of the interior gradient for the
// first n_real_boundaries markers (GreenGauss.cpp:175-234), zero otherwise
// (every gradient is zeroed first, GreenGauss.cpp:64-81).
__device__ __forceinline__ void ghost_gradient(const DMesh& m, int kind, int f, const double gi[15], double gg[15]) {
    if (kind & GK_REFLECT) {
        const double area = m.area[f];
        double n[3];
        load3(m.S, m.NF, f, n);
        n[0] /= area;
        n[1] /= area;
        n[2] /= area;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            double dot = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) dot += gi[v * 3 + d] * n[d];
#pragma unroll
            for (int d = 0; d < 3; ++d) gg[v * 3 + d] = gi[v * 3 + d] - 2.0 * dot * n[d];
        }
    } else {
#pragma unroll
        for (int i = 0; i < 15; ++i) gg[i] = 0.0;
    }
}

__global__ void __launch_bounds__(kWG) k_gradient_ghost(Range rg, DMesh m, DState s) {
    const int i = rg.g0 + blockIdx.x * kWG + threadIdx.x;
    if (i >= rg.g1) return;
    const int gcell = m.NRp + i, in = m.g_in[i], f = m.g_face[i];
    double gi[15], gg[15];
#pragma unroll
    for (int q = 0; q < 15; ++q) gi[q] = s.grad[(size_t)q * m.NA + in];
    ghost_gradient(m, m.g_kind[i], f, gi, gg);
#pragma unroll
    for (int q = 0; q < 15; ++q) s.grad[(size_t)q * m.NA + gcell] = gg[q];
}

// Limiter of one cell from its state, its gradient, its neighbour states and
// its own-side r vectors (Venkatakrishnan.cpp:87-324 / BarthJespersen.cpp:88-331).
template <bool VENKAT>
struct LimAcc {
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

template <bool VENKAT>
__global__ void __launch_bounds__(kWG) k_limiter(Range rg, DMesh m, DState s, double k3, double omega) {
    const int idx = blockIdx.x * kWG + threadIdx.x;
    const int nr = rg.c1 - rg.c0;
    if (idx >= nr + (rg.g1 - rg.g0)) return;
    const int c = idx < nr ? rg.c0 + idx : m.NRp + rg.g0 + (idx - nr);
    double g[15], w[5];
    LimAcc<VENKAT> L;
    load_w(s.W, m.NA, c, w);
    L.start(w);
#pragma unroll
    for (int q = 0; q < 15; ++q) g[q] = s.grad[(size_t)q * m.NA + c];
    const double eps2 = k3 * (m.vol[c]);
    if (c < m.NRp) {
        const int nf = m.nf[c];
        if (nf == 0) return;
        for (int k = 0; k < nf; ++k) {
            double wn[5];
            load_w(s.W, m.NA, slot(m, k, c).nb, wn);
            L.neighbour(wn);
        }
        for (int k = 0; k < nf; ++k) {
            const Slot sl = slot(m, k, c);
            double r[3];
            load3(sl.right ? m.rR : m.rL, m.NF, sl.f, r);
            L.face(g, r, eps2, omega);
        }
    } else {
        const int i = c - m.NRp;
        double wn[5], r[3];
        load_w(s.W, m.NA, m.g_in[i], wn);
        L.neighbour(wn);
        load3(m.rR, m.NF, m.g_face[i], r);  // a ghost is always the right cell of its face
        L.face(g, r, eps2, omega);
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) s.lim[(size_t)v * m.NA + c] = L.lim[v];
}

// Convective residual (FluxScheme.cpp:83-207 / AusmScheme.cpp:716-870) as a gather.
template <int FLUX>
__global__ void __launch_bounds__(kWG) k_flux_conv(Range rg, DMesh m, DState s, Gas gas) {
    const int c = rg.c0 + blockIdx.x * kWG + threadIdx.x;
    if (c >= rg.c1) return;
    const int nf = m.nf[c];
    if (nf == 0) return;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < nf; ++k) {
        const Slot sl = slot(m, k, c);
        const int L = sl.right ? sl.nb : c, R = sl.right ? c : sl.nb;
        double Sv[3], n[3], norm, r[3], wl[5], wr[5], F[5];
        load3(m.S, m.NF, sl.f, Sv);
        unit_normal(Sv[0], Sv[1], Sv[2], n, norm);
        load3(m.rL, m.NF, sl.f, r);
        recon_lim_grad(s, m.NA, L, r, wl);
        load3(m.rR, m.NF, sl.f, r);
        recon_lim_grad(s, m.NA, R, r, wr);
        if (FLUX == 0) central_flux(gas, wl, wr, n, norm, F);
        else ausm_conv(gas, wl, wr, n, norm, F);
        if (!sl.right) {
#pragma unroll
            for (int v = 0; v < 5; ++v) acc[v] += F[v];
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) acc[v] -= F[v];
        }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) s.conv[(size_t)v * m.NRp + c] = acc[v];
}

// Dissipative residual with the (1-beta) blend (RoeScheme.cpp:87-291 / AusmScheme.cpp:82-240).
template <int FLUX>
__global__ void __launch_bounds__(kWG) k_flux_diss(Range rg, DMesh m, DState s, Gas gas, double beta) {
    const int c = rg.c0 + blockIdx.x * kWG + threadIdx.x;
    if (c >= rg.c1) return;
    const int nf = m.nf[c];
    if (nf == 0) return;
    double acc[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) acc[v] = s.diss[(size_t)v * m.NRp + c] * (1.0 - beta);
    for (int k = 0; k < nf; ++k) {
        const Slot sl = slot(m, k, c);
        const int L = sl.right ? sl.nb : c, R = sl.right ? c : sl.nb;
        double Sv[3], n[3], norm, r[3], wl[5], wr[5], D[5];
        load3(m.S, m.NF, sl.f, Sv);
        unit_normal(Sv[0], Sv[1], Sv[2], n, norm);
        load3(m.rL, m.NF, sl.f, r);
        recon_lim_grad(s, m.NA, L, r, wl);
        load3(m.rR, m.NF, sl.f, r);
        recon_lim_grad(s, m.NA, R, r, wr);
        if (FLUX == 0) roe_dissipation(gas, wl, wr, n, norm, D);
        else ausm_diss(gas, wl, wr, n, norm, D);
        if (!sl.right) {
#pragma unroll
            for (int v = 0; v < 5; ++v) acc[v] += D[v] * beta;
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) acc[v] -= D[v] * beta;
        }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) s.diss[(size_t)v * m.NRp + c] = acc[v];
}

__global__ void __launch_bounds__(kWG) k_update(Range rg, DMesh m, DState s, double alpha, double gm1) {
    const int c = rg.c0 + blockIdx.x * kWG + threadIdx.x;
    if (c >= rg.c1 || m.nf[c] == 0) return;
    double w0[5], cv[5], ds[5], w[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        w0[v] = s.W0[(size_t)v * m.NRp + c];
        cv[v] = s.conv[(size_t)v * m.NRp + c];
        ds[v] = s.diss[(size_t)v * m.NRp + c];
    }
    rk_update(gm1, alpha * s.dt[c] / m.vol[c], w0, cv, ds, w);
#pragma unroll
    for (int v = 0; v < 5; ++v) s.W[(size_t)v * m.NA + c] = w[v];
}

// CentralIrsSmoother::smoothResidual (Solver/src/CentralIrsSmoother.cpp:12-183):
// one Jacobi sweep of central implicit residual smoothing (epsilon = 0.8) on
// conv and diss of real cells. Component q < 5 is conv variable q, q >= 5 diss
// variable q - 5. As in the reference, the w-dissipation Jacobi value lands in
// the v slot (:90) and the w slot is never set; it holds zero here.
constexpr double kIrsEpsilon = 0.8;

__device__ __forceinline__ double* irs_res(const DState& s, int NRp, int q) {
    return q < 5 ? s.conv + (size_t)q * NRp : s.diss + (size_t)(q - 5) * NRp;
}

__global__ void __launch_bounds__(kWG) k_irs_jacobi(Range rg, DMesh m, DState s) {
    const int c = rg.c0 + blockIdx.x * kWG + threadIdx.x;
    if (c >= rg.c1 || m.nf[c] == 0) return;
    const int na = m.nf[c];
    const size_t N = m.NRp;
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        if (q == 7) continue;
        const int dst = q == 8 ? 7 : q;
        s.jac[dst * N + c] = (irs_res(s, m.NRp, q)[c] + 0.) / (1. + kIrsEpsilon * na);
    }
    s.jac[8 * N + c] = 0.0;
}

__global__ void __launch_bounds__(kWG) k_irs_apply(Range rg, DMesh m, DState s) {
    const int c = rg.c0 + blockIdx.x * kWG + threadIdx.x;
    if (c >= rg.c1 || m.nf[c] == 0) return;
    const int na = m.nf[c];
    const size_t N = m.NRp;
    double sum[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) sum[q] = 0.;
    for (int k = 0; k < na; ++k) {
        const int nb = m.nbr[(size_t)k * N + c];
        if (nb < m.NRp)  // real neighbours only (ghosts are numbered after every real cell)
#pragma unroll
            for (int q = 0; q < 10; ++q) sum[q] += s.jac[q * N + nb];
    }
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        double* r = irs_res(s, m.NRp, q);
        r[c] = (r[c] + kIrsEpsilon * sum[q]) / (1. + kIrsEpsilon * na);
    }
}

__device__ __forceinline__ void apply_bc(const DMesh& m, int kind, int f, double gamma, const Free& fs,
                                         const double wi[5], double g[5]) {
    const double area = m.area[f];
    double n[3];
    load3(m.S, m.NF, f, n);
    n[0] /= area;
    n[1] /= area;
    n[2] /= area;
    if ((kind & GK_BC_MASK) == GK_SLIP) ghost_slip(wi, n, g);
    else ghost_farfield(gamma, wi, fs.v, n, g);
}

// BoundaryCellIds::updateBoundary of every physical marker (Updater.cpp:80-88).
__global__ void __launch_bounds__(kWG) k_bc(Range rg, DMesh m, DState s, double gamma, Free fs) {
    const int i = rg.g0 + blockIdx.x * kWG + threadIdx.x;
    if (i >= rg.g1) return;
    const int kind = m.g_kind[i];
    if ((kind & GK_BC_MASK) == 0) return;
    double wi[5], g[5];
    load_w(s.W, m.NA, m.g_in[i], wi);
    apply_bc(m, kind, m.g_face[i], gamma, fs, wi, g);
#pragma unroll
    for (int v = 0; v < 5; ++v) s.W[(size_t)v * m.NA + m.NRp + i] = g[v];
}

// BlockCommunicator::updateBoundaries (BlockCommunicator.cpp:212-268) on one
// device: origin cell -> destination interface ghost. Primitive state always;
// gradient and limiter too in the granular path (copy_all).
__global__ void __launch_bounds__(kWG) k_exchange(int n, const int* dst, const int* org, int NA, double* W,
                                                  double* grad, double* lim, int copy_all) {
    const int i = blockIdx.x * kWG + threadIdx.x;
    if (i >= n) return;
    const int d = dst[i], o = org[i];
#pragma unroll
    for (int v = 0; v < 5; ++v) W[(size_t)v * NA + d] = W[(size_t)v * NA + o];
    if (copy_all) {
#pragma unroll
        for (int v = 0; v < 5; ++v) lim[(size_t)v * NA + d] = lim[(size_t)v * NA + o];
#pragma unroll
        for (int q = 0; q < 15; ++q) grad[(size_t)q * NA + d] = grad[(size_t)q * NA + o];
    }
}

// Residual-norm partial sums from stored conv/diss (granular path), one per
// wavefront of 64 cells, as the last fused stage forms them.
__global__ void __launch_bounds__(kWG) k_rms_partials(DMesh m, DState s, int ntiles) {
    const int c = blockIdx.x * kWG + threadIdx.x;
    double res2[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (m.nf[c] != 0) {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            const double r = s.conv[(size_t)v * m.NRp + c] - s.diss[(size_t)v * m.NRp + c];
            res2[v] = r * r;
        }
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        double x = res2[v];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0) s.part[(size_t)(c >> 6) * 8 + v] = x;
    }
}

// Sum over a workgroup of five per-thread values: shuffle tree within each
// wavefront, then the wavefront sums in order. Valid in thread 0..4 (value v in
// thread v) after the call.
__device__ __forceinline__ double wg_sum5(const double a[5], double (*red)[5]) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
        double x = a[v];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][v] = x;
    }
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < 5)
#pragma unroll
        for (int w = 0; w < kWG / 64; ++w) t += red[w][threadIdx.x];
    __syncthreads();
    return t;
}

// Per-block monitors, first level: workgroup (b, p) of split workgroups per
// block sums the block's wall-pressure forces and moments
// (AerodynamicParameters.cpp:163-239) over wall faces p*kWG + t + j*split*kWG,
// and the squared-residual tile partials of the block (Convergence.cpp:27-144)
// over tiles in the same pattern; fixed trees, so the sums are deterministic.
__global__ void __launch_bounds__(kWG) k_monitor_partials(int split, const int* blk_tile0, const int* blk_ntile,
                                                          const int* wall_off, const int* wall_ghost,
                                                          const int* wall_face, const double* fcenter, int NF, DMesh m,
                                                          const double* W, const double* part,
                                                          const double* blk_qcmac, double* mpart) {
    const int b = blockIdx.x / split, p = blockIdx.x % split;
    const double qcmac = blk_qcmac[b];  // dynamic head x cmac of the block's instance
    __shared__ double red[kWG / 64][5];
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int q = wall_off[b] + p * kWG + threadIdx.x; q < wall_off[b + 1]; q += split * kWG) {
        const int gcell = wall_ghost[q], f = wall_face[q], in = m.g_in[gcell - m.NRp];
        const double ppbc = 0.5 * (W[(size_t)4 * m.NA + in] + W[(size_t)4 * m.NA + gcell]);
        const double cpbc = (ppbc - 1.) / qcmac;
        const double fx = cpbc * m.S[f], fy = cpbc * m.S[(size_t)NF + f], fz = cpbc * m.S[2 * (size_t)NF + f];
        const double x = fcenter[f], y = fcenter[(size_t)NF + f], z = fcenter[2 * (size_t)NF + f];
        a[0] += fy;
        a[1] += fx;
        a[2] += y * fz - z * fy;
        a[3] += z * fx - x * fz;
        a[4] += x * fy - y * fx;
    }
    double t = wg_sum5(a, red);
    if (threadIdx.x < 5) mpart[((size_t)b * split + p) * 10 + 5 + threadIdx.x] = t;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // tile q of kWG cells: its four wavefront partials in order
    for (int q = p * kWG + threadIdx.x; q < blk_ntile[b]; q += split * kWG) {
        const double* w = part + (size_t)(blk_tile0[b] + q) * (kWG / 64) * 8;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            double t = 0.0;
#pragma unroll
            for (int i = 0; i < kWG / 64; ++i) t += w[i * 8 + v];
            r[v] += t;
        }
    }
    t = wg_sum5(r, red);
    if (threadIdx.x < 5) mpart[((size_t)b * split + p) * 10 + threadIdx.x] = t;
}

// Second level: per block, the split partials in order; residual norm
// (1/N) sqrt(sum) (Convergence.cpp:35-42) and the force and moment sums.
__global__ void __launch_bounds__(64) k_block_monitors(int nb, int split, const int* blk_nreal, const double* mpart,
                                                       double* out) {
    const int b = blockIdx.x * 6 + threadIdx.x / 10, q = threadIdx.x % 10;
    if (threadIdx.x >= 60 || b >= nb) return;
    double t = 0.0;
    for (int p = 0; p < split; ++p) t += mpart[((size_t)b * split + p) * 10 + q];
    out[(size_t)b * 16 + q] = q < 5 ? (1. / blk_nreal[b]) * sqrt(t) : t;
}

// ------------------------------------------------------------ host launchers

#define GRID(n) dim3(((n) + kWG - 1) / kWG), dim3(kWG), 0, st
#define RGRID GRID(rg.c1 - rg.c0)

void launch_save_w0(const Range& rg, const DMesh& m, const DState& s, double inv_gm1, hipStream_t st) {
    if (rg.c1 > rg.c0) hipLaunchKernelGGL(k_save_w0, RGRID, rg, m, s, inv_gm1);
}
void launch_spectral_radius(const Range& rg, const DMesh& m, const DState& s, double gamma, hipStream_t st) {
    if (rg.c1 > rg.c0) hipLaunchKernelGGL(k_spectral_radius, RGRID, rg, m, s, gamma);
}
void launch_timestep(const Range& rg, const DMesh& m, const DState& s, double cfl, hipStream_t st) {
    if (rg.c1 > rg.c0) hipLaunchKernelGGL(k_timestep, RGRID, rg, m, s, cfl);
}
void launch_gradient(const Range& rg, const DMesh& m, const DState& s, int grad, hipStream_t st) {
    if (rg.c1 > rg.c0) {
        if (grad == 0) hipLaunchKernelGGL(k_gradient<0>, RGRID, rg, m, s);
        else hipLaunchKernelGGL(k_gradient<1>, RGRID, rg, m, s);
    }
    if (rg.g1 > rg.g0) hipLaunchKernelGGL(k_gradient_ghost, GRID(rg.g1 - rg.g0), rg, m, s);
}
void launch_limiter(const Range& rg, const DMesh& m, const DState& s, int venkat, double k3, double omega,
                    hipStream_t st) {
    const int n = (rg.c1 - rg.c0) + (rg.g1 - rg.g0);
    if (n == 0) return;
    if (venkat) hipLaunchKernelGGL(k_limiter<true>, GRID(n), rg, m, s, k3, omega);
    else hipLaunchKernelGGL(k_limiter<false>, GRID(n), rg, m, s, k3, omega);
}
void launch_flux_conv(const Range& rg, const DMesh& m, const DState& s, int flux, const Gas& g, hipStream_t st) {
    if (rg.c1 <= rg.c0) return;
    if (flux == 0) hipLaunchKernelGGL(k_flux_conv<0>, RGRID, rg, m, s, g);
    else hipLaunchKernelGGL(k_flux_conv<1>, RGRID, rg, m, s, g);
}
void launch_flux_diss(const Range& rg, const DMesh& m, const DState& s, int flux, const Gas& g, double beta,
                      hipStream_t st) {
    if (rg.c1 <= rg.c0) return;
    if (flux == 0) hipLaunchKernelGGL(k_flux_diss<0>, RGRID, rg, m, s, g, beta);
    else hipLaunchKernelGGL(k_flux_diss<1>, RGRID, rg, m, s, g, beta);
}
void launch_update(const Range& rg, const DMesh& m, const DState& s, double alpha, double gm1, hipStream_t st) {
    if (rg.c1 > rg.c0) hipLaunchKernelGGL(k_update, RGRID, rg, m, s, alpha, gm1);
}
void launch_smooth(const Range& rg, const DMesh& m, const DState& s, hipStream_t st) {
    if (rg.c1 <= rg.c0) return;
    hipLaunchKernelGGL(k_irs_jacobi, RGRID, rg, m, s);
    hipLaunchKernelGGL(k_irs_apply, RGRID, rg, m, s);
}
void launch_bc(const Range& rg, const DMesh& m, const DState& s, double gamma, const double fs[5], hipStream_t st) {
    if (rg.g1 <= rg.g0) return;
    Free f;
    for (int v = 0; v < 5; ++v) f.v[v] = fs[v];
    hipLaunchKernelGGL(k_bc, GRID(rg.g1 - rg.g0), rg, m, s, gamma, f);
}
void launch_exchange(int n, const int* dst, const int* org, int NA, double* W, double* grad, double* lim,
                     int copy_all, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_exchange, GRID(n), n, dst, org, NA, W, grad, lim, copy_all);
}
void launch_rms_partials(const DMesh& m, const DState& s, hipStream_t st) {
    const int nt = m.NRp / kWG;
    hipLaunchKernelGGL(k_rms_partials, dim3(nt), dim3(kWG), 0, st, m, s, nt);
}
void launch_block_monitors(int nb, int split, const int* tile0, const int* ntile, const int* nreal,
                           const int* wall_off, const int* wall_ghost, const int* wall_face, const double* fcenter,
                           const DMesh& m, const double* W, const double* part, const double* blk_qcmac,
                           double* mpart, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_monitor_partials, dim3(nb * split), dim3(kWG), 0, st, split, tile0, ntile, wall_off,
                       wall_ghost, wall_face, fcenter, m.NF, m, W, part, blk_qcmac, mpart);
    hipLaunchKernelGGL(k_block_monitors, dim3((nb + 5) / 6), dim3(64), 0, st, nb, split, nreal, mpart, out);
}

}  // namespace e3d
