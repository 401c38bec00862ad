// Face-flux, limiter and boundary physics shared by the granular and fused
// kernels. Every expression keeps the operand order of the reference statement
// it restates (file:line in each comment) and the file is compiled with
// -ffp-contract=off, so a cell-gather evaluation of a face reproduces the
// reference's face-loop value bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace e3d {

// x / y rounded to nearest, for many x over one y, from rv = 1.0 / y (one
// correctly rounded division per divisor) in five FMA-unit operations instead
// of a full division each: q0 = x * rv is within 1.5 ulp of x / y; one
// correction q1 = q0 + (x - q0 * y) * rv (remainder by FMA) makes it a
// faithful rounding, after which the remainder x - q1 * y is exact and the
// second correction q1 + (x - q1 * y) * rv, rounded once, is x / y rounded to
// nearest (P. Markstein, IBM J. Res. Dev. 34(1), 1990: rv within half an ulp
// of 1 / y and q1 faithful). Finite nonzero y; x = -0.0 gives +0.0, so it is
// used on sums that start from +0.0 (never -0.0).
__device__ __forceinline__ double div_by(double x, double y, double rv) {
    const double q0 = x * rv;
    const double q1 = fma(fma(-q0, y, x), rv, q0);
    return fma(fma(-q1, y, x), rv, q1);
}

struct Gas {
    double gamma;      // ratio of specific heats
    double g_over_gm1; // gamma/(gamma-1.0), the factor of H (FluxScheme.cpp:140)
    double gm1;        // gamma-1.0
};

// Face unit normal and |S| (FluxScheme.cpp:105-112).
__device__ __forceinline__ void unit_normal(double sx, double sy, double sz, double n[3], double& norm) {
    norm = sqrt(sx * sx + sy * sy + sz * sz);
    n[0] = sx / norm;
    n[1] = sy / norm;
    n[2] = sz / norm;
}

// Central convective flux of two reconstructed states (FluxScheme.cpp:139-182).
__device__ __forceinline__ void central_flux(const Gas& g, const double wl[5], const double wr[5],
                                             const double n[3], double norm, double F[5]) {
    const double qq_L = wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3];
    const double H_L = (0.5 * qq_L + g.g_over_gm1 * wl[4] / wl[0]);
    const double V_L = wl[1] * n[0] + wl[2] * n[1] + wl[3] * n[2];
    const double qq_R = wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3];
    const double H_R = (0.5 * qq_R + g.g_over_gm1 * wr[4] / wr[0]);
    const double V_R = wr[1] * n[0] + wr[2] * n[1] + wr[3] * n[2];
    F[0] = 0.5 * (wl[0] * V_L + wr[0] * V_R) * norm;
    F[1] = 0.5 * ((wl[0] * wl[1] * V_L + n[0] * wl[4]) + (wr[0] * wr[1] * V_R + n[0] * wr[4])) * norm;
    F[2] = 0.5 * ((wl[0] * wl[2] * V_L + n[1] * wl[4]) + (wr[0] * wr[2] * V_R + n[1] * wr[4])) * norm;
    F[3] = 0.5 * ((wl[0] * wl[3] * V_L + n[2] * wl[4]) + (wr[0] * wr[3] * V_R + n[2] * wr[4])) * norm;
    F[4] = 0.5 * (wl[0] * H_L * V_L + wr[0] * H_R * V_R) * norm;
}

// Roe dissipation with Harten's entropy fix (RoeScheme.cpp:183-254), before the
// beta factor. Returns 0.5*|A|dW*|S|.
__device__ __forceinline__ void roe_dissipation(const Gas& g, const double wl[5], const double wr[5],
                                                const double n[3], double norm, double D[5]) {
    const double rho_L = wl[0], u_L = wl[1], v_L = wl[2], w_L = wl[3], p_L = wl[4];
    const double rho_R = wr[0], u_R = wr[1], v_R = wr[2], w_R = wr[3], p_R = wr[4];
    const double H_L = (0.5 * (u_L * u_L + v_L * v_L + w_L * w_L) + g.g_over_gm1 * p_L / rho_L);
    const double V_L = u_L * n[0] + v_L * n[1] + w_L * n[2];
    const double H_R = (0.5 * (u_R * u_R + v_R * v_R + w_R * w_R) + g.g_over_gm1 * p_R / rho_R);
    const double V_R = u_R * n[0] + v_R * n[1] + w_R * n[2];
    const double delta_p = (p_R - p_L), delta_V = (V_R - V_L), delta_rho = (rho_R - rho_L);
    const double delta_u = (u_R - u_L), delta_v = (v_R - v_L), delta_w = (w_R - w_L);
    const double sL = sqrt(rho_L), sR = sqrt(rho_R);
    const double rho_wave = sqrt(rho_L * rho_R);
    const double u_wave = (u_L * sL + u_R * sR) / (sL + sR);
    const double v_wave = (v_L * sL + v_R * sR) / (sL + sR);
    const double w_wave = (w_L * sL + w_R * sR) / (sL + sR);
    const double H_wave = (H_L * sL + H_R * sR) / (sL + sR);
    const double q2 = u_wave * u_wave + v_wave * v_wave + w_wave * w_wave;
    const double c_wave = sqrt(g.gm1 * (H_wave - q2 / 2.));
    const double V_wave = u_wave * n[0] + v_wave * n[1] + w_wave * n[2];
    const double delta = 0.1 * c_wave;  // 1./10.*c_wave
    const double a1 = fabs(V_wave - c_wave), a2 = fabs(V_wave), a5 = fabs(V_wave + c_wave);
    const double L1 = a1 <= delta ? (a1 * a1 + delta * delta) / (2. * delta) : a1;
    const double L234 = a2 <= delta ? (a2 * a2 + delta * delta) / (2. * delta) : a2;
    const double L5 = a5 <= delta ? (a5 * a5 + delta * delta) / (2. * delta) : a5;
    const double two_c2 = 2. * c_wave * c_wave;
    const double F1 = L1 * (delta_p - rho_wave * c_wave * delta_V) / two_c2;
    const double F234 = (delta_rho - delta_p / (c_wave * c_wave));
    const double F5 = L5 * (delta_p + rho_wave * c_wave * delta_V) / two_c2;
    const double m234 = L234 * (F234 * 1. + rho_wave * 0.);
    const double u234 = L234 * ((F234 * u_wave) + rho_wave * (delta_u - delta_V * n[0]));
    const double v234 = L234 * ((F234 * v_wave) + rho_wave * (delta_v - delta_V * n[1]));
    const double w234 = L234 * ((F234 * w_wave) + rho_wave * (delta_w - delta_V * n[2]));
    const double e234 = L234 * ((F234 * q2 / 2.) +
                                rho_wave * (u_wave * delta_u + v_wave * delta_v + w_wave * delta_w - V_wave * delta_V));
    D[0] = 0.5 * (F1 * 1. + m234 + F5 * 1.) * norm;
    D[1] = 0.5 * (F1 * (u_wave - c_wave * n[0]) + u234 + F5 * (u_wave + c_wave * n[0])) * norm;
    D[2] = 0.5 * (F1 * (v_wave - c_wave * n[1]) + v234 + F5 * (v_wave + c_wave * n[1])) * norm;
    D[3] = 0.5 * (F1 * (w_wave - c_wave * n[2]) + w234 + F5 * (w_wave + c_wave * n[2])) * norm;
    D[4] = 0.5 * (F1 * (H_wave - c_wave * V_wave) + e234 + F5 * (H_wave + c_wave * V_wave)) * norm;
}

// AUSM split Mach numbers (AusmScheme.cpp:143-156, 182-195, 777-833).
__device__ __forceinline__ double mach_plus(double M) {
    return M <= -1.0 ? 0.0 : (M < 1.0 ? 0.25 * (M + 1.0) * (M + 1.0) : M);
}
__device__ __forceinline__ double mach_minus(double M) {
    return M <= -1.0 ? M : (M < 1.0 ? -0.25 * (M - 1.0) * (M - 1.0) : 0.0);
}

// AUSM convective flux (AusmScheme.cpp:770-854).
__device__ __forceinline__ void ausm_conv(const Gas& g, const double wl[5], const double wr[5], const double n[3],
                                          double norm, double F[5]) {
    const double rho_L = wl[0], u_L = wl[1], v_L = wl[2], w_L = wl[3], p_L = wl[4];
    const double rho_R = wr[0], u_R = wr[1], v_R = wr[2], w_R = wr[3], p_R = wr[4];
    const double H_L = (0.5 * (u_L * u_L + v_L * v_L + w_L * w_L) + g.g_over_gm1 * p_L / rho_L);
    const double V_L = u_L * n[0] + v_L * n[1] + w_L * n[2];
    const double c_L = sqrt(g.gamma * p_L / rho_L);
    const double M_L = V_L / c_L;
    double p_L_plus, M_L_plus;
    if (M_L <= -1.0) {
        p_L_plus = 0.0;
        M_L_plus = 0.0;
    } else if (M_L < 1.0) {
        p_L_plus = p_L * 0.25 * (M_L + 1.0) * (M_L + 1.0) * (2.0 - M_L);
        M_L_plus = 0.25 * (M_L + 1.0) * (M_L + 1.0);
    } else {
        M_L_plus = M_L;
        p_L_plus = p_L;
    }
    const double H_R = (0.5 * (u_R * u_R + v_R * v_R + w_R * w_R) + g.g_over_gm1 * p_R / rho_R);
    const double V_R = u_R * n[0] + v_R * n[1] + w_R * n[2];
    const double c_R = sqrt(g.gamma * p_R / rho_R);
    const double M_R = V_R / c_R;
    double p_R_minus, M_R_minus;
    if (M_R <= -1.0) {
        p_R_minus = p_R;
        M_R_minus = M_R;
    } else if (M_R < 1.0) {
        p_R_minus = p_R * 0.25 * (1.0 - M_R) * (1.0 - M_R) * (2.0 + M_R);
        M_R_minus = -0.25 * (M_R - 1.0) * (M_R - 1.0);
    } else {
        p_R_minus = 0.0;
        M_R_minus = 0.0;
    }
    const double M_L_R = M_L_plus + M_R_minus;
    const double pp = p_L_plus + p_R_minus;
    F[0] = 0.5 * M_L_R * (rho_L * c_L + rho_R * c_R) * norm;
    F[1] = 0.5 * M_L_R * (rho_L * u_L * c_L + rho_R * u_R * c_R) * norm + n[0] * pp * norm;
    F[2] = 0.5 * M_L_R * (rho_L * v_L * c_L + rho_R * v_R * c_R) * norm + n[1] * pp * norm;
    F[3] = 0.5 * M_L_R * (rho_L * w_L * c_L + rho_R * w_R * c_R) * norm + n[2] * pp * norm;
    F[4] = 0.5 * M_L_R * (rho_L * H_L * c_L + rho_R * H_R * c_R) * norm;
}

// AUSM dissipation before the beta factor (AusmScheme.cpp:138-227).
__device__ __forceinline__ void ausm_diss(const Gas& g, const double wl[5], const double wr[5], const double n[3],
                                          double norm, double D[5]) {
    const double rho_L = wl[0], u_L = wl[1], v_L = wl[2], w_L = wl[3], p_L = wl[4];
    const double rho_R = wr[0], u_R = wr[1], v_R = wr[2], w_R = wr[3], p_R = wr[4];
    const double H_L = (0.5 * (u_L * u_L + v_L * v_L + w_L * w_L) + g.g_over_gm1 * p_L / rho_L);
    const double V_L = u_L * n[0] + v_L * n[1] + w_L * n[2];
    const double c_L = sqrt(g.gamma * p_L / rho_L);
    const double H_R = (0.5 * (u_R * u_R + v_R * v_R + w_R * w_R) + g.g_over_gm1 * p_R / rho_R);
    const double V_R = u_R * n[0] + v_R * n[1] + w_R * n[2];
    const double c_R = sqrt(g.gamma * p_R / rho_R);
    double M = mach_plus(V_L / c_L) + mach_minus(V_R / c_R);
    const double delta = 0.25;
    if (!(fabs(M) > delta)) M = (M * M + delta * delta) / (2.0 * delta);
    const double a = 0.5 * fabs(M);
    D[0] = a * (-(rho_L * c_L) + rho_R * c_R) * norm;
    D[1] = a * (-(rho_L * u_L * c_L) + rho_R * u_R * c_R) * norm;
    D[2] = a * (-(rho_L * v_L * c_L) + rho_R * v_R * c_R) * norm;
    D[3] = a * (-(rho_L * w_L * c_L) + rho_R * w_R * c_R) * norm;
    D[4] = a * (-(rho_L * H_L * c_L) + rho_R * H_R * c_R) * norm;
}

// One face limiter value (Venkatakrishnan.cpp:227-238 / BarthJespersen.cpp:225-245).
// The reference's d2 > 0 and d2 < 0 branches differ only in the bound (dmax or
// dmin): the bound is selected first and the one expression evaluated, so a
// wavefront whose lanes differ in the sign of d2 does not run both, and every
// lane gets the value of its own branch (1 where d2 is 0 or NaN, as there).
// On the 4.7M-cell channel: Barth-Jespersen interp 0.369 -> 0.357 ms;
// Venkatakrishnan the same over the first 100 iterations (uniform signs) and
// 0.430 -> 0.395 ms by iteration 1,000, where the flow has developed and the
// signs mix within wavefronts.
template <bool VENKAT>
__device__ __forceinline__ double face_limiter(double d2, double dmax, double dmin, double eps2, double omega) {
    if (VENKAT) {
        const double dm = d2 > 0.0 ? dmax : dmin;
        const double phi = (1.0 / d2) * ((dm * dm + eps2) * d2 + 2.0 * d2 * d2 * dm) /
                           (dm * dm + 2.0 * d2 * d2 + dm * d2 + eps2);
        return (d2 > 0.0 || d2 < 0.0) ? phi : 1.0;
    } else {
        d2 = (copysign(1.0, d2)) * (fabs(d2) + omega);
        const double dm = d2 > 0.0 ? dmax : dmin;
        const double phi = fmin(1.0, dm / d2);
        return (d2 > 0.0 || d2 < 0.0) ? phi : 1.0;
    }
}

// Slip-wall / symmetry ghost from its interior state (WallCellIds.cpp:39-60).
__device__ __forceinline__ void ghost_slip(const double wi[5], const double n[3], double g[5]) {
    const double un1 = wi[1] * n[0] + wi[2] * n[1] + wi[3] * n[2];
    g[0] = wi[0];
    g[1] = 2.0 * (wi[1] - un1 * n[0]) - wi[1];
    g[2] = 2.0 * (wi[2] - un1 * n[1]) - wi[2];
    g[3] = 2.0 * (wi[3] - un1 * n[2]) - wi[3];
    g[4] = wi[4];
}

// Characteristic farfield ghost (FarfieldCellIds.cpp:55-113).
__device__ __forceinline__ void ghost_farfield(double gamma, const double wi[5], const double fs[5],
                                               const double n[3], double g[5]) {
    const double ri = wi[0], ui = wi[1], vi = wi[2], wv = wi[3], pi = wi[4];
    const double rf = fs[0], uf = fs[1], vf = fs[2], wf = fs[3], pf = fs[4];
    const double cc = sqrt(gamma * pi / ri);
    const double un1 = ui * n[0] + vi * n[1] + wv * n[2];
    if (fabs(un1 / cc) > 1.0) {
        if (un1 < 0.0) {
            g[0] = 2.0 * rf - ri; g[1] = 2.0 * uf - ui; g[2] = 2.0 * vf - vi; g[3] = 2.0 * wf - wv; g[4] = 2.0 * pf - pi;
        } else {
            g[0] = ri; g[1] = ui; g[2] = vi; g[3] = wv; g[4] = pi;
        }
    } else if (un1 < 0.0) {
        const double pb = 0.5 * (pf + pi - ri * cc * (n[0] * (uf - ui) + n[1] * (vf - vi) + n[2] * (wf - wv)));
        const double rb = rf + (pb - pf) / (cc * cc);
        const double ub = uf - n[0] * (pf - pb) / (ri * cc);
        const double vb = vf - n[1] * (pf - pb) / (ri * cc);
        const double wb = wf - n[2] * (pf - pb) / (ri * cc);
        g[0] = 2.0 * rb - ri; g[1] = 2.0 * ub - ui; g[2] = 2.0 * vb - vi; g[3] = 2.0 * wb - wv; g[4] = 2.0 * pb - pi;
    } else {
        const double pb = pf;
        const double rb = ri + (pb - pi) / (cc * cc);
        const double ub = ui + n[0] * (pi - pb) / (ri * cc);
        const double vb = vi + n[1] * (pi - pb) / (ri * cc);
        const double wb = wv + n[2] * (pi - pb) / (ri * cc);
        g[0] = 2.0 * rb - ri; g[1] = 2.0 * ub - ui; g[2] = 2.0 * vb - vi; g[3] = 2.0 * wb - wv; g[4] = 2.0 * pb - pi;
    }
}

// Conservative update and back to primitives (Updater.cpp:54-64).
__device__ __forceinline__ void rk_update(double gm1, double coef, const double w0[5], const double conv[5],
                                          const double diss[5], double w[5]) {
    double nw[5];
    for (int v = 0; v < 5; ++v) nw[v] = w0[v] - coef * (conv[v] - diss[v]);
    w[0] = nw[0];
    w[1] = nw[1] / nw[0];
    w[2] = nw[2] / nw[0];
    w[3] = nw[3] / nw[0];
    w[4] = gm1 * (nw[4] - 0.5 * (nw[1] * nw[1] + nw[2] * nw[2] + nw[3] * nw[3]) / nw[0]);
}

}  // namespace e3d
