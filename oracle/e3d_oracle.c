/*
 * CPU restatement of euler3D's residual-and-update path -- TEST INFRASTRUCTURE
 * (the oracle). See e3d_oracle.h. Every expression keeps the operand order of
 * the reference statement it restates; build with -ffp-contract=off.
 */
#include "e3d_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#define NV 5

static inline double* var(double* a, int nall, int v) { return a + (size_t)v * nall; }
static inline double* gcomp(double* g, int nall, int v, int d) { return g + (size_t)(v * 3 + d) * nall; }

/* PrimitiveVariables::calculateFreeVariables (DataStructure/src/PrimitiveVariables.cpp:41-56) */
void e3do_freestream(const e3do_params* p, double fs[5]) {
    const double pi = atan(1.0) * 4.0;
    const double alpha_rad = p->aoa_deg * pi / 180.0;
    const double c = cos(alpha_rad), s = sin(alpha_rad);
    fs[0] = 1.0;
    fs[1] = p->mach * sqrt(p->gamma) * c;
    fs[2] = p->mach * sqrt(p->gamma) * s;
    fs[3] = 0.0;
    fs[4] = 1.0;
}

/* PrimitiveVariables::initializeFlowField (:58-69): every cell, ghosts included */
void e3do_init_flow(e3do_block* b, const double fs[5]) {
    for (int i = 0; i < b->nall; ++i)
        for (int v = 0; v < NV; ++v) var(b->W, b->nall, v)[i] = fs[v];
}

/* Solver::saveW0 (Solver/src/Solver.cpp:161-210) */
void e3do_save_w0(e3do_block* b, double gamma) {
    const int na = b->nall;
    for (int c = 0; c < b->nreal; ++c) {
        const double ro = b->W[c], uu = b->W[na + c], vv = b->W[2 * na + c], ww = b->W[3 * na + c],
                     pp = b->W[4 * na + c];
        b->W0[c] = ro;
        b->W0[na + c] = ro * uu;
        b->W0[2 * na + c] = ro * vv;
        b->W0[3 * na + c] = ro * ww;
        b->W0[4 * na + c] = 0.5 * ro * (uu * uu + vv * vv + ww * ww) + 1. / (gamma - 1.0) * pp;
    }
}

/* Timestep::computeSpectralRadius (Solver/src/Timestep.cpp:26-121): own-cell state only */
void e3do_spectral_radius(e3do_block* b, double gamma) {
    const int na = b->nall;
    for (int c = 0; c < b->nreal; ++c) {
        const double c_center = sqrt(gamma * b->W[4 * na + c] / b->W[c]);
        b->specrad[c] = 0.0;
        for (int k = b->cell_off[c]; k < b->cell_off[c + 1]; ++k) {
            const int f = b->cell2face[k];
            double nx = b->S[3 * f], ny = b->S[3 * f + 1], nz = b->S[3 * f + 2];
            const double ds = sqrt(nx * nx + ny * ny + nz * nz);
            nx /= ds;
            ny /= ds;
            nz /= ds;
            b->specrad[c] += (fabs((b->W[na + c]) * nx + (b->W[2 * na + c]) * ny + (b->W[3 * na + c]) * nz) + (c_center)) * ds;
        }
    }
}

/* Timestep::computeTimestep (:123-148) */
void e3do_timestep(e3do_block* b, double cfl) {
    for (int c = 0; c < b->nreal; ++c) b->dt[c] = cfl * b->volume[c] / b->specrad[c];
}

/* Ghost-gradient reflection over the first n_real_bnd markers
 * (Solver/src/GreenGauss.cpp:175-234, Solver/src/LeastSquares.cpp:127-186). */
static void reflect_ghost_gradients(e3do_block* b) {
    const int na = b->nall;
    for (int j = 0; j < b->n_real_bnd; ++j)
        for (int q = b->bnd_off[j]; q < b->bnd_off[j + 1]; ++q) {
            const int ext = b->bnd_ghosts[q];
            const int f = b->cell2face[b->cell_off[ext]];
            const int in = b->face2cell[2 * f];
            const double area = b->area[f];
            double n[3];
            for (int d = 0; d < 3; ++d) {
                n[d] = b->S[3 * f + d];
                n[d] /= area;
            }
            for (int v = 0; v < NV; ++v) {
                double dot = 0.0;
                for (int d = 0; d < 3; ++d) dot += gcomp(b->grad, na, v, d)[in] * n[d];
                for (int d = 0; d < 3; ++d)
                    gcomp(b->grad, na, v, d)[ext] = gcomp(b->grad, na, v, d)[in] - 2.0 * dot * n[d];
            }
        }
}

/* GreenGauss::computeGradients (Solver/src/GreenGauss.cpp:12-234) */
void e3do_green_gauss(e3do_block* b) {
    const int na = b->nall;
    for (size_t i = 0; i < (size_t)15 * na; ++i) b->grad[i] = 0.0;
    for (int c = 0; c < b->nreal; ++c) {
        for (int k = b->cell_off[c]; k < b->cell_off[c + 1]; ++k) {
            const int f = b->cell2face[k];
            double n[3];
            for (int d = 0; d < 3; ++d) n[d] = b->S[3 * f + d];
            if (c != b->face2cell[2 * f])
                for (int d = 0; d < 3; ++d) n[d] *= -1.0;
            const int L = b->face2cell[2 * f], R = b->face2cell[2 * f + 1];
            for (int v = 0; v < NV; ++v) {
                const double wl = var(b->W, na, v)[L], wr = var(b->W, na, v)[R];
                for (int d = 0; d < 3; ++d) gcomp(b->grad, na, v, d)[c] += 0.5 * (wl + wr) * n[d];
            }
        }
        const double vol = b->volume[c];
        for (int v = 0; v < NV; ++v)
            for (int d = 0; d < 3; ++d) gcomp(b->grad, na, v, d)[c] /= vol;
    }
    reflect_ghost_gradients(b);
}

/* LeastSquares::computeGradients (Solver/src/LeastSquares.cpp:12-186) */
void e3do_least_squares(e3do_block* b) {
    const int na = b->nall;
    for (size_t i = 0; i < (size_t)15 * na; ++i) b->grad[i] = 0.0;
    for (int c = 0; c < b->nreal; ++c)
        for (int k = b->cell_off[c]; k < b->cell_off[c + 1]; ++k) {
            const int o = b->cell2cell[k];
            const int j = k - b->cell_off[c];
            const double* w = b->lsqw + 3 * (size_t)(b->cell_off[c] + j);
            for (int v = 0; v < NV; ++v) {
                const double p0 = var(b->W, na, v)[c], pi = var(b->W, na, v)[o];
                for (int d = 0; d < 3; ++d) gcomp(b->grad, na, v, d)[c] += w[d] * (pi - p0);
            }
        }
    reflect_ghost_gradients(b);
}

/* Shared neighbour scan of both limiters: Delta_max / Delta_min over cell->cells
 * (Venkatakrishnan.cpp:87-177, BarthJespersen.cpp:88-177). */
static void delta_minmax(const e3do_block* b, int c, double dmax[5], double dmin[5]) {
    const int na = b->nall;
    for (int v = 0; v < NV; ++v) {
        const double own = b->W[(size_t)v * na + c];
        double mx = own, mn = own;
        for (int k = b->cell_off[c]; k < b->cell_off[c + 1]; ++k) {
            const double nb = b->W[(size_t)v * na + b->cell2cell[k]];
            if (nb > mx) mx = nb;
            if (nb < mn) mn = nb;
        }
        dmax[v] = mx - own;
        dmin[v] = mn - own;
    }
}

/* Venkatakrishnan::computeLimiters (Solver/src/Venkatakrishnan.cpp:12-326): all cells, uncapped */
void e3do_venkatakrishnan(e3do_block* b, double k) {
    const int na = b->nall;
    for (int c = 0; c < na; ++c) {
        const double eps2 = k * k * k * (b->volume[c]);
        double dmax[5], dmin[5];
        delta_minmax(b, c, dmax, dmin);
        for (int v = 0; v < NV; ++v) var(b->lim, na, v)[c] = INFINITY;
        for (int q = b->cell_off[c]; q < b->cell_off[c + 1]; ++q) {
            const int f = b->cell2face[q];
            const double* r = (c == b->face2cell[2 * f]) ? b->rL + 3 * (size_t)f : b->rR + 3 * (size_t)f;
            for (int v = 0; v < NV; ++v) {
                double d2 = 0.0;
                for (int d = 0; d < 3; ++d) d2 += gcomp(b->grad, na, v, d)[c] * r[d];
                double phi;
                if (d2 > 0.0)
                    phi = (1.0 / d2) * ((dmax[v] * dmax[v] + eps2) * d2 + 2.0 * d2 * d2 * dmax[v]) /
                          (dmax[v] * dmax[v] + 2.0 * d2 * d2 + dmax[v] * d2 + eps2);
                else if (d2 < 0.0)
                    phi = (1.0 / d2) * ((dmin[v] * dmin[v] + eps2) * d2 + 2.0 * d2 * d2 * dmin[v]) /
                          (dmin[v] * dmin[v] + 2.0 * d2 * d2 + dmin[v] * d2 + eps2);
                else
                    phi = 1.0;
                if (phi < var(b->lim, na, v)[c]) var(b->lim, na, v)[c] = phi;
            }
        }
    }
}

/* BarthJespersen::computeLimiters (Solver/src/BarthJespersen.cpp:14-333) */
void e3do_barth_jespersen(e3do_block* b, double omega) {
    const int na = b->nall;
    for (int c = 0; c < na; ++c) {
        double dmax[5], dmin[5];
        delta_minmax(b, c, dmax, dmin);
        for (int v = 0; v < NV; ++v) var(b->lim, na, v)[c] = INFINITY;
        for (int q = b->cell_off[c]; q < b->cell_off[c + 1]; ++q) {
            const int f = b->cell2face[q];
            const double* r = (c == b->face2cell[2 * f]) ? b->rL + 3 * (size_t)f : b->rR + 3 * (size_t)f;
            for (int v = 0; v < NV; ++v) {
                double d2 = 0.0;
                for (int d = 0; d < 3; ++d) d2 += gcomp(b->grad, na, v, d)[c] * r[d];
                d2 = (copysign(1.0, d2)) * (fabs(d2) + omega);
                double phi;
                if (d2 > 0.0) phi = fmin(1.0, dmax[v] / d2);
                else if (d2 < 0.0) phi = fmin(1.0, dmin[v] / d2);
                else phi = 1.0;
                if (phi < var(b->lim, na, v)[c]) var(b->lim, na, v)[c] = phi;
            }
        }
    }
}

/* Reconstructed face state of cell c with r vector r (FluxScheme.cpp:118-134). */
static void reconstruct(const e3do_block* b, int c, const double* r, double out[5]) {
    const int na = b->nall;
    for (int v = 0; v < NV; ++v) out[v] = b->W[(size_t)v * na + c];
    for (int d = 0; d < 3; ++d)
        for (int v = 0; v < NV; ++v)
            out[v] += b->lim[(size_t)v * na + c] * b->grad[(size_t)(v * 3 + d) * na + c] * r[d];
}

static void face_normal(const e3do_block* b, int f, double n[3], double* norm) {
    n[0] = b->S[3 * f];
    n[1] = b->S[3 * f + 1];
    n[2] = b->S[3 * f + 2];
    *norm = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= *norm;
    n[1] /= *norm;
    n[2] /= *norm;
}

/* FluxScheme::computeFluxConv, central flux used by Roe (Solver/src/FluxScheme.cpp:16-211) */
void e3do_flux_conv_central(e3do_block* b, double gamma) {
    const int na = b->nall;
    for (int c = 0; c < b->nreal; ++c)
        for (int v = 0; v < NV; ++v) var(b->conv, na, v)[c] = 0.0;
    for (int f = 0; f < b->nface; ++f) {
        double n[3], norm, wl[5], wr[5];
        face_normal(b, f, n, &norm);
        const int L = b->face2cell[2 * f], R = b->face2cell[2 * f + 1];
        reconstruct(b, L, b->rL + 3 * (size_t)f, wl);
        const double qq_L = wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3];
        const double H_L = (0.5 * qq_L + gamma / (gamma - 1.0) * wl[4] / wl[0]);
        const double V_L = wl[1] * n[0] + wl[2] * n[1] + wl[3] * n[2];
        reconstruct(b, R, b->rR + 3 * (size_t)f, wr);
        const double qq_R = wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3];
        const double H_R = (0.5 * qq_R + gamma / (gamma - 1.0) * wr[4] / wr[0]);
        const double V_R = wr[1] * n[0] + wr[2] * n[1] + wr[3] * n[2];
        double FL[5], FR[5], F[5];
        FL[0] = wl[0] * V_L;
        FR[0] = wr[0] * V_R;
        FL[1] = wl[0] * wl[1] * V_L + n[0] * wl[4];
        FR[1] = wr[0] * wr[1] * V_R + n[0] * wr[4];
        FL[2] = wl[0] * wl[2] * V_L + n[1] * wl[4];
        FR[2] = wr[0] * wr[2] * V_R + n[1] * wr[4];
        FL[3] = wl[0] * wl[3] * V_L + n[2] * wl[4];
        FR[3] = wr[0] * wr[3] * V_R + n[2] * wr[4];
        FL[4] = wl[0] * H_L * V_L;
        FR[4] = wr[0] * H_R * V_R;
        for (int v = 0; v < NV; ++v) F[v] = 0.5 * (FL[v] + FR[v]) * norm;
        for (int v = 0; v < NV; ++v) var(b->conv, na, v)[L] += F[v];
        for (int v = 0; v < NV; ++v) var(b->conv, na, v)[R] -= F[v];
    }
}

/* RoeScheme::computeFluxDiss (Solver/src/RoeScheme.cpp:12-311) */
void e3do_flux_diss_roe(e3do_block* b, double gamma, double beta) {
    const int na = b->nall;
    for (int c = 0; c < b->nreal; ++c)
        for (int v = 0; v < NV; ++v) var(b->diss, na, v)[c] *= (1.0 - beta);
    for (int f = 0; f < b->nface; ++f) {
        double n[3], norm, wl[5], wr[5];
        face_normal(b, f, n, &norm);
        const int L = b->face2cell[2 * f], R = b->face2cell[2 * f + 1];
        reconstruct(b, L, b->rL + 3 * (size_t)f, wl);
        const double rho_L = wl[0], u_L = wl[1], v_L = wl[2], w_L = wl[3], p_L = wl[4];
        const double qq_L = u_L * u_L + v_L * v_L + w_L * w_L;
        const double H_L = (0.5 * qq_L + gamma / (gamma - 1.0) * p_L / rho_L);
        const double V_L = u_L * n[0] + v_L * n[1] + w_L * n[2];
        reconstruct(b, R, b->rR + 3 * (size_t)f, wr);
        const double rho_R = wr[0], u_R = wr[1], v_R = wr[2], w_R = wr[3], p_R = wr[4];
        const double qq_R = u_R * u_R + v_R * v_R + w_R * w_R;
        const double H_R = (0.5 * qq_R + gamma / (gamma - 1.0) * p_R / rho_R);
        const double V_R = u_R * n[0] + v_R * n[1] + w_R * n[2];

        const double delta_p = (p_R - p_L), delta_V = (V_R - V_L), delta_rho = (rho_R - rho_L);
        const double delta_u = (u_R - u_L), delta_v = (v_R - v_L), delta_w = (w_R - w_L);
        const double rho_wave = sqrt(rho_L * rho_R);
        const double u_wave = (u_L * sqrt(rho_L) + u_R * sqrt(rho_R)) / (sqrt(rho_L) + sqrt(rho_R));
        const double v_wave = (v_L * sqrt(rho_L) + v_R * sqrt(rho_R)) / (sqrt(rho_L) + sqrt(rho_R));
        const double w_wave = (w_L * sqrt(rho_L) + w_R * sqrt(rho_R)) / (sqrt(rho_L) + sqrt(rho_R));
        const double H_wave = (H_L * sqrt(rho_L) + H_R * sqrt(rho_R)) / (sqrt(rho_L) + sqrt(rho_R));
        const double q_square_wave = u_wave * u_wave + v_wave * v_wave + w_wave * w_wave;
        const double c_wave = sqrt((gamma - 1.) * (H_wave - q_square_wave / 2.));
        const double V_wave = u_wave * n[0] + v_wave * n[1] + w_wave * n[2];

        const double delta = 1. / 10. * c_wave;  /* Harten entropy fix (:193-217) */
        double L1, L234, L5;
        if (fabs(V_wave - c_wave) <= delta)
            L1 = (fabs(V_wave - c_wave) * fabs(V_wave - c_wave) + delta * delta) / (2. * delta);
        else
            L1 = fabs(V_wave - c_wave);
        if (fabs(V_wave) <= delta)
            L234 = (fabs(V_wave) * fabs(V_wave) + delta * delta) / (2. * delta);
        else
            L234 = fabs(V_wave);
        if (fabs(V_wave + c_wave) <= delta)
            L5 = (fabs(V_wave + c_wave) * fabs(V_wave + c_wave) + delta * delta) / (2. * delta);
        else
            L5 = fabs(V_wave + c_wave);

        const double F_1_cst = L1 * (delta_p - rho_wave * c_wave * delta_V) / (2. * c_wave * c_wave);
        const double F_1_mass = F_1_cst * 1.;
        const double F_1_u = F_1_cst * (u_wave - c_wave * n[0]);
        const double F_1_v = F_1_cst * (v_wave - c_wave * n[1]);
        const double F_1_w = F_1_cst * (w_wave - c_wave * n[2]);
        const double F_1_energy = F_1_cst * (H_wave - c_wave * V_wave);
        const double F_234_cst = (delta_rho - delta_p / (c_wave * c_wave));
        const double F_234_mass = L234 * (F_234_cst * 1. + rho_wave * 0.);
        const double F_234_u = L234 * ((F_234_cst * u_wave) + rho_wave * (delta_u - delta_V * n[0]));
        const double F_234_v = L234 * ((F_234_cst * v_wave) + rho_wave * (delta_v - delta_V * n[1]));
        const double F_234_w = L234 * ((F_234_cst * w_wave) + rho_wave * (delta_w - delta_V * n[2]));
        const double F_234_energy =
            L234 * ((F_234_cst * q_square_wave / 2.) +
                    rho_wave * (u_wave * delta_u + v_wave * delta_v + w_wave * delta_w - V_wave * delta_V));
        const double F_5_cst = L5 * (delta_p + rho_wave * c_wave * delta_V) / (2. * c_wave * c_wave);
        const double F_5_mass = F_5_cst * 1.;
        const double F_5_u = F_5_cst * (u_wave + c_wave * n[0]);
        const double F_5_v = F_5_cst * (v_wave + c_wave * n[1]);
        const double F_5_w = F_5_cst * (w_wave + c_wave * n[2]);
        const double F_5_energy = F_5_cst * (H_wave + c_wave * V_wave);

        double D[5];
        D[0] = 0.5 * (F_1_mass + F_234_mass + F_5_mass) * norm;
        D[1] = 0.5 * (F_1_u + F_234_u + F_5_u) * norm;
        D[2] = 0.5 * (F_1_v + F_234_v + F_5_v) * norm;
        D[3] = 0.5 * (F_1_w + F_234_w + F_5_w) * norm;
        D[4] = 0.5 * (F_1_energy + F_234_energy + F_5_energy) * norm;
        for (int v = 0; v < NV; ++v) var(b->diss, na, v)[L] += D[v] * beta;
        for (int v = 0; v < NV; ++v) var(b->diss, na, v)[R] -= D[v] * beta;
    }
}

/* AusmScheme::computeFluxConv (Solver/src/AusmScheme.cpp:647-871) */
void e3do_flux_conv_ausm(e3do_block* b, double gamma) {
    const int na = b->nall;
    for (int c = 0; c < b->nreal; ++c)
        for (int v = 0; v < NV; ++v) var(b->conv, na, v)[c] = 0.0;
    for (int f = 0; f < b->nface; ++f) {
        double n[3], norm, wl[5], wr[5];
        face_normal(b, f, n, &norm);
        const int L = b->face2cell[2 * f], R = b->face2cell[2 * f + 1];
        reconstruct(b, L, b->rL + 3 * (size_t)f, wl);
        const double rho_L = wl[0], u_L = wl[1], v_L = wl[2], w_L = wl[3], p_L = wl[4];
        const double qq_L = u_L * u_L + v_L * v_L + w_L * w_L;
        const double H_L = (0.5 * qq_L + gamma / (gamma - 1.0) * p_L / rho_L);
        const double V_L = u_L * n[0] + v_L * n[1] + w_L * n[2];
        const double c_L = sqrt(gamma * p_L / rho_L);
        const double M_L = V_L / c_L;
        double p_L_plus, M_L_plus;
        if (M_L <= -1.0) {
            p_L_plus = 0;
            M_L_plus = 0;
        } else if (M_L < 1.0) {
            p_L_plus = p_L * 0.25 * (M_L + 1.0) * (M_L + 1.0) * (2.0 - M_L);
            M_L_plus = 0.25 * (M_L + 1.0) * (M_L + 1.0);
        } else {
            M_L_plus = M_L;
            p_L_plus = p_L;
        }
        reconstruct(b, R, b->rR + 3 * (size_t)f, wr);
        const double rho_R = wr[0], u_R = wr[1], v_R = wr[2], w_R = wr[3], p_R = wr[4];
        const double qq_R = u_R * u_R + v_R * v_R + w_R * w_R;
        const double H_R = (0.5 * qq_R + gamma / (gamma - 1.0) * p_R / rho_R);
        const double V_R = u_R * n[0] + v_R * n[1] + w_R * n[2];
        const double c_R = sqrt(gamma * p_R / rho_R);
        const double M_R = V_R / c_R;
        double p_R_minus, M_R_minus;
        if (M_R <= -1.0) {
            p_R_minus = p_R;
            M_R_minus = M_R;
        } else if (M_R < 1.0) {
            p_R_minus = p_R * 0.25 * (1.0 - M_R) * (1.0 - M_R) * (2.0 + M_R);
            M_R_minus = -0.25 * (M_R - 1.0) * (M_R - 1.0);
        } else {
            p_R_minus = 0;
            M_R_minus = 0;
        }
        const double M_L_R = M_L_plus + M_R_minus;
        double FL[5], FR[5], F[5];
        FL[0] = rho_L * c_L;
        FR[0] = rho_R * c_R;
        FL[1] = rho_L * u_L * c_L;
        FR[1] = rho_R * u_R * c_R;
        FL[2] = rho_L * v_L * c_L;
        FR[2] = rho_R * v_R * c_R;
        FL[3] = rho_L * w_L * c_L;
        FR[3] = rho_R * w_R * c_R;
        FL[4] = rho_L * H_L * c_L;
        FR[4] = rho_R * H_R * c_R;
        F[0] = 0.5 * M_L_R * (FL[0] + FR[0]) * norm;
        F[1] = 0.5 * M_L_R * (FL[1] + FR[1]) * norm + n[0] * (p_L_plus + p_R_minus) * norm;
        F[2] = 0.5 * M_L_R * (FL[2] + FR[2]) * norm + n[1] * (p_L_plus + p_R_minus) * norm;
        F[3] = 0.5 * M_L_R * (FL[3] + FR[3]) * norm + n[2] * (p_L_plus + p_R_minus) * norm;
        F[4] = 0.5 * M_L_R * (FL[4] + FR[4]) * norm;
        for (int v = 0; v < NV; ++v) var(b->conv, na, v)[L] += F[v];
        for (int v = 0; v < NV; ++v) var(b->conv, na, v)[R] -= F[v];
    }
}

/* AusmScheme::computeFluxDiss, active part (Solver/src/AusmScheme.cpp:12-243) */
void e3do_flux_diss_ausm(e3do_block* b, double gamma, double beta) {
    const int na = b->nall;
    const double delta = 0.25;
    for (int c = 0; c < b->nreal; ++c)
        for (int v = 0; v < NV; ++v) var(b->diss, na, v)[c] *= (1.0 - beta);
    for (int f = 0; f < b->nface; ++f) {
        double n[3], norm, wl[5], wr[5];
        face_normal(b, f, n, &norm);
        const int L = b->face2cell[2 * f], R = b->face2cell[2 * f + 1];
        reconstruct(b, L, b->rL + 3 * (size_t)f, wl);
        const double rho_L = wl[0], u_L = wl[1], v_L = wl[2], w_L = wl[3], p_L = wl[4];
        const double qq_L = u_L * u_L + v_L * v_L + w_L * w_L;
        const double H_L = (0.5 * qq_L + gamma / (gamma - 1.0) * p_L / rho_L);
        const double V_L = u_L * n[0] + v_L * n[1] + w_L * n[2];
        const double c_L = sqrt(gamma * p_L / rho_L);
        const double M_L = V_L / c_L;
        double M_L_plus;
        if (M_L <= -1.0) M_L_plus = 0;
        else if (M_L < 1.0) M_L_plus = 0.25 * (M_L + 1.0) * (M_L + 1.0);
        else M_L_plus = M_L;
        reconstruct(b, R, b->rR + 3 * (size_t)f, wr);
        const double rho_R = wr[0], u_R = wr[1], v_R = wr[2], w_R = wr[3], p_R = wr[4];
        const double qq_R = u_R * u_R + v_R * v_R + w_R * w_R;
        const double H_R = (0.5 * qq_R + gamma / (gamma - 1.0) * p_R / rho_R);
        const double V_R = u_R * n[0] + v_R * n[1] + w_R * n[2];
        const double c_R = sqrt(gamma * p_R / rho_R);
        const double M_R = V_R / c_R;
        double M_R_minus;
        if (M_R <= -1.0) M_R_minus = M_R;
        else if (M_R < 1.0) M_R_minus = -0.25 * (M_R - 1.0) * (M_R - 1.0);
        else M_R_minus = 0;
        double M_L_R = M_L_plus + M_R_minus;
        if (fabs(M_L_R) > delta) {
        } else {
            M_L_R = (M_L_R * M_L_R + delta * delta) / (2.0 * delta);
        }
        double FL[5], FR[5], D[5];
        FL[0] = rho_L * c_L;
        FR[0] = rho_R * c_R;
        FL[1] = rho_L * u_L * c_L;
        FR[1] = rho_R * u_R * c_R;
        FL[2] = rho_L * v_L * c_L;
        FR[2] = rho_R * v_R * c_R;
        FL[3] = rho_L * w_L * c_L;
        FR[3] = rho_R * w_R * c_R;
        FL[4] = rho_L * H_L * c_L;
        FR[4] = rho_R * H_R * c_R;
        for (int v = 0; v < NV; ++v) D[v] = 0.5 * fabs(M_L_R) * (-FL[v] + FR[v]) * norm;
        for (int v = 0; v < NV; ++v) var(b->diss, na, v)[L] += D[v] * beta;
        for (int v = 0; v < NV; ++v) var(b->diss, na, v)[R] -= D[v] * beta;
    }
}

/* Updater::updateInternalBlock (Solver/src/Updater.cpp:12-78) */
void e3do_update(e3do_block* b, double gamma, double alpha) {
    const int na = b->nall;
    for (int c = 0; c < b->nreal; ++c) {
        const double dt = b->dt[c], vol = b->volume[c];
        double nw[5];
        for (int v = 0; v < NV; ++v)
            nw[v] = var(b->W0, na, v)[c] - (alpha * dt / vol) * (var(b->conv, na, v)[c] - var(b->diss, na, v)[c]);
        b->W[c] = nw[0];
        b->W[na + c] = nw[1] / nw[0];
        b->W[2 * na + c] = nw[2] / nw[0];
        b->W[3 * na + c] = nw[3] / nw[0];
        b->W[4 * na + c] = (gamma - 1.0) * (nw[4] - 0.5 * (nw[1] * nw[1] + nw[2] * nw[2] + nw[3] * nw[3]) / nw[0]);
    }
}

/* WallCellIds / SymmetryCellIds::updateBoundary (DataStructure/src/WallCellIds.cpp:9-95,
 * SymmetryCellIds.cpp:9-96: identical reflection) */
static void bc_slip(e3do_block* b, int ext) {
    const int na = b->nall;
    const int f = b->cell2face[b->cell_off[ext]];
    const int in = b->cell2cell[b->cell_off[ext]];
    double nx = b->S[3 * f], ny = b->S[3 * f + 1], nz = b->S[3 * f + 2];
    const double area = b->area[f];
    nx /= area;
    ny /= area;
    nz /= area;
    const double ro = b->W[in], uu = b->W[na + in], vv = b->W[2 * na + in], ww = b->W[3 * na + in],
                 pp = b->W[4 * na + in];
    const double un1 = uu * nx + vv * ny + ww * nz;
    const double ub = uu - un1 * nx, vb = vv - un1 * ny, wb = ww - un1 * nz;
    b->W[ext] = ro;
    b->W[na + ext] = 2.0 * ub - uu;
    b->W[2 * na + ext] = 2.0 * vb - vv;
    b->W[3 * na + ext] = 2.0 * wb - ww;
    b->W[4 * na + ext] = pp;
}

/* FarfieldCellIds::updateBoundary (DataStructure/src/FarfieldCellIds.cpp:10-114) */
static void bc_farfield(e3do_block* b, int ext, const double fs[5], double gamma) {
    const int na = b->nall;
    const int f = b->cell2face[b->cell_off[ext]];
    const int in = b->cell2cell[b->cell_off[ext]];
    double nx = b->S[3 * f], ny = b->S[3 * f + 1], nz = b->S[3 * f + 2];
    const double area = b->area[f];
    nx /= area;
    ny /= area;
    nz /= area;
    const double ri = b->W[in], ui = b->W[na + in], vi = b->W[2 * na + in], wi = b->W[3 * na + in],
                 pi = b->W[4 * na + in];
    const double rf = fs[0], uf = fs[1], vf = fs[2], wf = fs[3], pf = fs[4];
    const double cc = sqrt(gamma * pi / ri);
    const double un1 = ui * nx + vi * ny + wi * nz;
    double g[5];
    if (fabs(un1 / cc) > 1.0) {
        if (un1 < 0.0) {
            g[0] = 2.0 * rf - ri; g[1] = 2.0 * uf - ui; g[2] = 2.0 * vf - vi; g[3] = 2.0 * wf - wi; g[4] = 2.0 * pf - pi;
        } else {
            g[0] = ri; g[1] = ui; g[2] = vi; g[3] = wi; g[4] = pi;
        }
    } else if (un1 < 0.0) {
        const double pb = 0.5 * (pf + pi - ri * cc * (nx * (uf - ui) + ny * (vf - vi) + nz * (wf - wi)));
        const double rb = rf + (pb - pf) / (cc * cc);
        const double ub = uf - nx * (pf - pb) / (ri * cc);
        const double vb = vf - ny * (pf - pb) / (ri * cc);
        const double wb = wf - nz * (pf - pb) / (ri * cc);
        g[0] = 2.0 * rb - ri; g[1] = 2.0 * ub - ui; g[2] = 2.0 * vb - vi; g[3] = 2.0 * wb - wi; g[4] = 2.0 * pb - pi;
    } else {
        const double pb = pf;
        const double rb = ri + (pb - pi) / (cc * cc);
        const double ub = ui + nx * (pi - pb) / (ri * cc);
        const double vb = vi + ny * (pi - pb) / (ri * cc);
        const double wb = wi + nz * (pi - pb) / (ri * cc);
        g[0] = 2.0 * rb - ri; g[1] = 2.0 * ub - ui; g[2] = 2.0 * vb - vi; g[3] = 2.0 * wb - wi; g[4] = 2.0 * pb - pi;
    }
    for (int v = 0; v < NV; ++v) b->W[(size_t)v * na + ext] = g[v];
}

/* Updater::updateBoundary (Solver/src/Updater.cpp:80-88): the first n_real_bnd markers */
void e3do_bc(e3do_block* b, const double fs[5], double gamma) {
    for (int j = 0; j < b->n_real_bnd; ++j)
        for (int q = b->bnd_off[j]; q < b->bnd_off[j + 1]; ++q) {
            const int ext = b->bnd_ghosts[q];
            switch (b->bnd_type[j]) {
                case 0:
                case 2: bc_slip(b, ext); break;
                case 1: bc_farfield(b, ext, fs, gamma); break;
                default: break;  /* ConnexionBlockCellIds::updateBoundary is empty */
            }
        }
}

/* RungeKutta constructor tables (Solver/src/RungeKutta.cpp:23-68) */
void e3do_rk_coefficients(int stages, double a[5], double bt[5]) {
    static const double A[5][5] = {{1.0}, {0.6, 1.0}, {0.6, 0.6, 1.0}, {0.25, 1.0 / 3.0, 0.5, 1.0},
                                   {0.2742, 0.2067, 0.5020, 0.5142, 1.0}};
    static const double B[5][5] = {{1.0}, {1.0, 0.0}, {1.0, 0.0, 0.0}, {1.0, 0.5, 0.0, 0.0},
                                   {1.0, 0.0, 0.56, 0.0, 0.44}};
    for (int i = 0; i < 5; ++i) {
        a[i] = (stages >= 1 && stages <= 5) ? A[stages - 1][i] : 0.0;
        bt[i] = (stages >= 1 && stages <= 5) ? B[stages - 1][i] : 0.0;
    }
}

/* ResidualCalculator::computeResidual (Solver/src/ResidualCalculator.cpp:24-41); the
 * dissipation always runs because `10^-16` is an integer XOR (= -6). */
/* CentralIrsSmoother::smoothResidual (Solver/src/CentralIrsSmoother.cpp:12-183):
 * one Jacobi sweep of central implicit residual smoothing, epsilon = 0.8, over
 * real cells and their real neighbours, on conv and diss. Kept as the
 * reference computes it: the dissipative Jacobi value of w is stored into the
 * v array (:90), so the v sweep reads w values, and the w array is never
 * written; it is read here as zero (a freshly zeroed allocation). */
void e3do_smooth_residual(e3do_block* b) {
    const int n = b->nreal;
    const double epsilon = 0.8;
    double* jac = (double*)calloc((size_t)10 * n, sizeof(double));
    double* out = (double*)malloc(sizeof(double) * 10 * (size_t)n);
    double* res[10];
    for (int v = 0; v < 5; ++v) {
        res[v] = b->conv + (size_t)v * b->nall;
        res[5 + v] = b->diss + (size_t)v * b->nall;
    }
    for (int c = 0; c < n; ++c) {
        const int na = b->cell_off[c + 1] - b->cell_off[c];
        for (int q = 0; q < 10; ++q) {
            const int dst = (q == 8) ? 7 : q;          /* w diss -> v diss slot (:90) */
            if (q == 7) continue;                      /* v diss: overwritten by w diss */
            jac[(size_t)dst * n + c] = (res[q][c] + 0.) / (1. + epsilon * na);
        }
    }
    for (int c = 0; c < n; ++c) {
        double sum[10] = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
        const int na = b->cell_off[c + 1] - b->cell_off[c];
        for (int j = b->cell_off[c]; j < b->cell_off[c + 1]; ++j) {
            const int nb = b->cell2cell[j];
            if (nb < n)
                for (int q = 0; q < 10; ++q) sum[q] += jac[(size_t)q * n + nb];
        }
        for (int q = 0; q < 10; ++q) out[(size_t)q * n + c] = (res[q][c] + epsilon * sum[q]) / (1. + epsilon * na);
    }
    for (int q = 0; q < 10; ++q)
        for (int c = 0; c < n; ++c) res[q][c] = out[(size_t)q * n + c];
    free(jac);
    free(out);
}

void e3do_residual(e3do_block* b, const e3do_params* p, int stage) {
    double a[5], bt[5];
    e3do_rk_coefficients(p->stages, a, bt);
    if (p->order == 2) {
        if (p->gradient == 1) e3do_least_squares(b);
        else e3do_green_gauss(b);
        if (p->limiter == 1) e3do_venkatakrishnan(b, p->k);
        else e3do_barth_jespersen(b, p->omega);
    }
    if (p->flux == 1) {
        e3do_flux_conv_ausm(b, p->gamma);
        e3do_flux_diss_ausm(b, p->gamma, bt[stage]);
    } else {
        e3do_flux_conv_central(b, p->gamma);
        e3do_flux_diss_roe(b, p->gamma, bt[stage]);
    }
    if (p->smoothing) e3do_smooth_residual(b);
}

/* RungeKutta::computeRungeKutta (Solver/src/RungeKutta.cpp:80-94) */
void e3do_runge_kutta(e3do_block* b, const e3do_params* p, const double fs[5]) {
    double a[5], bt[5];
    e3do_rk_coefficients(p->stages, a, bt);
    for (int s = 0; s < p->stages; ++s) {
        e3do_residual(b, p, s);
        e3do_update(b, p->gamma, a[s]);
        e3do_bc(b, fs, p->gamma);
    }
}

/* Convergence::computeConvergence (PostProcessing/src/Convergence.cpp:27-144) */
void e3do_convergence(const e3do_block* b, double rms[5]) {
    const int na = b->nall;
    for (int v = 0; v < NV; ++v) {
        double s = 0.0;
        for (int i = 0; i < b->nreal; ++i) {
            const double r = b->conv[(size_t)v * na + i] - b->diss[(size_t)v * na + i];
            s += r * r;
        }
        rms[v] = (1. / b->nreal) * pow(s, 0.5);
    }
}

/* AerodynamicParameters::computeAerodynamic (PostProcessing/src/AerodynamicParameters.cpp:163-239):
 * out = {cl_global, cd_global, cmx, cmy, cmz} of the block. */
void e3do_aerodynamic(const e3do_block* b, const e3do_params* p, double out[5]) {
    const int na = b->nall;
    const double aoa_rad = p->aoa_deg * 4. * atan(1.) / 180;
    const double dyn_head = 0.5 * p->gamma * p->mach * p->mach;
    double cl = 0., cd = 0., cmx = 0., cmy = 0., cmz = 0.;
    for (int i = 0; i < b->n_wall; ++i) {
        const int f = b->wall_faces[i];
        const int c0 = b->face2cell[2 * f], c1 = b->face2cell[2 * f + 1];
        const double x = b->fcenter[3 * f], y = b->fcenter[3 * f + 1], z = b->fcenter[3 * f + 2];
        const double ppbc = 0.5 * (b->W[4 * na + c0] + b->W[4 * na + c1]);
        const double cpbc = (ppbc - 1.) / (dyn_head * p->cmac);
        const double fx = cpbc * b->S[3 * f], fy = cpbc * b->S[3 * f + 1], fz = cpbc * b->S[3 * f + 2];
        cl += fy;
        cd += fx;
        cmx += y * fz - z * fy;
        cmy += z * fx - x * fz;
        cmz += x * fy - y * fx;
    }
    out[0] = cl * cos(aoa_rad) - cd * sin(aoa_rad);
    out[1] = cl * sin(aoa_rad) + cd * cos(aoa_rad);
    out[2] = cmx;
    out[3] = cmy;
    out[4] = cmz;
}

/* BlockCommunicator::updateBoundaries, serial branch (MPI/src/BlockCommunicator.cpp:212-268):
 * all 40 per-cell values of every origin cell into its destination ghost. */
void e3do_exchange(e3do_block* blocks, const e3do_interface* itf, int n_itf) {
    for (int i = 0; i < n_itf; ++i) {
        e3do_block* o = &blocks[itf[i].origin_block];
        e3do_block* d = &blocks[itf[i].dst_block];
        for (int j = 0; j < itf[i].n; ++j) {
            const int co = itf[i].origin_cells[j], cd = itf[i].dst_cells[j];
            for (int v = 0; v < NV; ++v) {
                d->W[(size_t)v * d->nall + cd] = o->W[(size_t)v * o->nall + co];
                d->conv[(size_t)v * d->nall + cd] = o->conv[(size_t)v * o->nall + co];
                d->diss[(size_t)v * d->nall + cd] = o->diss[(size_t)v * o->nall + co];
                d->W0[(size_t)v * d->nall + cd] = o->W0[(size_t)v * o->nall + co];
                d->lim[(size_t)v * d->nall + cd] = o->lim[(size_t)v * o->nall + co];
                for (int dd = 0; dd < 3; ++dd)
                    d->grad[(size_t)(v * 3 + dd) * d->nall + cd] = o->grad[(size_t)(v * 3 + dd) * o->nall + co];
            }
        }
    }
}

/* Solver::solve (Solver/src/Solver.cpp:36-98) with PostProcessing::computeFlowData / process
 * (PostProcessing/src/PostProcessing.cpp:198-308) and the serial getGlobal sum
 * (MPI/src/BlockCommunicator.cpp:341-369). */
/* Solver::solve (Solver/src/Solver.cpp:36-98). With nthreads > 1 the blocks of
 * an iteration are split into contiguous ranges, one per thread, as the
 * reference's MPI build gives each rank a contiguous block range
 * (MPI/src/BlockCommunicator.cpp:284-294); the exchange and the sums over
 * blocks stay sequential, so the result does not depend on nthreads. */
int e3do_iterate_mt(e3do_block* blocks, int nb, const e3do_interface* itf, int n_itf, const e3do_params* p,
                    int n, int* it, double rms0[5], e3do_monitor* history, int nthreads) {
    double fs[5];
    e3do_freestream(p, fs);
    if (*it == 0) e3do_exchange(blocks, itf, n_itf);
    int done = 0;
    double* brms = (double*)__builtin_alloca(sizeof(double) * 10 * (size_t)nb);
    if (nthreads < 1) nthreads = 1;
    for (int k = 0; k < n; ++k) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int i = 0; i < nb; ++i) {
            e3do_block* b = &blocks[i];
            e3do_save_w0(b, p->gamma);
            e3do_spectral_radius(b, p->gamma);
            e3do_timestep(b, p->cfl);
            e3do_runge_kutta(b, p, fs);
            e3do_convergence(b, brms + 10 * i);
            e3do_aerodynamic(b, p, brms + 10 * i + 5);
        }
        e3do_exchange(blocks, itf, n_itf);
        double sum[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < nb; ++i)
            for (int q = 0; q < 10; ++q) sum[q] += brms[10 * i + q];
        if (*it == 0)
            for (int v = 0; v < 5; ++v) rms0[v] = sum[v];
        e3do_monitor m;
        m.iter = *it;
        for (int v = 0; v < 5; ++v) {
            m.rms[v] = sum[v];
            m.drop[v] = log10(sum[v]) - log10(rms0[v]);
        }
        m.cl = sum[5];
        m.cd = sum[6];
        m.cmx = sum[7];
        m.cmy = sum[8];
        m.cmz = sum[9];
        m.stop = (*it + 1 == p->max_iter) || (m.drop[0] <= p->criterion);
        if (history) history[done] = m;
        ++done;
        ++*it;
        if (m.stop) break;
    }
    return done;
}

int e3do_iterate(e3do_block* blocks, int nb, const e3do_interface* itf, int n_itf, const e3do_params* p,
                 int n, int* it, double rms0[5], e3do_monitor* history) {
    return e3do_iterate_mt(blocks, nb, itf, n_itf, p, n, it, rms0, history, 1);
}
