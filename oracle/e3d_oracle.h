/*
 * CPU restatement of euler3D's residual-and-update path -- TEST INFRASTRUCTURE.
 *
 * This is the oracle: the checker that tests, the smoke check and the CPU
 * baseline leg of the benchmark compare the product against. The product
 * (libeuler3d_amd.so) never calls it. It follows the reference statement by
 * statement in the reference's own loop form (global face loops scattering to
 * both cells, cell loops in index order), in plain C compiled with
 * -ffp-contract=off, so that on the reference's inputs it reproduces the
 * reference's outputs bit for bit. Each function cites the reference lines it
 * restates. Arrays are per block, in the reference's numbering, variable-major:
 * W[v*nall + c] (v = rho,u,v,w,p), grad[(v*3+d)*nall + c].
 */
#ifndef E3D_ORACLE_H
#define E3D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct e3do_block {
    int nreal, nall, nface;
    const int* face2cell;                              /* 2F */
    const double *S, *area, *fcenter, *rL, *rR;        /* 3F, F, 3F, 3F, 3F */
    const int *cell_off, *cell2face, *cell2cell;       /* CSR over nall */
    const double *volume, *lsqw;                       /* nall, 3*cell_off[nreal] */
    int n_bnd, n_real_bnd;
    const int *bnd_type, *bnd_off, *bnd_ghosts;        /* per marker: 0 wall 1 farfield 2 symmetry 3 connexion */
    int n_wall;
    const int* wall_faces;
    /* state, owned by the caller */
    double *W, *conv, *diss, *W0, *grad, *lim;         /* 5nall, 5nall, 5nall, 5nall, 15nall, 5nall */
    double *dt, *specrad;                              /* nreal */
} e3do_block;

typedef struct e3do_params {
    double gamma, cfl, aoa_deg, mach, cmac, omega, k, criterion;
    int stages, order, flux, gradient, limiter, max_iter;   /* flux 0 roe 1 ausm; gradient 0 GG 1 LSQ; limiter 0 BJ 1 VK */
    int smoothing;                                          /* 1: CentralIrsSmoother after each residual */
} e3do_params;

typedef struct e3do_interface {        /* one CONNEXION boundary: ghosts of dst filled from origin cells */
    int dst_block, origin_block, n;
    const int *dst_cells, *origin_cells;
} e3do_interface;

typedef struct e3do_monitor {
    int iter, stop;
    double rms[5], drop[5], cl, cd, cmx, cmy, cmz;
} e3do_monitor;

void e3do_freestream(const e3do_params* p, double fs[5]);
void e3do_init_flow(e3do_block* b, const double fs[5]);
void e3do_save_w0(e3do_block* b, double gamma);
void e3do_spectral_radius(e3do_block* b, double gamma);
void e3do_timestep(e3do_block* b, double cfl);
void e3do_green_gauss(e3do_block* b);
void e3do_least_squares(e3do_block* b);
void e3do_venkatakrishnan(e3do_block* b, double k);
void e3do_barth_jespersen(e3do_block* b, double omega);
void e3do_flux_conv_central(e3do_block* b, double gamma);
void e3do_flux_diss_roe(e3do_block* b, double gamma, double beta);
void e3do_flux_conv_ausm(e3do_block* b, double gamma);
void e3do_flux_diss_ausm(e3do_block* b, double gamma, double beta);
void e3do_update(e3do_block* b, double gamma, double alpha);
void e3do_bc(e3do_block* b, const double fs[5], double gamma);
void e3do_rk_coefficients(int stages, double alpha[5], double beta[5]);
void e3do_smooth_residual(e3do_block* b);
void e3do_residual(e3do_block* b, const e3do_params* p, int stage);
void e3do_runge_kutta(e3do_block* b, const e3do_params* p, const double fs[5]);
void e3do_convergence(const e3do_block* b, double rms[5]);
void e3do_aerodynamic(const e3do_block* b, const e3do_params* p, double out[5]);
void e3do_exchange(e3do_block* blocks, const e3do_interface* itf, int n_itf);
/* Runs Solver::solve for at most n iterations (continuing the iteration count
 * in *it and the iteration-0 reference in rms0); one monitor row per iteration.
 * Returns the number of iterations run. */
int e3do_iterate(e3do_block* blocks, int nb, const e3do_interface* itf, int n_itf, const e3do_params* p,
                 int n, int* it, double rms0[5], e3do_monitor* history);
/* The same with the blocks of each iteration spread over nthreads threads
 * (contiguous block ranges); identical results for any nthreads. */
int e3do_iterate_mt(e3do_block* blocks, int nb, const e3do_interface* itf, int n_itf, const e3do_params* p,
                    int n, int* it, double rms0[5], e3do_monitor* history, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
