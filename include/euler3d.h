/*
 * euler3d_amd -- C ABI of the MI355X residual-and-update path of euler3D.
 *
 * One e3d_ctx owns the device copy of a multi-block mesh and its flow state,
 * one HIP stream and the solver parameters. Calls on a context are ordered on
 * its stream; getters synchronise. Every function returns E3D_OK (0) or a
 * negative status and never exits or throws; e3d_last_error() gives the text.
 *
 * Cell numbering inside a block is the reference's: [0, nreal) real cells,
 * [nreal, nall) ghost cells in marker order (physical markers first, CONNEXION
 * last), faces numbered by first appearance. Arrays passed across the ABI are
 * structure-of-arrays in that numbering: variable-major, e.g. W[v*nall + c]
 * with v = rho, u, v, w, p; gradients G[(v*3 + d)*nall + c].
 */
#ifndef EULER3D_AMD_H
#define EULER3D_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define E3D_OK 0
#define E3D_EINVAL (-1)      /* bad argument or unsupported option */
#define E3D_ENOMEM (-2)      /* host or device allocation failed */
#define E3D_EIO (-3)         /* file missing or malformed */
#define E3D_EHIP (-4)        /* HIP runtime error (no device, launch failure) */
#define E3D_ESTATE (-5)      /* call out of order (no mesh, no params) */
#define E3D_ENONFINITE (-6)  /* residual norm became NaN or Inf */

/* Scheme selectors: the reference picks them by string in its factories
 * (Solver/src/ResidualCalculator.cpp:72-95, Solver/src/SecondOrder.cpp:50-61). */
#define E3D_FLUX_ROE 0
#define E3D_FLUX_AUSM 1
#define E3D_GRAD_GREEN_GAUSS 0
#define E3D_GRAD_LEAST_SQUARES 1
#define E3D_LIM_BARTH_JESPERSEN 0
#define E3D_LIM_VENKATAKRISHNAN 1

/* Array selectors for e3d_get_array / e3d_set_array. */
#define E3D_ARR_W 0        /* primitive rho,u,v,w,p      5 x nall  */
#define E3D_ARR_W0 1       /* conservative state at iteration start 5 x nreal */
#define E3D_ARR_CONV 2     /* convective residual        5 x nreal */
#define E3D_ARR_DISS 3     /* dissipative residual       5 x nreal */
#define E3D_ARR_GRAD 4     /* gradients                 15 x nall  */
#define E3D_ARR_LIM 5      /* limiters                   5 x nall  */
#define E3D_ARR_DT 6       /* local time step            1 x nreal */
#define E3D_ARR_SPECRAD 7  /* convective spectral radius 1 x nreal */

typedef struct e3d_ctx e3d_ctx;
typedef struct e3d_mesh e3d_mesh;

/* Bump channel [0,3]x[0,1]x[0,1], circular-arc bump of thickness h on the
 * lower wall for 1 <= x <= 2, nx*ny*nz hexes split into 6 tets each.
 * WALL = lower wall, SYMMETRY = top and sides, FARFIELD = inlet and outlet.
 * Cells are numbered brick by brick (brick^3 hexes, 0 = lexicographic;
 * brick < 0: tets in a pseudo-random order seeded by -brick, a badly numbered
 * mesh for memory-order studies);
 * px*py*pz > 1 splits the hexes into that many blocks with CONNEXION markers. */
typedef struct e3d_bump_desc {
    int nx, ny, nz;
    double h;
    int brick;
    int px, py, pz;
} e3d_bump_desc;

/* Read-only view of one built block, in the reference's numbering. Pointers
 * stay valid until e3d_mesh_free. */
typedef struct e3d_block_view {
    int nreal, nall, nface, nnode;
    int n_boundaries, n_real_boundaries, n_wall_faces;
    const int *face2cell, *face_nnode, *face_nodes;      /* 2F, F, 4F        */
    const double *S, *area, *fcenter, *rL, *rR;          /* 3F, F, 3F, 3F, 3F */
    const int *cell_off, *cell2face, *cell2cell;         /* nall+1, CSR      */
    const double *volume, *ccenter, *lsqw;               /* nall, 3nall, 3*cell_off[nreal] */
    const int *bnd_type, *bnd_off, *bnd_ghosts;          /* per marker, CSR of ghost ids */
    const int *wall_faces;
    int ghost_count;                                     /* topology GhostCount */
    int n_interfaces;
    const int *itf_origin_block, *itf_off, *itf_origin_cells;
    const int *cell_node_off, *cell_nodes;               /* nall+1 CSR: element nodes in file order */
    const double* node_xyz;                              /* 3*nnode, x y z per node */
} e3d_block_view;

/* Solver parameters: the numeric content of the reference's input file
 * (Solver/src/Interface.cpp:33-198) and of its factory choices. */
typedef struct e3d_params {
    double gamma, cfl, aoa_deg, mach, cmac;
    int stages;           /* RK stages 1..5, coefficient tables of Solver/src/RungeKutta.cpp:23-68 */
    int max_iter;
    double criterion;     /* stop when log10 density residual drop <= criterion */
    int smoothing;        /* 0 NoSmoother, 1 CentralIrsSmoother (Solver/src/ResidualCalculator.cpp:90-95) */
    int flux;             /* E3D_FLUX_* */
    int order;            /* 1 or 2 */
    int gradient;         /* E3D_GRAD_* */
    int limiter;          /* E3D_LIM_* */
    double omega, k;      /* Barth-Jespersen omega, Venkatakrishnan K */
} e3d_params;

/* One row of the monitor history (PostProcessing/src/PostProcessing.cpp:220-308). */
typedef struct e3d_monitor {
    int iter;
    int stop;             /* 1 when the reference's stop test fires */
    double rms[5];        /* sum over blocks of (1/N_b) * sqrt(sum_c (conv-diss)^2) */
    double drop[5];       /* log10(rms) - log10(rms at iteration 0) */
    double cl, cd, cmx, cmy, cmz;
} e3d_monitor;

/* ---------------------------------------------------------------- errors */
const char* e3d_last_error(const e3d_ctx* ctx);   /* ctx may be NULL: last error of this thread */
const char* e3d_version(void);

/* ------------------------------------------------------------ host mesh  */
/* Replaces CompleteMesh::InitializeMyBlocks (DataStructure/src/CompleteMesh.cpp:79-293):
 * reads the topology file and every block file, builds faces and connectivity
 * (BlockBuilder/src/ConcreteBlockBuilder.cpp:22-742, BlockBuilder/src/BlockBuilder.cpp:147-458)
 * and geometry (Metrics/src/MetricsInitializer.cpp:27-800). */
int e3d_mesh_read(const char* topology_path, e3d_mesh** out);
/* Generates the bump channel in memory and builds it like e3d_mesh_read. */
int e3d_mesh_bump(const e3d_bump_desc* desc, e3d_mesh** out);
/* Writes <dir>/<stem>_<b>.su2 per block and <dir>/<stem>.top (reference dialect). */
int e3d_mesh_write(const e3d_mesh* mesh, const char* dir, const char* stem);
int e3d_mesh_nblocks(const e3d_mesh* mesh);
int e3d_mesh_block(const e3d_mesh* mesh, int block, e3d_block_view* out);
void e3d_mesh_free(e3d_mesh* mesh);

/* Parses the reference's positional input file (Solver/src/Interface.cpp:14-237).
 * topology_path receives the topology file name (buffer of path_len bytes). */
int e3d_read_input(const char* input_path, e3d_params* params, char* topology_path, int path_len);

/* ------------------------------------------------------- device context */
int e3d_create(int device, e3d_ctx** out);
void e3d_destroy(e3d_ctx* ctx);
/* Replaces the Solver/RungeKutta/ResidualCalculator constructors
 * (Solver/src/Solver.cpp:16-22, Solver/src/RungeKutta.cpp:15-72). */
int e3d_set_params(e3d_ctx* ctx, const e3d_params* params);
/* Copies every block of the mesh and its interfaces into device memory. */
int e3d_upload_mesh(e3d_ctx* ctx, const e3d_mesh* mesh);
int e3d_nblocks(const e3d_ctx* ctx);
/* PrimitiveVariables::calculateFreeVariables + initializeFlowField
 * (DataStructure/src/PrimitiveVariables.cpp:41-69): freestream in every cell. */
int e3d_init_freestream(e3d_ctx* ctx);
/* Host <-> device copies of one block's arrays (E3D_ARR_*), reference layout. */
int e3d_get_array(e3d_ctx* ctx, int block, int which, double* out);
int e3d_set_array(e3d_ctx* ctx, int block, int which, const double* in);

/* ----------------------------------------------- granular operators
 * block = -1 applies the operator to every block (one launch). */
/* Solver::saveW0 (Solver/src/Solver.cpp:161-210) */
int e3d_save_w0(e3d_ctx* ctx, int block);
/* Timestep::computeSpectralRadius / computeTimestep (Solver/src/Timestep.cpp:26-148) */
int e3d_spectral_radius(e3d_ctx* ctx, int block);
int e3d_timestep(e3d_ctx* ctx, int block);
/* Gradient::computeGradients (Solver/head/Gradient.h:15): kind E3D_GRAD_GREEN_GAUSS
 * (GreenGauss.cpp:12-234) or E3D_GRAD_LEAST_SQUARES (LeastSquares.cpp:12-186) */
int e3d_gradients(e3d_ctx* ctx, int block, int kind);
/* Limiter::computeLimiters (Solver/head/Limiter.h:15): kind E3D_LIM_VENKATAKRISHNAN with
 * param = K (Venkatakrishnan.cpp:12-326) or E3D_LIM_BARTH_JESPERSEN with param = omega
 * (BarthJespersen.cpp:14-333) */
int e3d_limiters(e3d_ctx* ctx, int block, int kind, double param);
/* Interpolation::computeInterpolation (Solver/head/Interpolation.h:15): FirstOrder no-op, SecondOrder gradient+limiter */
int e3d_interpolation(e3d_ctx* ctx, int block);
/* FluxScheme::computeFluxConv (Solver/head/FluxScheme.h:20): scheme E3D_FLUX_ROE (the central
 * flux of FluxScheme.cpp:16-211) or E3D_FLUX_AUSM (AusmScheme.cpp:647-871) */
int e3d_flux_conv(e3d_ctx* ctx, int block, int scheme);
/* FluxScheme::computeFluxDiss with current_beta_ = beta (Solver/head/FluxScheme.h:19):
 * RoeScheme.cpp:12-311 or AusmScheme.cpp:12-243 */
int e3d_flux_diss(e3d_ctx* ctx, int block, int scheme, double beta);
/* CentralIrsSmoother::smoothResidual (Solver/src/CentralIrsSmoother.cpp:12-183): one central
 * implicit-smoothing Jacobi sweep on conv and diss of the block's real cells, including the
 * reference's use of the w-dissipation value in the v sweep (:90); the w sweep, whose Jacobi
 * array the reference never writes, smooths with zero. NoSmoother is no call at all.
 * e3d_residual calls it when params.smoothing == 1; e3d_iterate then runs the operators
 * block by block instead of the fused stage kernels. */
int e3d_smooth_residual(e3d_ctx* ctx, int block);
/* ResidualCalculator::computeResidual (Solver/src/ResidualCalculator.cpp:24-41) for RK stage `stage` */
int e3d_residual(e3d_ctx* ctx, int block, int stage);
/* Updater::updateInternalBlock (Solver/src/Updater.cpp:12-78) for RK stage `stage` */
int e3d_update(e3d_ctx* ctx, int block, int stage);
/* Updater::updateBoundary -> BoundaryCellIds::updateBoundary for every physical marker
 * (Solver/src/Updater.cpp:80-88; WallCellIds.cpp:9-95, SymmetryCellIds.cpp:9-96, FarfieldCellIds.cpp:10-114) */
int e3d_bc(e3d_ctx* ctx, int block);
/* BoundaryCellIds::updateBoundary (DataStructure/head/BoundaryCellIds.h:13) of one marker of a
 * block (marker index in file order); a CONNEXION marker is a no-op as in ConnexionBlockCellIds. */
int e3d_bc_marker(e3d_ctx* ctx, int block, int marker);
/* RungeKutta::computeRungeKutta (Solver/src/RungeKutta.cpp:80-94) */
int e3d_runge_kutta(e3d_ctx* ctx, int block);
/* BlockCommunicator::updateBoundaries (MPI/src/BlockCommunicator.cpp:212-268): interface ghosts
 * take the primitive state of their origin cells (the only copied field the path reads later). */
int e3d_exchange(e3d_ctx* ctx);
/* PostProcessing::computeFlowData + BlockCommunicator::getGlobal + PostProcessing::process
 * (PostProcessing/src/PostProcessing.cpp:198-308, Convergence.cpp:27-144,
 * AerodynamicParameters.cpp:163-239): fills one monitor row and advances the iteration count. */
int e3d_monitors(e3d_ctx* ctx, e3d_monitor* out);

/* ------------------------------------------------------------ fused path
 * Solver::solve (Solver/src/Solver.cpp:36-98) for up to n iterations with
 * stage-fused kernels over all blocks; history (may be NULL) receives one row
 * per iteration. Returns the number of iterations run (>= 0) or a status < 0.
 * Stops early on the reference's stop test or on e3d_request_stop. */
int e3d_iterate(e3d_ctx* ctx, int n, e3d_monitor* history);
/* Equivalent of the reference's STOP file (PostProcessing.cpp:72-80): the next
 * iteration boundary ends e3d_iterate. */
int e3d_request_stop(e3d_ctx* ctx);
/* Resets the iteration counter and the reference residual of the drop. */
int e3d_reset_monitors(e3d_ctx* ctx);
/* Iteration counter and iteration-0 residual norms (PostProcessing's current_iter_ and
 * ro_rms0_mesh_ ... pp_rms0_mesh_, PostProcessing/src/PostProcessing.cpp:86-123): read them
 * to write a restart file, set them to resume a run from one (no counterpart in the
 * reference, whose restart file is empty: PostProcessing/src/OutputTecplot.cpp:475-495). */
int e3d_monitor_state(const e3d_ctx* ctx, int* iter, double rms0[5]);
int e3d_set_monitor_state(e3d_ctx* ctx, int iter, const double rms0[5]);

/* ------------------------------------------------------------- batches
 * Many independent meshes ("instances") of one scheme, each with its own flow
 * conditions, advanced together: every RK stage is one launch over all of them
 * (no counterpart in the reference, which runs one case per process). An
 * instance stops on its own stop test (PostProcessing.cpp:59-83, per instance;
 * a non-finite residual norm also stops it) and its cells are skipped from then
 * on; its state stays as it was at its last iteration. */
typedef struct e3d_instance {
    double mach, aoa_deg;  /* freestream (PrimitiveVariables.cpp:41-56) */
    double cmac;           /* reference length of the coefficients */
    double criterion;      /* stop when log10 density residual drop <= criterion */
    int max_iter;          /* stop at iteration max_iter - 1 */
    double rms_ref[5];     /* residual norms the drop is measured from; all 0: the instance's own
                              iteration-0 norms (PostProcessing.cpp:240-250). A warm start passes the
                              cold start's iteration-0 norms here. */
} e3d_instance;
/* Uploads n meshes as one batch. Scheme, gamma, CFL and RK stages come from
 * e3d_set_params; mach/aoa/cmac/criterion/max_iter of params are not used.
 * Blocks are numbered across the batch: instance i owns blocks
 * [first, first + count) of e3d_batch_blocks, and the per-block entry points
 * (e3d_get_array, e3d_set_array, granular operators) take those numbers.
 * e3d_init_freestream then sets each instance to its own freestream. */
int e3d_upload_batch(e3d_ctx* ctx, const e3d_mesh* const* meshes, const e3d_instance* inst, int n);
int e3d_batch_size(const e3d_ctx* ctx);
int e3d_batch_blocks(const e3d_ctx* ctx, int inst, int* first_block, int* n_blocks);
/* Per-instance outcome of e3d_batch_run. */
typedef struct e3d_batch_status {
    int iterations;        /* iterations the instance has run */
    int stopped;           /* 1 once its stop test fired */
    e3d_monitor last;      /* its last monitor row */
    double stop_seconds;   /* device time from the first e3d_batch_run after e3d_init_freestream
                              to the end of its last iteration (0 while running) */
} e3d_batch_status;
/* Runs up to max_iterations more iterations, stopping early once every instance
 * has stopped. use_graph = 1 replays a hipGraph of two iterations (stage
 * kernels, exchange, per-block monitors, per-instance stop test) with one host
 * check per 16 iterations; 0 launches the same kernels one by one. status
 * (may be NULL) receives one entry per instance. Returns the number of
 * iterations launched or a status < 0. */
int e3d_batch_run(e3d_ctx* ctx, int max_iterations, int use_graph, e3d_batch_status* status);
/* Monitor rows of one instance, iterations [0, n) with n = min(max_rows, rows
 * recorded); returns n. Rows beyond the history capacity (max over the
 * instances' max_iter, at most 200000) are not kept. */
int e3d_batch_history(e3d_ctx* ctx, int inst, e3d_monitor* rows, int max_rows);
/* Workgroup-to-tile order of the stage kernels (no counterpart in the
 * reference, which runs one cell loop on one core). Workgroups are dealt to
 * the GPU's eight XCDs in turn; E3D_TILES_LINEAR maps workgroup i to tile i,
 * E3D_TILES_XCD_CONTIGUOUS gives each XCD one contiguous range of tiles,
 * E3D_TILES_XCD_CHUNKS deals runs of 4096 consecutive cells to the XCDs in
 * turn, E3D_TILES_AUTO (default) takes XCD_CONTIGUOUS for meshes of fewer
 * than 262,144 cells and XCD_CHUNKS for larger ones. Addresses only: results
 * are the same. */
#define E3D_TILES_AUTO (-1)
#define E3D_TILES_LINEAR 0
#define E3D_TILES_XCD_CONTIGUOUS 1
#define E3D_TILES_XCD_CHUNKS 2
int e3d_set_tile_order(e3d_ctx* ctx, int order);
/* Cells per workgroup of the stage kernels: 64, 128 or 256 (default). No
 * counterpart in the reference (one cell loop on one core). A thread owns one
 * cell whatever the tile size and the residual-norm partial sums are formed
 * per 64 cells, so results are the same bit for bit. */
int e3d_set_tile_cells(e3d_ctx* ctx, int cells);
/* Where the second-order Green-Gauss stage kernel takes each face slot's area
 * vector S and r vector from. No counterpart in the reference, whose cells
 * hold pointers to stored face metrics (Metrics/src/MetricsInitializer.cpp).
 * E3D_GEOMETRY_AUTO (default): on a mesh of tetrahedra, rebuilt in the kernel
 * from a per-cell copy of the cell's four node coordinates with the metric
 * expressions of MetricsInitializer.cpp:64-142, 337-504 (100 bytes read per
 * cell instead of 192); the upload rebuilds every slot on the host the same
 * way and keeps this form only if all equal the stored metrics bit for bit.
 * Other meshes, least squares, and E3D_GEOMETRY_STREAMED: the stored per-slot
 * copies. Addresses and arithmetic only: results are the same. */
#define E3D_GEOMETRY_AUTO (-1)
#define E3D_GEOMETRY_STREAMED 0
int e3d_set_slot_geometry(e3d_ctx* ctx, int mode);
/* 1 when the Green-Gauss stage kernel of the uploaded mesh and current
 * parameters rebuilds the slot geometry from node copies, 0 when it streams
 * the stored copies. */
int e3d_slot_geometry_nodes(const e3d_ctx* ctx);

/* Memory order of the cells and faces of the next uploaded mesh or batch.
 * E3D_ORDER_REFERENCE (default) keeps the reference's numbering
 * (BlockBuilder/src/BlockBuilder.cpp:147-458, cells as the mesh file lists them,
 * faces by first appearance). E3D_ORDER_RCM places each block's real cells in
 * reverse Cuthill-McKee order, E3D_ORDER_MORTON along a Morton (Z-order) curve
 * of their centroids, and the faces by first appearance over them, so that
 * neighbour gathers stay local whatever order the mesh file came in. Only
 * addresses change: every cell keeps its faces in the reference's order, so
 * every per-cell value is bit for bit the same; the residual norms, summed
 * over cells in memory order, agree to rounding. Arrays are still read and
 * written in the reference's numbering (e3d_get_array, e3d_set_array). The
 * reference has no counterpart (its Block arrays are in file order,
 * DataStructure/head/Block.h:17-60). */
#define E3D_ORDER_REFERENCE 0
#define E3D_ORDER_RCM 1
#define E3D_ORDER_MORTON 2
int e3d_set_ordering(e3d_ctx* ctx, int mode);
/* The device positions that mode gives one block of a host mesh:
 * cell_pos[q] for its nreal real cells, face_pos[f] for its nface faces (both
 * permutations, 0-based within the block). Host only; for callers that keep
 * their own arrays in device order. */
int e3d_mesh_layout(const e3d_mesh* mesh, int block, int mode, int* cell_pos, int* face_pos);
/* Host only: the per-cell node copies and slot codes from which the
 * Green-Gauss stage kernel rebuilds each face slot's S and r on a mesh of
 * tetrahedra (E3D_GEOMETRY_AUTO). node_xyz[12 q + 3 j + d]: coordinate d of
 * node j of real cell q in its element node order; slot_code[q]: one byte per
 * face slot k (bits 8k..8k+7) with the local node indices of the face's three
 * nodes in the face's order (2 bits each) and bit 6 set where the slot holds
 * the metric formula's S negated. Every slot is rebuilt with the expressions
 * of Metrics/src/MetricsInitializer.cpp:64-142, 337-504 and compared bit for
 * bit with the block's stored metrics; E3D_EINVAL if the block holds other
 * cells or a rebuilt value differs. No counterpart in the reference. */
int e3d_mesh_tet_geometry(const e3d_mesh* mesh, int block, double* node_xyz, unsigned* slot_code);

/* ------------------------------------------------------ learned warm start
 * A coordinate MLP that predicts an initial flow field, replacing the uniform
 * freestream of PrimitiveVariables::initializeFlowField
 * (DataStructure/src/PrimitiveVariables.cpp:58-69; no counterpart in the
 * reference otherwise). n_in features (<= 32) -> depth hidden layers of width
 * units (a multiple of 128) with SiLU -> n_out outputs (<= 32). Trained and
 * evaluated on the GPU with BF16 operands and FP32 accumulation on the matrix
 * cores; master weights and Adam state FP32. The solver's own residual decides
 * convergence: a prediction only replaces the starting state. */
typedef struct e3d_mlp e3d_mlp;
int e3d_mlp_create(int device, int n_in, int width, int depth, int n_out, unsigned long long seed, e3d_mlp** out);
void e3d_mlp_destroy(e3d_mlp* mlp);
/* Number of parameters; flat layout per layer: W [out x in] row-major, then b [out]. */
long e3d_mlp_param_count(const e3d_mlp* mlp);
int e3d_mlp_get_params(e3d_mlp* mlp, float* out);
int e3d_mlp_set_params(e3d_mlp* mlp, const float* in);
/* Training set in device memory: x [n x n_in], targets t [n x n_out] (host arrays). */
int e3d_mlp_set_data(e3d_mlp* mlp, const float* x, const float* t, long n);
/* Mean-squared-error loss and its gradients (flat, as the parameters) on the
 * samples idx[0..batch) of the training set, without an update; batch a
 * multiple of 512. */
int e3d_mlp_gradients(e3d_mlp* mlp, const int* idx, int batch, float* loss, float* grads);
/* steps Adam steps on mini-batches (multiple of 512) drawn without replacement
 * from a seeded shuffle of the training set; cosine != 0 decays lr to lr/20
 * over the steps; loss (may be NULL) receives one value per step. */
int e3d_mlp_train(e3d_mlp* mlp, int steps, int batch, float lr, int cosine, unsigned long long seed, float* loss);
/* Outputs y [n x n_out] for features x [n x n_in] (host arrays). */
int e3d_mlp_predict(e3d_mlp* mlp, const float* x, long n, float* y);
/* e3d_mlp_predict for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc); it
 * has no counterpart in the reference. x [n x n_in] and y [n x n_out] are
 * device memory of the network's GPU, filled and read by the caller. The
 * library allocates nothing in this call: rows go through a static per-GPU
 * scratch of 8 MiB in chunks of the largest multiple of 512 rows that fits
 * (3,584 rows at width 512), so any n is accepted; widths whose 512 rows do
 * not fit (above 3,968) are rejected with E3D_EINVAL. Outputs equal
 * e3d_mlp_predict's bit for bit. Runs on the network's stream and returns
 * after it drains; the caller must have finished writing x. Calls that
 * overlap on one GPU share the scratch and race. */
int e3d_mlp_predict_noalloc(e3d_mlp* mlp, const float* x, long n, float* y);
/* Average time of one training step's forward and backward passes on a batch
 * (HIP events, reps repetitions) and its matrix-core flops. */
int e3d_mlp_time_step(e3d_mlp* mlp, int batch, int reps, double* ms, double* flops);

/* --------------------------------------------------------- measurement */
/* Average duration in ms of the named kernel over the last e3d_iterate call
 * (HIP events on the context's stream); kernel = "interp", "flux_update", ... */
int e3d_kernel_time(e3d_ctx* ctx, const char* kernel, double* avg_ms, int* launches);
/* Streaming HBM bandwidth of this GPU for a mix of reads and writes (no
 * counterpart in the reference; it calibrates the stage kernels' roofline
 * fractions): `bytes` of traffic per launch, `reads` planes read and `writes`
 * planes written per element (1:0, 7:1, 3:1, 2:1, 1:1, 1:2 or 0:1), 16-byte
 * nontemporal loads and stores, best of 1 and 4 elements per thread; GB/s
 * over reps launches timed with HIP events after 3 warmup launches. */
int e3d_stream_bandwidth(int device, long bytes, int reads, int writes, int reps, double* gbs);
int e3d_enable_timing(e3d_ctx* ctx, int on);
int e3d_sync(e3d_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* EULER3D_AMD_H */
