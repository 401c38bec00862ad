// C ABI: device context, mesh upload, granular operators and the fused
// iteration loop (Solver::solve, Solver/src/Solver.cpp:36-98).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/euler3d.h"
#include "abi_internal.hpp"
#include "device.hpp"
#include "mesh_internal.hpp"
#include "ordering.hpp"
#include "parallel.hpp"
#include "physics.hpp"
#include "tet_geometry.hpp"

namespace e3d {
void launch_save_w0(const Range&, const DMesh&, const DState&, double, hipStream_t);
void launch_spectral_radius(const Range&, const DMesh&, const DState&, double, hipStream_t);
void launch_timestep(const Range&, const DMesh&, const DState&, double, hipStream_t);
void launch_gradient(const Range&, const DMesh&, const DState&, int, hipStream_t);
void launch_limiter(const Range&, const DMesh&, const DState&, int, double, double, hipStream_t);
void launch_flux_conv(const Range&, const DMesh&, const DState&, int, const Gas&, hipStream_t);
void launch_flux_diss(const Range&, const DMesh&, const DState&, int, const Gas&, double, hipStream_t);
void launch_update(const Range&, const DMesh&, const DState&, double, double, hipStream_t);
void launch_bc(const Range&, const DMesh&, const DState&, double, const double*, hipStream_t);
void launch_exchange(int, const int*, const int*, int, double*, double*, double*, int, hipStream_t);
void launch_rms_partials(const DMesh&, const DState&, hipStream_t);
void launch_smooth(const Range&, const DMesh&, const DState&, hipStream_t);
void launch_block_monitors(int, int, const int*, const int*, const int*, const int*, const int*, const int*,
                           const double*, const DMesh&, const double*, const double*, const double*, double*, double*,
                           hipStream_t);
void launch_interp_fused(const DMesh&, const DState&, int, int, double, double, int, bool, hipStream_t);
void launch_batch_reduce(int, const int*, const double*, const void*, int*, double*, int*, int*, void*, int, int*,
                         long long*, hipStream_t);
void launch_clock_mark(long long*, hipStream_t);
void launch_batch_freeze(const DMesh&, const int*, const int*, const double*, double*, hipStream_t);

void launch_flux_update(const DMesh&, const DState&, int, bool, bool, bool, bool, const Gas&, double, double, double,
                        double, int, hipStream_t);
}  // namespace e3d

using namespace e3d;

namespace {
enum KernelId { KI_INTERP, KI_FLUX_UPDATE0, KI_FLUX_UPDATE, KI_FLUX_UPDATE_ND, KI_EXCHANGE, KI_MONITORS, KI_COUNT };
// flux_update0: stage 0 (with W0, spectral radius and time step); flux_update /
// flux_update_nodiss: later stages with beta != 0 and beta == 0.
const char* kKernelNames[KI_COUNT] = {"interp", "flux_update0", "flux_update", "flux_update_nodiss", "exchange",
                                      "monitors"};
}  // namespace

struct e3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool have_params = false, have_mesh = false;
    e3d_params params{};
    double fs[5] = {0, 0, 0, 0, 0};
    double alpha[5] = {0, 0, 0, 0, 0}, beta[5] = {0, 0, 0, 0, 0};
    Gas gas{};
    // layout
    int nb = 0;
    std::vector<int> nreal, nall, roff, goff, foff, ghost_count;
    std::vector<std::vector<int>> marker_off;  // per block: ghost offsets of its markers (relative to the block's first ghost)
    DMesh m{};
    DState s{};
    // device allocations
    std::vector<void*> allocs;
    int* d_itf_dst = nullptr;
    int* d_itf_org = nullptr;
    int n_itf = 0;
    int *d_tile0 = nullptr, *d_ntile = nullptr, *d_nreal = nullptr, *d_wall_off = nullptr, *d_wall_ghost = nullptr,
        *d_wall_face = nullptr;
    double* d_fcenter = nullptr;
    double* d_mon = nullptr;   // [nb][16]
    double* d_mpart = nullptr; // [nb][mon_split][10] first-level monitor sums
    int mon_split = 1;         // workgroups per block of the first monitor level
    double* h_mon = nullptr;   // pinned
    // instances: one mesh is instance 0 with the conditions of params; a batch
    // (e3d_upload_batch) has one instance per mesh with its own conditions
    bool batch = false;
    int n_inst = 1;
    std::vector<e3d_instance> inst;
    std::vector<int> blk_inst, inst_b0;  // block -> instance; first block of each instance, then nb
    std::vector<double> inst_fs;         // [n_inst][5]
    int *d_blk_inst = nullptr, *d_g_blk = nullptr, *d_inst_b0 = nullptr;
    double *d_inst_fs = nullptr, *d_blk_qcmac = nullptr;
    // batched-run state on the device
    int *d_active = nullptr, *d_frozen = nullptr, *d_iter = nullptr, *d_nactive = nullptr;
    double* d_rms0 = nullptr;
    void* d_prm = nullptr;
    void* d_hist = nullptr;
    int hist_cap = 0;
    int* h_nactive = nullptr;  // pinned
    long long* d_clock = nullptr;  // [0] device wall clock at the start of the run, [1 + i] when instance i stopped
    bool clock_started = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    double* graph_W = nullptr;   // W buffer the graph was captured with
    bool capturing = false;
    int tile_pref = E3D_TILES_AUTO;       // tile order asked for (e3d_set_tile_order)
    int tile_cells = 256;                 // cells per workgroup of the stage kernels (e3d_set_tile_cells)
    int slot_geometry = E3D_GEOMETRY_AUTO;  // Green-Gauss slot geometry source (e3d_set_slot_geometry)
    int tile_map = E3D_TILES_XCD_CHUNKS;  // tile order in use (resolve_tile_order)
    int ordering = ORDER_REFERENCE;  // memory order of the next upload (e3d_set_ordering)
    // per block: device position of each real cell and face (empty: the reference's numbering)
    std::vector<std::vector<int>> cell_pos, face_pos;
    // monitors
    int iter = 0;
    bool started = false;      // initial exchange done
    double rms0[5] = {0, 0, 0, 0, 0};
    std::atomic<int> stop_request{0};
    // timing
    bool timing = false;
    struct Ev {
        hipEvent_t a, b;
        int id;
    };
    std::vector<Ev> ev_pool;
    size_t ev_used = 0;
    double t_ms[KI_COUNT] = {};
    int t_n[KI_COUNT] = {};
};

namespace e3d {
void ctx_record_error(e3d_ctx* ctx, const std::string& msg) { ctx->err = msg; }
const char* ctx_error_text(const e3d_ctx* ctx) { return ctx->err.empty() ? nullptr : ctx->err.c_str(); }
}  // namespace e3d

#define HIP_TRY(ctx, call)                                                                              \
    do {                                                                                                \
        hipError_t _e = (call);                                                                         \
        if (_e != hipSuccess)                                                                           \
            return set_error(ctx, E3D_EHIP, std::string(#call) + ": " + hipGetErrorString(_e));         \
    } while (0)

namespace {

void drop_graph(e3d_ctx* c) {
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->graph_W = nullptr;
}

// The tile order in use: the one asked for, or by size (E3D_TILES_AUTO): a
// mesh of fewer than kAutoChunkCells cells keeps each XCD on one contiguous
// range (runs of 4096 cells would leave some XCDs a run short), a larger one
// deals the runs to the XCDs in turn.
constexpr int kAutoChunkCells = 1024 * 256;
void resolve_tile_order(e3d_ctx* c) {
    c->tile_map = c->tile_pref != E3D_TILES_AUTO ? c->tile_pref
                  : c->m.NRp < kAutoChunkCells   ? E3D_TILES_XCD_CONTIGUOUS
                                                 : E3D_TILES_XCD_CHUNKS;
    c->m.tile_map = c->tile_map;
}

void free_device(e3d_ctx* c) {
    drop_graph(c);
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    if (c->h_mon) (void)hipHostFree(c->h_mon);
    c->h_mon = nullptr;
    if (c->h_nactive) (void)hipHostFree(c->h_nactive);
    c->h_nactive = nullptr;
    c->have_mesh = false;
}

template <class T>
int dalloc(e3d_ctx* c, T** p, size_t n) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess)
        return set_error(c, E3D_ENOMEM, "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
    e = hipMemsetAsync(q, 0, bytes, c->stream);
    if (e != hipSuccess) return set_error(c, E3D_EHIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
    c->allocs.push_back(q);
    *p = (T*)q;
    return E3D_OK;
}

template <class T>
int upload(e3d_ctx* c, T** p, const std::vector<T>& h) {
    int rc = dalloc(c, p, h.size());
    if (rc) return rc;
    if (!h.empty()) HIP_TRY(c, hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return E3D_OK;
}

// Device position of a block's cell (reference-numbered, real or ghost) and face.
int global_cell(const e3d_ctx* c, int b, int local) {
    if (local >= c->nreal[b]) return c->m.NRp + c->goff[b] + (local - c->nreal[b]);
    return c->roff[b] + (c->cell_pos[b].empty() ? local : c->cell_pos[b][local]);
}
int global_face(const e3d_ctx* c, int b, int f) { return c->foff[b] + (c->face_pos[b].empty() ? f : c->face_pos[b][f]); }

Range block_range(const e3d_ctx* c, int b) {
    if (b < 0) return Range{0, c->m.NRp, 0, c->m.NG};
    return Range{c->roff[b], c->roff[b] + c->nreal[b], c->goff[b], c->goff[b] + (c->nall[b] - c->nreal[b])};
}

int check_ready(e3d_ctx* c, int block) {
    if (!c) return set_error(nullptr, E3D_EINVAL, "null context");
    if (!c->have_mesh) return set_error(c, E3D_ESTATE, "no mesh uploaded");
    if (!c->have_params) return set_error(c, E3D_ESTATE, "no parameters set");
    if (block < -1 || block >= c->nb) return set_error(c, E3D_EINVAL, "block index out of range");
    return E3D_OK;
}

int launch_status(e3d_ctx* c) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(c, E3D_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return E3D_OK;
}

void ev_begin(e3d_ctx* c, int id, size_t* slot) {
    if (!c->timing || c->capturing) return;
    if (c->ev_used == c->ev_pool.size()) {
        e3d_ctx::Ev e;
        (void)hipEventCreate(&e.a);
        (void)hipEventCreate(&e.b);
        c->ev_pool.push_back(e);
    }
    *slot = c->ev_used++;
    c->ev_pool[*slot].id = id;
    (void)hipEventRecord(c->ev_pool[*slot].a, c->stream);
}
void ev_end(e3d_ctx* c, size_t slot) {
    if (!c->timing || c->capturing) return;
    (void)hipEventRecord(c->ev_pool[slot].b, c->stream);
}
// After a synchronisation: fold the recorded launches into the per-kernel totals.
void ev_collect(e3d_ctx* c) {
    for (size_t i = 0; i < c->ev_used; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_pool[i].a, c->ev_pool[i].b) == hipSuccess) {
            c->t_ms[c->ev_pool[i].id] += ms;
            c->t_n[c->ev_pool[i].id] += 1;
        }
    }
    c->ev_used = 0;
}

// Residual norms and force coefficients of every block, summed over blocks in
// block order (BlockCommunicator::getGlobal serial branch, :341-369), then the
// drop and the stop test (PostProcessing::process, :220-308).
int finish_monitor(e3d_ctx* c, e3d_monitor* out) {
    size_t ev = 0;
    ev_begin(c, KI_MONITORS, &ev);
    launch_block_monitors(c->nb, c->mon_split, c->d_tile0, c->d_ntile, c->d_nreal, c->d_wall_off, c->d_wall_ghost,
                          c->d_wall_face, c->d_fcenter, c->m, c->s.W, c->s.part, c->d_blk_qcmac, c->d_mpart, c->d_mon,
                          c->stream);
    ev_end(c, ev);
    int rc = launch_status(c);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpyAsync(c->h_mon, c->d_mon, sizeof(double) * 16 * c->nb, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    ev_collect(c);
    const double aoa_rad = c->params.aoa_deg * 4. * std::atan(1.) / 180;
    double sum[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < c->nb; ++b) {
        const double* r = c->h_mon + 16 * b;
        for (int v = 0; v < 5; ++v) sum[v] += r[v];
        const double cl = r[5], cd = r[6];
        sum[5] += cl * std::cos(aoa_rad) - cd * std::sin(aoa_rad);
        sum[6] += cl * std::sin(aoa_rad) + cd * std::cos(aoa_rad);
        sum[7] += r[7];
        sum[8] += r[8];
        sum[9] += r[9];
    }
    if (c->iter == 0)
        for (int v = 0; v < 5; ++v) c->rms0[v] = sum[v];
    e3d_monitor m{};
    m.iter = c->iter;
    for (int v = 0; v < 5; ++v) {
        m.rms[v] = sum[v];
        m.drop[v] = std::log10(sum[v]) - std::log10(c->rms0[v]);
    }
    m.cl = sum[5];
    m.cd = sum[6];
    m.cmx = sum[7];
    m.cmy = sum[8];
    m.cmz = sum[9];
    m.stop = (c->iter + 1 == c->params.max_iter) || (m.drop[0] <= c->params.criterion) ? 1 : 0;
    if (out) *out = m;
    ++c->iter;
    for (int v = 0; v < 5; ++v)
        if (!std::isfinite(sum[v])) return set_error(c, E3D_ENONFINITE, "residual norm is not finite at iteration " +
                                                                             std::to_string(m.iter));
    return E3D_OK;
}

int exchange(e3d_ctx* c, int copy_all) {
    size_t ev = 0;
    ev_begin(c, KI_EXCHANGE, &ev);
    launch_exchange(c->n_itf, c->d_itf_dst, c->d_itf_org, c->m.NA, c->s.W, c->s.grad, c->s.lim, copy_all, c->stream);
    ev_end(c, ev);
    return launch_status(c);
}

// Freestream state (PrimitiveVariables::calculateFreeVariables,
// DataStructure/src/PrimitiveVariables.cpp:41-56).
void freestream_state(double gamma, double mach, double aoa_deg, double fs[5]) {
    const double pi = std::atan(1.0) * 4.0;
    const double alpha_rad = aoa_deg * pi / 180.0;
    const double cs = std::cos(alpha_rad), sn = std::sin(alpha_rad);
    fs[0] = 1.0;
    fs[1] = mach * std::sqrt(gamma) * cs;
    fs[2] = mach * std::sqrt(gamma) * sn;
    fs[3] = 0.0;
    fs[4] = 1.0;
}

// Per-instance freestream, monitor scale and stop constants, on the host and the
// device. One mesh: instance 0 takes its conditions from params.
int sync_instances(e3d_ctx* c) {
    if (!c->have_params || !c->have_mesh) return E3D_OK;
    const e3d_params& p = c->params;
    if (!c->batch) c->inst.assign(1, e3d_instance{p.mach, p.aoa_deg, p.cmac, p.criterion, p.max_iter, {0, 0, 0, 0, 0}});
    c->inst_fs.assign(5 * (size_t)c->n_inst, 0.0);
    std::vector<InstParam> prm(c->n_inst);
    for (int i = 0; i < c->n_inst; ++i) {
        const e3d_instance& q = c->inst[i];
        freestream_state(p.gamma, q.mach, q.aoa_deg, &c->inst_fs[5 * (size_t)i]);
        const double aoa_rad = q.aoa_deg * 4. * std::atan(1.) / 180;  // as the monitor sums of finish_monitor
        prm[i] = InstParam{std::cos(aoa_rad), std::sin(aoa_rad), q.criterion, q.max_iter, q.rms_ref[0] > 0.0 ? 1 : 0};
    }
    for (int v = 0; v < 5; ++v) c->fs[v] = c->inst_fs[v];
    std::vector<double> qcmac(c->nb);
    for (int b = 0; b < c->nb; ++b) {
        const e3d_instance& q = c->inst[c->blk_inst[b]];
        qcmac[b] = 0.5 * p.gamma * q.mach * q.mach * q.cmac;
    }
    // the context's stream does not synchronise with the null stream: let the
    // work queued on it (which reads these constants) finish first
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(c->d_inst_fs, c->inst_fs.data(), sizeof(double) * c->inst_fs.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_blk_qcmac, qcmac.data(), sizeof(double) * qcmac.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_prm, prm.data(), sizeof(InstParam) * prm.size(), hipMemcpyHostToDevice));
    return E3D_OK;
}

const double* block_fs(const e3d_ctx* c, int b) { return &c->inst_fs[5 * (size_t)c->blk_inst[b]]; }

}  // namespace

extern "C" {

int e3d_create(int device, e3d_ctx** out) {
    if (!out) return set_error(nullptr, E3D_EINVAL, "e3d_create: null output");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return set_error(nullptr, E3D_EHIP, std::string("no HIP device: ") + (e != hipSuccess ? hipGetErrorString(e) : "count 0"));
    if (device < 0 || device >= n) return set_error(nullptr, E3D_EINVAL, "device index out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return set_error(nullptr, E3D_EHIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    auto* c = new e3d_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return set_error(nullptr, E3D_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return E3D_OK;
}

void e3d_destroy(e3d_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_device(c);
    for (auto& e : c->ev_pool) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int e3d_set_params(e3d_ctx* c, const e3d_params* p) {
    if (!c || !p) return set_error(c, E3D_EINVAL, "e3d_set_params: null argument");
    if (p->stages < 1 || p->stages > 5) return set_error(c, E3D_EINVAL, "RK stages must be 1..5");
    if (p->order != 1 && p->order != 2) return set_error(c, E3D_EINVAL, "order must be 1 or 2");
    if (p->flux != E3D_FLUX_ROE && p->flux != E3D_FLUX_AUSM) return set_error(c, E3D_EINVAL, "unknown flux scheme");
    if (p->gradient != E3D_GRAD_GREEN_GAUSS && p->gradient != E3D_GRAD_LEAST_SQUARES)
        return set_error(c, E3D_EINVAL, "unknown gradient");
    if (p->limiter != E3D_LIM_BARTH_JESPERSEN && p->limiter != E3D_LIM_VENKATAKRISHNAN)
        return set_error(c, E3D_EINVAL, "unknown limiter");
    if (p->smoothing != 0 && p->smoothing != 1) return set_error(c, E3D_EINVAL, "smoothing must be 0 or 1");
    if (!(p->gamma > 1.0)) return set_error(c, E3D_EINVAL, "gamma must exceed 1");
    c->params = *p;
    // RungeKutta constructor tables (Solver/src/RungeKutta.cpp:23-68)
    static const double A[5][5] = {{1.0}, {0.6, 1.0}, {0.6, 0.6, 1.0}, {0.25, 1.0 / 3.0, 0.5, 1.0},
                                   {0.2742, 0.2067, 0.5020, 0.5142, 1.0}};
    static const double B[5][5] = {{1.0}, {1.0, 0.0}, {1.0, 0.0, 0.0}, {1.0, 0.5, 0.0, 0.0},
                                   {1.0, 0.0, 0.56, 0.0, 0.44}};
    for (int i = 0; i < 5; ++i) {
        c->alpha[i] = A[p->stages - 1][i];
        c->beta[i] = B[p->stages - 1][i];
    }
    freestream_state(p->gamma, p->mach, p->aoa_deg, c->fs);
    c->gas.gamma = p->gamma;
    c->gas.g_over_gm1 = p->gamma / (p->gamma - 1.0);
    c->gas.gm1 = p->gamma - 1.0;
    c->have_params = true;
    drop_graph(c);
    return sync_instances(c);
}

// Every block of every mesh, in upload order, with the instance it belongs to.
struct BlockSrc {
    const HostBlock* B;
    const RawBlock* R;
    const e3d_mesh::Flat* F;
    int ghost_count;
    int base;  // global index of the mesh's block 0
    int inst;
};

static int upload_meshes(e3d_ctx* c, const std::vector<const e3d_mesh*>& meshes) {
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    free_device(c);
    std::vector<BlockSrc> src;
    c->inst_b0.assign(1, 0);
    for (size_t i = 0; i < meshes.size(); ++i) {
        const int base = (int)src.size();
        for (size_t b = 0; b < meshes[i]->built.size(); ++b)
            src.push_back({&meshes[i]->built[b], &meshes[i]->raw.blocks[b], &meshes[i]->flat[b], meshes[i]->raw.topo[b].ghost_count, base,
                           (int)i});
        c->inst_b0.push_back((int)src.size());
    }
    c->n_inst = (int)meshes.size();
    const int nb = (int)src.size();
    c->blk_inst.assign(nb, 0);
    for (int b = 0; b < nb; ++b) c->blk_inst[b] = src[b].inst;
    c->nb = nb;
    c->nreal.assign(nb, 0);
    c->nall.assign(nb, 0);
    c->roff.assign(nb, 0);
    c->goff.assign(nb, 0);
    c->foff.assign(nb, 0);
    c->ghost_count.assign(nb, 0);
    c->marker_off.assign(nb, std::vector<int>(1, 0));
    long NRp = 0, NG = 0, NF = 0;
    int nslot = 1;
    for (int b = 0; b < nb; ++b) {
        const HostBlock& B = *src[b].B;
        c->nreal[b] = B.nreal;
        c->nall[b] = B.nall;
        c->roff[b] = (int)NRp;
        c->goff[b] = (int)NG;
        c->foff[b] = (int)NF;
        c->ghost_count[b] = src[b].ghost_count;
        NRp += (long)(B.nreal + kWG - 1) / kWG * kWG;
        NG += B.nall - B.nreal;
        NF += B.nface;
        for (int q = 0; q < B.nreal; ++q) nslot = std::max(nslot, B.cell_off[q + 1] - B.cell_off[q]);
    }
    if (nslot > kMaxSlots) return set_error(c, E3D_EINVAL, "cells with more than 6 faces are not supported");
    c->cell_pos.assign(nb, {});
    c->face_pos.assign(nb, {});
    if (c->ordering != ORDER_REFERENCE)
        for (int b = 0; b < nb; ++b) block_layout(*src[b].B, c->ordering, c->cell_pos[b], c->face_pos[b]);
    if (NRp + NG >= (1L << 28) || NF >= (1L << 31) - 1)  // kernels address cell arrays with 32-bit byte offsets
        return set_error(c, E3D_EINVAL, "mesh too large: at most 2^28 cells including ghosts");
    DMesh& m = c->m;
    m.NRp = (int)NRp;
    m.NG = (int)NG;
    m.NA = (int)(NRp + NG);
    m.NF = (int)NF;
    m.nslot = nslot;
    const size_t NR = NRp, NA = m.NA;

    std::vector<int> nf(NR, 0), nbr((size_t)nslot * NR, 0), fid((size_t)nslot * NR, 0);
    std::vector<int> g_in(NG), g_face(NG), g_kind(NG);
    std::vector<double> S(3 * (size_t)NF), area(NF), rL(3 * (size_t)NF), rR(3 * (size_t)NF), fc(3 * (size_t)NF);
    std::vector<double> vol(NA, 1.0), lsqw((size_t)nslot * 3 * NR, 0.0);
    std::vector<int> wg_block(NR / kWG, 0), tile0(nb), ntile(nb), wall_off(nb + 1, 0), wall_ghost, wall_face;
    std::vector<int> itf_dst, itf_org;
    // per-slot face geometry of the fused path: S oriented away from the cell
    // and the cell's r vector, each as an (x, y) pair plane and a z plane
    const size_t NS = (size_t)nslot * NR;
    std::vector<double> sSxy(2 * NS, 0.0), sSz(NS, 0.0), sRxy(2 * NS, 0.0), sRz(NS, 0.0), g_r(3 * (size_t)NG, 0.0),
        g_n(3 * (size_t)NG, 0.0);
    std::vector<int> mate(NS, 0);
    std::vector<int> g_blk(NG, 0);
    for (int b = 0; b < nb; ++b) {
        const HostBlock& B = *src[b].B;
        const auto& F = *src[b].F;
        for (int gl = B.nreal; gl < B.nall; ++gl) g_blk[c->goff[b] + (gl - B.nreal)] = b;
        tile0[b] = c->roff[b] / kWG;
        ntile[b] = (B.nreal + kWG - 1) / kWG;
        for (int t = 0; t < ntile[b]; ++t) wg_block[tile0[b] + t] = b;
        parallel_for(B.nface, [&](size_t f0, size_t f1) {
            for (int f = (int)f0; f < (int)f1; ++f) {
                const size_t g = global_face(c, b, f);
                for (int d = 0; d < 3; ++d) {
                    S[d * (size_t)NF + g] = B.S[3 * f + d];
                    rL[d * (size_t)NF + g] = B.rL[3 * f + d];
                    rR[d * (size_t)NF + g] = B.rR[3 * f + d];
                    fc[d * (size_t)NF + g] = B.fcenter[3 * f + d];
                }
                area[g] = B.area[f];
            }
        });
        // per cell and face slot (cells write disjoint entries; several host threads)
        std::atomic<int> bad{0};
        parallel_for(B.nreal, [&](size_t q0, size_t q1) {
            for (int q = (int)q0; q < (int)q1; ++q) {
                const int gc = global_cell(c, b, q);
                nf[gc] = B.cell_off[q + 1] - B.cell_off[q];
                vol[gc] = B.volume[q];
                for (int k = 0; k < nf[gc]; ++k) {
                    const int f = B.cell2face[B.cell_off[q] + k];
                    const bool right = B.face2cell[2 * f] != q;
                    const int other = B.cell2cell[B.cell_off[q] + k];
                    nbr[(size_t)k * NR + gc] = global_cell(c, b, other);
                    fid[(size_t)k * NR + gc] = global_face(c, b, f) | (right ? (int)0x80000000u : 0);
                    for (int d = 0; d < 3; ++d)
                        lsqw[(size_t)(k * 3 + d) * NR + gc] = B.lsqw[3 * (size_t)(B.cell_off[q] + k) + d];
                    // geometry copies: S oriented away from this cell, as the gradient uses it
                    // (GreenGauss.cpp:104-118), and the r vector on this cell's side
                    const double* rv = right ? &B.rR[3 * f] : &B.rL[3 * f];
                    const size_t q2 = (size_t)k * NR + gc;
                    for (int d = 0; d < 3; ++d) {
                        const double sd = right ? B.S[3 * f + d] * -1.0 : B.S[3 * f + d];
                        if (d < 2) {
                            sSxy[2 * q2 + d] = sd;
                            sRxy[2 * q2 + d] = rv[d];
                        } else {
                            sSz[q2] = sd;
                            sRz[q2] = rv[d];
                        }
                    }
                    // face-state record of the other cell on this face
                    size_t rec;
                    if (other < B.nreal) {
                        int ko = -1;
                        for (int j = B.cell_off[other]; j < B.cell_off[other + 1]; ++j)
                            if (B.cell2face[j] == f) ko = j - B.cell_off[other];
                        if (ko < 0) {
                            bad |= 1;
                            continue;
                        }
                        rec = (size_t)ko * NR + (size_t)global_cell(c, b, other);
                    } else {
                        rec = ghost_record(nslot, (int)NR, c->goff[b] + (other - B.nreal));
                    }
                    if (rec >= 0x80000000ull) {
                        bad |= 2;
                        continue;
                    }
                    mate[(size_t)k * NR + gc] = (int)rec | (right ? (int)0x80000000u : 0);
                }
            }
        });
        if (bad & 1) return set_error(c, E3D_EINVAL, "upload: inconsistent face connectivity");
        if (bad & 2) return set_error(c, E3D_EINVAL, "mesh too large for face-state records");
        for (size_t j = 0; j < B.bnds.size(); ++j) {
            const bool real_bnd = (int)j < B.n_real_boundaries;
            const BcType t = B.bnds[j].type;
            for (int gl : B.bnds[j].ghosts) {
                const int gi = c->goff[b] + (gl - B.nreal);
                g_in[gi] = global_cell(c, b, B.cell2cell[B.cell_off[gl]]);
                g_face[gi] = global_face(c, b, B.cell2face[B.cell_off[gl]]);
                const int gf = B.cell2face[B.cell_off[gl]];
                // unit normal of the ghost's face: S / |S| with |S| = sqrt(Sx*Sx + Sy*Sy + Sz*Sz), as
                // FluxScheme.cpp:105-112 forms it (equal to S / face area, MetricsInitializer.cpp:546-558)
                const double sx = B.S[3 * gf], sy = B.S[3 * gf + 1], sz = B.S[3 * gf + 2];
                const double norm = std::sqrt(sx * sx + sy * sy + sz * sz);
                for (int d = 0; d < 3; ++d) {
                    g_r[(size_t)d * NG + gi] = B.rR[3 * (size_t)gf + d];
                    g_n[(size_t)d * NG + gi] = B.S[3 * (size_t)gf + d] / norm;
                }
                int kind = 0;
                if (real_bnd) {
                    kind |= GK_REFLECT;
                    if (t == BC_WALL || t == BC_SYMMETRY) kind |= GK_SLIP;
                    else if (t == BC_FARFIELD) kind |= GK_FARFIELD;
                }
                if (t == BC_WALL) {
                    kind |= GK_WALL;
                    wall_ghost.push_back(m.NRp + gi);
                    wall_face.push_back(g_face[gi]);
                }
                g_kind[gi] = kind;
                vol[m.NRp + gi] = B.volume[gl];
            }
        }
        wall_off[b + 1] = (int)wall_ghost.size();
        for (size_t j = 0; j < B.bnds.size(); ++j)
            c->marker_off[b].push_back(c->marker_off[b].back() + (int)B.bnds[j].ghosts.size());
        // interfaces: destination ghosts numbered from GhostCount in topology order
        int k = 0;
        for (size_t i = 0; i < F.itf_origin_block.size(); ++i) {
            const int o = src[b].base + F.itf_origin_block[i];
            for (int q = F.itf_off[i]; q < F.itf_off[i + 1]; ++q, ++k) {
                itf_dst.push_back(global_cell(c, b, c->ghost_count[b] + k));
                itf_org.push_back(global_cell(c, o, F.itf_origin_cells[q]));
            }
        }
    }
    c->n_itf = (int)itf_dst.size();
    // Tet meshes: per-cell node copies and slot codes from which the
    // Green-Gauss kernel rebuilds S and r (tet_geometry.hpp); tet_slot_codes
    // rebuilds every slot on the host with the kernel's own expressions and
    // compares it bit for bit with the stored metrics, and the copies are
    // kept only if every block agrees.
    std::vector<double> cnode;
    std::vector<unsigned> cgeo;
    if (nslot == 4) {
        cnode.assign(12 * NR, 0.0);
        cgeo.assign(NR, 0u);
        bool ok = true;
        std::vector<double> xyz;
        std::vector<unsigned> code;
        for (int b = 0; b < nb && ok; ++b) {
            const HostBlock& B = *src[b].B;
            xyz.assign(12 * (size_t)B.nreal, 0.0);
            code.assign(B.nreal, 0u);
            ok = tet_slot_codes(B, *src[b].R, xyz.data(), code.data()) == 0;
            for (int q = 0; ok && q < B.nreal; ++q) {
                const size_t gc = (size_t)global_cell(c, b, q);
                cgeo[gc] = code[q];
                for (int i = 0; i < 12; ++i)  // pair plane i / 2 holds values 2(i/2), 2(i/2) + 1
                    cnode[2 * ((size_t)(i >> 1) * NR + gc) + (i & 1)] = xyz[12 * (size_t)q + i];
            }
        }
        if (!ok) {
            cnode.clear();
            cgeo.clear();
        }
    }
    // monitor reduction: enough workgroups per block that the largest block's
    // tiles and wall faces are spread over the chip (at most 64 per block)
    {
        int most = 1;
        for (int b = 0; b < nb; ++b)
            most = std::max(most, std::max(ntile[b], wall_off[b + 1] - wall_off[b]));
        c->mon_split = std::min(64, std::max(1, (most + 4 * kWG - 1) / (4 * kWG)));
    }
    int rc = 0;
    const int* cnf;
    (void)cnf;
    int *d_nf, *d_nbr, *d_fid, *d_gin, *d_gface, *d_gkind, *d_wgb, *d_mate;
    double *d_S, *d_area, *d_rL, *d_rR, *d_vol, *d_lsqw, *d_sSxy, *d_sSz, *d_sRxy, *d_sRz, *d_gr, *d_gn;
    if ((rc = upload(c, &d_nf, nf)) || (rc = upload(c, &d_nbr, nbr)) || (rc = upload(c, &d_fid, fid)) ||
        (rc = upload(c, &d_gin, g_in)) || (rc = upload(c, &d_gface, g_face)) || (rc = upload(c, &d_gkind, g_kind)) ||
        (rc = upload(c, &d_wgb, wg_block)) || (rc = upload(c, &d_S, S)) || (rc = upload(c, &d_area, area)) ||
        (rc = upload(c, &d_rL, rL)) || (rc = upload(c, &d_rR, rR)) || (rc = upload(c, &d_vol, vol)) ||
        (rc = upload(c, &d_lsqw, lsqw)) || (rc = upload(c, &c->d_fcenter, fc)) || (rc = upload(c, &c->d_itf_dst, itf_dst)) ||
        (rc = upload(c, &c->d_itf_org, itf_org)) || (rc = upload(c, &c->d_tile0, tile0)) ||
        (rc = upload(c, &c->d_ntile, ntile)) || (rc = upload(c, &c->d_nreal, c->nreal)) ||
        (rc = upload(c, &c->d_wall_off, wall_off)) || (rc = upload(c, &c->d_wall_ghost, wall_ghost)) ||
        (rc = upload(c, &c->d_wall_face, wall_face)) || (rc = upload(c, &d_mate, mate)) ||
        (rc = upload(c, &d_sSxy, sSxy)) || (rc = upload(c, &d_sSz, sSz)) || (rc = upload(c, &d_sRxy, sRxy)) ||
        (rc = upload(c, &d_sRz, sRz)) || (rc = upload(c, &d_gr, g_r)) || (rc = upload(c, &d_gn, g_n)) || (rc = upload(c, &c->d_blk_inst, c->blk_inst)) ||
        (rc = upload(c, &c->d_g_blk, g_blk)) || (rc = upload(c, &c->d_inst_b0, c->inst_b0)) ||
        (rc = dalloc(c, &c->d_inst_fs, 5 * (size_t)c->n_inst)) || (rc = dalloc(c, &c->d_blk_qcmac, (size_t)nb)) ||
        (rc = dalloc(c, &c->d_active, (size_t)c->n_inst)) || (rc = dalloc(c, &c->d_frozen, (size_t)c->n_inst)) ||
        (rc = dalloc(c, &c->d_iter, (size_t)c->n_inst)) || (rc = dalloc(c, &c->d_nactive, 1)) ||
        (rc = dalloc(c, &c->d_rms0, 5 * (size_t)c->n_inst)) || (rc = dalloc(c, &c->d_clock, 1 + (size_t)c->n_inst)) ||
        (rc = dalloc(c, (char**)&c->d_prm, sizeof(InstParam) * (size_t)c->n_inst))) {
        free_device(c);
        return rc;
    }
    m.nf = d_nf; m.nbr = d_nbr; m.fid = d_fid; m.S = d_S; m.area = d_area; m.rL = d_rL; m.rR = d_rR;
    m.vol = d_vol; m.lsqw = d_lsqw; m.g_in = d_gin; m.g_face = d_gface; m.g_kind = d_gkind; m.wg_block = d_wgb;
    m.sSxy = reinterpret_cast<const double2*>(d_sSxy);
    m.sSz = d_sSz;
    m.sRxy = reinterpret_cast<const double2*>(d_sRxy);
    m.sRz = d_sRz;
    m.mate = d_mate; m.g_r = d_gr; m.g_n = d_gn;
    m.cnode = nullptr;
    m.cgeo = nullptr;
    if (!cgeo.empty()) {
        double* d_cnode;
        unsigned* d_cgeo;
        if ((rc = upload(c, &d_cnode, cnode)) || (rc = upload(c, &d_cgeo, cgeo))) {
            free_device(c);
            return rc;
        }
        m.cnode = reinterpret_cast<const double2*>(d_cnode);
        m.cgeo = d_cgeo;
    }
    m.blk_inst = c->d_blk_inst; m.inst_fs = c->d_inst_fs; m.active = nullptr;
    resolve_tile_order(c);
    DState& s = c->s;
    s.jac = nullptr;
    if ((rc = dalloc(c, &s.W, 5 * NA)) || (rc = dalloc(c, &s.Wn, 5 * NA)) || (rc = dalloc(c, &s.W0, 5 * NR)) ||
        (rc = dalloc(c, &s.conv, 5 * NR)) || (rc = dalloc(c, &s.diss, 5 * NR)) || (rc = dalloc(c, &s.grad, 15 * NA)) ||
        (rc = dalloc(c, &s.lim, 5 * NA)) || (rc = dalloc(c, &s.Wf, 5 * (NS + NG))) || (rc = dalloc(c, &s.dt, NR)) ||
        (rc = dalloc(c, &s.specrad, NR)) || (rc = dalloc(c, &s.part, 8 * (NR / 64 + 1))) ||
        (rc = dalloc(c, &c->d_mon, 16 * (size_t)nb)) ||
        (rc = dalloc(c, &c->d_mpart, 10 * (size_t)nb * c->mon_split))) {
        free_device(c);
        return rc;
    }
    hipError_t e = hipHostMalloc((void**)&c->h_mon, sizeof(double) * 16 * nb, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->h_nactive, sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) {
        free_device(c);
        return set_error(c, E3D_ENOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return set_error(c, E3D_EHIP, std::string("upload: ") + hipGetErrorString(e));
    c->have_mesh = true;
    c->iter = 0;
    c->started = false;
    return sync_instances(c);
}

int e3d_upload_mesh(e3d_ctx* c, const e3d_mesh* mesh) {
    if (!c || !mesh) return set_error(c, E3D_EINVAL, "e3d_upload_mesh: null argument");
    c->batch = false;
    c->inst.assign(1, e3d_instance{});
    return upload_meshes(c, {mesh});
}

int e3d_nblocks(const e3d_ctx* c) { return c ? c->nb : E3D_EINVAL; }

int e3d_init_freestream(e3d_ctx* c) {
    if (!c || !c->have_mesh || !c->have_params) return set_error(c, E3D_ESTATE, "e3d_init_freestream: need mesh and params");
    const size_t NA = c->m.NA;
    std::vector<double> w(5 * NA, 0.0);  // one host buffer, alive until the copies complete
    for (int b = 0; b < c->nb; ++b) {  // every cell of a block gets its instance's freestream
        const double* fs = block_fs(c, b);
        const Range r = block_range(c, b);
        for (int v = 0; v < 5; ++v) {
            std::fill(w.begin() + v * NA + r.c0, w.begin() + v * NA + r.c1, fs[v]);
            std::fill(w.begin() + v * NA + c->m.NRp + r.g0, w.begin() + v * NA + c->m.NRp + r.g1, fs[v]);
        }
    }
    HIP_TRY(c, hipMemcpyAsync(c->s.W, w.data(), sizeof(double) * 5 * NA, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->s.Wn, w.data(), sizeof(double) * 5 * NA, hipMemcpyHostToDevice, c->stream));
    // batched-run state: every instance active at iteration 0
    std::vector<int> ones(c->n_inst, 1), zeros(c->n_inst, 0);
    HIP_TRY(c, hipMemcpyAsync(c->d_active, ones.data(), sizeof(int) * c->n_inst, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_frozen, zeros.data(), sizeof(int) * c->n_inst, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(c->d_iter, zeros.data(), sizeof(int) * c->n_inst, hipMemcpyHostToDevice, c->stream));
    std::vector<double> ref(5 * (size_t)c->n_inst);
    for (int i = 0; i < c->n_inst; ++i)
        for (int v = 0; v < 5; ++v) ref[5 * i + v] = c->inst[i].rms_ref[v];
    HIP_TRY(c, hipMemcpyAsync(c->d_rms0, ref.data(), sizeof(double) * ref.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->iter = 0;
    c->started = false;
    c->clock_started = false;
    return E3D_OK;
}

namespace {
// (device array, components, real-only?) of an E3D_ARR_* selector
int array_desc(e3d_ctx* c, int which, double** base, int* ncomp, bool* real_only) {
    switch (which) {
        case E3D_ARR_W: *base = c->s.W; *ncomp = 5; *real_only = false; break;
        case E3D_ARR_W0: *base = c->s.W0; *ncomp = 5; *real_only = true; break;
        case E3D_ARR_CONV: *base = c->s.conv; *ncomp = 5; *real_only = true; break;
        case E3D_ARR_DISS: *base = c->s.diss; *ncomp = 5; *real_only = true; break;
        case E3D_ARR_GRAD: *base = c->s.grad; *ncomp = 15; *real_only = false; break;
        case E3D_ARR_LIM: *base = c->s.lim; *ncomp = 5; *real_only = false; break;
        case E3D_ARR_DT: *base = c->s.dt; *ncomp = 1; *real_only = true; break;
        case E3D_ARR_SPECRAD: *base = c->s.specrad; *ncomp = 1; *real_only = true; break;
        default: return set_error(c, E3D_EINVAL, "unknown array selector");
    }
    return E3D_OK;
}

int copy_array(e3d_ctx* c, int b, int which, double* host, bool to_device) {
    if (!c || !c->have_mesh) return set_error(c, E3D_ESTATE, "no mesh uploaded");
    if (b < 0 || b >= c->nb || !host) return set_error(c, E3D_EINVAL, "bad block or null buffer");
    double* base;
    int nc;
    bool real_only;
    int rc = array_desc(c, which, &base, &nc, &real_only);
    if (rc) return rc;
    const size_t stride = real_only ? (size_t)c->m.NRp : (size_t)c->m.NA;
    const int nr = c->nreal[b], ng = c->nall[b] - c->nreal[b];
    const size_t n_host = real_only ? nr : (size_t)c->nall[b];
    const hipMemcpyKind kind = to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    // real cells in device order: staged through a host buffer and permuted to
    // or from the reference's numbering
    const std::vector<int>& pos = c->cell_pos[b];
    std::vector<double> stage(pos.empty() ? 0 : (size_t)nc * nr);
    for (int q = 0; q < nc; ++q) {
        double* d_real = base + q * stride + c->roff[b];
        double* h = host + q * n_host;
        double* hr = pos.empty() ? h : stage.data() + (size_t)q * nr;
        if (to_device) {
            if (!pos.empty())
                for (int i = 0; i < nr; ++i) hr[pos[i]] = h[i];
            HIP_TRY(c, hipMemcpyAsync(d_real, hr, sizeof(double) * nr, kind, c->stream));
        } else {
            HIP_TRY(c, hipMemcpyAsync(hr, d_real, sizeof(double) * nr, kind, c->stream));
        }
        if (!real_only && ng > 0) {
            double* d_g = base + q * stride + c->m.NRp + c->goff[b];
            if (to_device) HIP_TRY(c, hipMemcpyAsync(d_g, h + nr, sizeof(double) * ng, kind, c->stream));
            else HIP_TRY(c, hipMemcpyAsync(h + nr, d_g, sizeof(double) * ng, kind, c->stream));
        }
    }
    if (to_device && which == E3D_ARR_W) {  // keep the fused path's second buffer in step
        for (int q = 0; q < 5; ++q) {
            HIP_TRY(c, hipMemcpyAsync(c->s.Wn + q * stride + c->roff[b], c->s.W + q * stride + c->roff[b],
                                      sizeof(double) * nr, hipMemcpyDeviceToDevice, c->stream));
            if (ng > 0)
                HIP_TRY(c, hipMemcpyAsync(c->s.Wn + q * stride + c->m.NRp + c->goff[b],
                                          c->s.W + q * stride + c->m.NRp + c->goff[b], sizeof(double) * ng,
                                          hipMemcpyDeviceToDevice, c->stream));
        }
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (!to_device && !pos.empty())
        for (int q = 0; q < nc; ++q) {
            const double* hr = stage.data() + (size_t)q * nr;
            double* h = host + q * n_host;
            for (int i = 0; i < nr; ++i) h[i] = hr[pos[i]];
        }
    return E3D_OK;
}
}  // namespace

int e3d_get_array(e3d_ctx* c, int block, int which, double* out) { return copy_array(c, block, which, out, false); }
int e3d_set_array(e3d_ctx* c, int block, int which, const double* in) {
    return copy_array(c, block, which, const_cast<double*>(in), true);
}

// ------------------------------------------------------------ granular ops

int e3d_save_w0(e3d_ctx* c, int block) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    launch_save_w0(block_range(c, block), c->m, c->s, 1. / (c->params.gamma - 1.0), c->stream);
    return launch_status(c);
}

int e3d_spectral_radius(e3d_ctx* c, int block) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    launch_spectral_radius(block_range(c, block), c->m, c->s, c->params.gamma, c->stream);
    return launch_status(c);
}

int e3d_timestep(e3d_ctx* c, int block) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    launch_timestep(block_range(c, block), c->m, c->s, c->params.cfl, c->stream);
    return launch_status(c);
}

int e3d_gradients(e3d_ctx* c, int block, int kind) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    if (kind != E3D_GRAD_GREEN_GAUSS && kind != E3D_GRAD_LEAST_SQUARES) return set_error(c, E3D_EINVAL, "unknown gradient");
    launch_gradient(block_range(c, block), c->m, c->s, kind, c->stream);
    return launch_status(c);
}

int e3d_limiters(e3d_ctx* c, int block, int kind, double param) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    if (kind != E3D_LIM_BARTH_JESPERSEN && kind != E3D_LIM_VENKATAKRISHNAN) return set_error(c, E3D_EINVAL, "unknown limiter");
    const bool vk = kind == E3D_LIM_VENKATAKRISHNAN;
    launch_limiter(block_range(c, block), c->m, c->s, vk, vk ? param * param * param : 0.0, vk ? 0.0 : param, c->stream);
    return launch_status(c);
}

int e3d_interpolation(e3d_ctx* c, int block) {
    int rc = check_ready(c, block);
    if (rc || c->params.order != 2) return rc;  // FirstOrder::computeInterpolation is a no-op
    if ((rc = e3d_gradients(c, block, c->params.gradient))) return rc;
    return e3d_limiters(c, block, c->params.limiter,
                        c->params.limiter == E3D_LIM_VENKATAKRISHNAN ? c->params.k : c->params.omega);
}

int e3d_flux_conv(e3d_ctx* c, int block, int scheme) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    if (scheme != E3D_FLUX_ROE && scheme != E3D_FLUX_AUSM) return set_error(c, E3D_EINVAL, "unknown flux scheme");
    launch_flux_conv(block_range(c, block), c->m, c->s, scheme, c->gas, c->stream);
    return launch_status(c);
}

int e3d_flux_diss(e3d_ctx* c, int block, int scheme, double beta) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    if (scheme != E3D_FLUX_ROE && scheme != E3D_FLUX_AUSM) return set_error(c, E3D_EINVAL, "unknown flux scheme");
    launch_flux_diss(block_range(c, block), c->m, c->s, scheme, c->gas, beta, c->stream);
    return launch_status(c);
}

int e3d_smooth_residual(e3d_ctx* c, int block) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    if (!c->s.jac && (rc = dalloc(c, &c->s.jac, 10 * (size_t)c->m.NRp))) return rc;
    launch_smooth(block_range(c, block), c->m, c->s, c->stream);
    return launch_status(c);
}

int e3d_residual(e3d_ctx* c, int block, int stage) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    if (stage < 0 || stage >= c->params.stages) return set_error(c, E3D_EINVAL, "stage out of range");
    if ((rc = e3d_interpolation(c, block)) || (rc = e3d_flux_conv(c, block, c->params.flux)) ||
        (rc = e3d_flux_diss(c, block, c->params.flux, c->beta[stage])))
        return rc;
    return c->params.smoothing ? e3d_smooth_residual(c, block) : E3D_OK;  // NoSmoother otherwise
}

int e3d_update(e3d_ctx* c, int block, int stage) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    if (stage < 0 || stage >= c->params.stages) return set_error(c, E3D_EINVAL, "stage out of range");
    launch_update(block_range(c, block), c->m, c->s, c->alpha[stage], c->gas.gm1, c->stream);
    return launch_status(c);
}

int e3d_bc(e3d_ctx* c, int block) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    for (int b = block < 0 ? 0 : block; b < (block < 0 ? c->nb : block + 1); ++b)
        launch_bc(block_range(c, b), c->m, c->s, c->params.gamma, block_fs(c, b), c->stream);
    return launch_status(c);
}

int e3d_bc_marker(e3d_ctx* c, int block, int marker) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    if (block < 0 || marker < 0 || marker >= (int)c->marker_off[block].size() - 1)
        return set_error(c, E3D_EINVAL, "e3d_bc_marker: bad block or marker");
    Range rg = block_range(c, block);
    rg.c0 = rg.c1 = 0;
    rg.g1 = rg.g0 + c->marker_off[block][marker + 1];
    rg.g0 = rg.g0 + c->marker_off[block][marker];
    launch_bc(rg, c->m, c->s, c->params.gamma, block_fs(c, block), c->stream);  // kernels skip ghosts without a BC
    return launch_status(c);
}

int e3d_runge_kutta(e3d_ctx* c, int block) {
    int rc = check_ready(c, block);
    if (rc) return rc;
    for (int st = 0; st < c->params.stages; ++st)
        if ((rc = e3d_residual(c, block, st)) || (rc = e3d_update(c, block, st)) || (rc = e3d_bc(c, block))) return rc;
    return E3D_OK;
}

int e3d_exchange(e3d_ctx* c) {
    int rc = check_ready(c, -1);
    if (rc) return rc;
    c->started = true;
    return exchange(c, 1);
}

int e3d_monitors(e3d_ctx* c, e3d_monitor* out) {
    int rc = check_ready(c, -1);
    if (rc) return rc;
    if (c->batch) return set_error(c, E3D_ESTATE, "e3d_monitors: a batch has per-instance monitors (e3d_batch_history)");
    launch_rms_partials(c->m, c->s, c->stream);
    if ((rc = launch_status(c))) return rc;
    return finish_monitor(c, out);
}

int e3d_reset_monitors(e3d_ctx* c) {
    if (!c) return set_error(nullptr, E3D_EINVAL, "null context");
    c->iter = 0;
    c->started = false;
    for (double& r : c->rms0) r = 0.0;
    return E3D_OK;
}

int e3d_monitor_state(const e3d_ctx* c, int* iter, double rms0[5]) {
    if (!c || !iter || !rms0) return set_error(nullptr, E3D_EINVAL, "e3d_monitor_state: null argument");
    *iter = c->iter;
    for (int v = 0; v < 5; ++v) rms0[v] = c->rms0[v];
    return E3D_OK;
}

int e3d_set_monitor_state(e3d_ctx* c, int iter, const double rms0[5]) {
    if (!c || !rms0) return set_error(c, E3D_EINVAL, "e3d_set_monitor_state: null argument");
    if (iter < 0) return set_error(c, E3D_EINVAL, "e3d_set_monitor_state: negative iteration");
    c->iter = iter;
    for (int v = 0; v < 5; ++v) c->rms0[v] = rms0[v];
    return E3D_OK;
}

int e3d_request_stop(e3d_ctx* c) {
    if (!c) return set_error(nullptr, E3D_EINVAL, "null context");
    c->stop_request.store(1);
    return E3D_OK;
}

// ------------------------------------------------------------- fused path

// Solver::solve with one granular operator call per reference method, block by
// block (Solver/src/Solver.cpp:76-98). Used by e3d_iterate when the residual
// smoother is on: its Jacobi sweep reads the neighbours' residuals, which the
// fused stage kernels do not hold.
static int iterate_granular(e3d_ctx* c, int n, e3d_monitor* history) {
    int rc;
    if (!c->started) {  // Solver.cpp:48
        if ((rc = exchange(c, 1))) return rc;
        c->started = true;
    }
    int done = 0;
    for (int it = 0; it < n; ++it) {
        if (c->stop_request.exchange(0)) break;
        for (int b = 0; b < c->nb; ++b)
            if ((rc = e3d_save_w0(c, b)) || (rc = e3d_spectral_radius(c, b)) || (rc = e3d_timestep(c, b)) ||
                (rc = e3d_runge_kutta(c, b)))
                return rc;
        if ((rc = exchange(c, 1))) return rc;  // Solver.cpp:94
        e3d_monitor mon;
        rc = e3d_monitors(c, &mon);
        if (history) history[done] = mon;
        ++done;
        if (rc) return rc;
        if (mon.stop) break;
    }
    return done;
}

int e3d_iterate(e3d_ctx* c, int n, e3d_monitor* history) {
    int rc = check_ready(c, -1);
    if (rc) return rc;
    if (n < 0) return set_error(c, E3D_EINVAL, "negative iteration count");
    HIP_TRY(c, hipSetDevice(c->device));
    if (c->batch) return set_error(c, E3D_ESTATE, "e3d_iterate: a batch advances with e3d_batch_run");
    if (c->params.smoothing) return iterate_granular(c, n, history);
    const e3d_params& p = c->params;
    const bool order2 = p.order == 2;
    const double k = p.k;
    if (!c->started) {  // Solver.cpp:48
        if ((rc = exchange(c, 0))) return rc;
        c->started = true;
    }
    int done = 0;
    for (int it = 0; it < n; ++it) {
        if (c->stop_request.exchange(0)) break;
        for (int st = 0; st < p.stages; ++st) {
            size_t ev = 0;
            if (order2) {
                ev_begin(c, KI_INTERP, &ev);
                launch_interp_fused(c->m, c->s, p.gradient, p.limiter == E3D_LIM_VENKATAKRISHNAN, k * k * k, p.omega,
                                    c->tile_cells, c->slot_geometry != E3D_GEOMETRY_STREAMED, c->stream);
                ev_end(c, ev);
            }
            ev_begin(c, st == 0 ? KI_FLUX_UPDATE0 : c->beta[st] != 0.0 ? KI_FLUX_UPDATE : KI_FLUX_UPDATE_ND, &ev);
            launch_flux_update(c->m, c->s, p.flux, order2, c->beta[st] != 0.0, st == p.stages - 1, st == 0, c->gas,
                               c->beta[st], c->alpha[st], p.cfl, 1. / (p.gamma - 1.0), c->tile_cells, c->stream);
            ev_end(c, ev);
            if ((rc = launch_status(c))) return rc;
            std::swap(c->s.W, c->s.Wn);
        }
        if ((rc = exchange(c, 0))) return rc;  // Solver.cpp:94
        e3d_monitor mon;
        rc = finish_monitor(c, &mon);
        if (history) history[done] = mon;
        ++done;
        if (rc) return rc;
        if (mon.stop) break;
    }
    return done;
}

// ---------------------------------------------------------------- batches

static_assert(sizeof(MonRow) == sizeof(e3d_monitor), "monitor row layout");

int e3d_upload_batch(e3d_ctx* c, const e3d_mesh* const* meshes, const e3d_instance* inst, int n) {
    if (!c || !meshes || !inst || n < 1) return set_error(c, E3D_EINVAL, "e3d_upload_batch: bad arguments");
    std::vector<const e3d_mesh*> ms(meshes, meshes + n);
    int cap = 1;
    for (int i = 0; i < n; ++i) {
        if (!ms[i]) return set_error(c, E3D_EINVAL, "e3d_upload_batch: null mesh " + std::to_string(i));
        if (inst[i].max_iter < 1) return set_error(c, E3D_EINVAL, "e3d_upload_batch: max_iter must be >= 1");
        cap = std::max(cap, std::min(inst[i].max_iter, 200000));
    }
    c->batch = true;
    c->inst.assign(inst, inst + n);
    int rc = upload_meshes(c, ms);
    if (rc) return rc;
    c->hist_cap = cap;
    if ((rc = dalloc(c, (char**)&c->d_hist, sizeof(MonRow) * (size_t)n * cap))) {
        free_device(c);
        return rc;
    }
    c->m.active = c->d_active;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return E3D_OK;
}

int e3d_batch_size(const e3d_ctx* c) { return c ? c->n_inst : E3D_EINVAL; }

int e3d_batch_blocks(const e3d_ctx* c, int i, int* first, int* count) {
    if (!c || !first || !count || i < 0 || i >= c->n_inst || !c->have_mesh)
        return set_error(nullptr, E3D_EINVAL, "e3d_batch_blocks: bad argument");
    *first = c->inst_b0[i];
    *count = c->inst_b0[i + 1] - c->inst_b0[i];
    return E3D_OK;
}

// One iteration of a batch: the stage kernels over every active instance, the
// interface exchange, per-block monitor rows, then the per-instance rows, stop
// tests and freezing. Launches only, so it can be captured.
static int batch_iteration(e3d_ctx* c) {
    const e3d_params& p = c->params;
    const bool order2 = p.order == 2;
    int rc;
    for (int st = 0; st < p.stages; ++st) {
        size_t ev = 0;
        if (order2) {
            ev_begin(c, KI_INTERP, &ev);
            launch_interp_fused(c->m, c->s, p.gradient, p.limiter == E3D_LIM_VENKATAKRISHNAN, p.k * p.k * p.k, p.omega,
                                c->tile_cells, c->slot_geometry != E3D_GEOMETRY_STREAMED, c->stream);
            ev_end(c, ev);
        }
        ev_begin(c, st == 0 ? KI_FLUX_UPDATE0 : c->beta[st] != 0.0 ? KI_FLUX_UPDATE : KI_FLUX_UPDATE_ND, &ev);
        launch_flux_update(c->m, c->s, p.flux, order2, c->beta[st] != 0.0, st == p.stages - 1, st == 0, c->gas,
                           c->beta[st], c->alpha[st], p.cfl, 1. / (p.gamma - 1.0), c->tile_cells, c->stream);
        ev_end(c, ev);
        if ((rc = launch_status(c))) return rc;
        std::swap(c->s.W, c->s.Wn);
    }
    if ((rc = exchange(c, 0))) return rc;
    size_t ev = 0;
    ev_begin(c, KI_MONITORS, &ev);
    launch_block_monitors(c->nb, c->mon_split, c->d_tile0, c->d_ntile, c->d_nreal, c->d_wall_off, c->d_wall_ghost,
                          c->d_wall_face, c->d_fcenter, c->m, c->s.W, c->s.part, c->d_blk_qcmac, c->d_mpart, c->d_mon,
                          c->stream);
    launch_batch_reduce(c->n_inst, c->d_inst_b0, c->d_mon, c->d_prm, c->d_iter, c->d_rms0, c->d_active, c->d_frozen,
                        c->d_hist, c->hist_cap, c->d_nactive, c->d_clock + 1, c->stream);
    launch_batch_freeze(c->m, c->d_frozen, c->d_g_blk, c->s.W, c->s.Wn, c->stream);
    ev_end(c, ev);
    return launch_status(c);
}

static int batch_active(e3d_ctx* c, int* n_active) {
    HIP_TRY(c, hipMemcpyAsync(c->h_nactive, c->d_nactive, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    ev_collect(c);
    *n_active = *c->h_nactive;
    return E3D_OK;
}

static int batch_capture(e3d_ctx* c) {
    drop_graph(c);
    double* w_start = c->s.W;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = batch_iteration(c);
    if (!rc) rc = batch_iteration(c);  // two iterations: an even number of buffer swaps
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return set_error(c, E3D_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (c->s.W != w_start) return set_error(c, E3D_EHIP, "batch graph: unbalanced state buffers");
    c->graph = g;
    HIP_TRY(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    c->graph_W = w_start;
    return E3D_OK;
}

int e3d_batch_run(e3d_ctx* c, int max_it, int use_graph, e3d_batch_status* status) {
    int rc = check_ready(c, -1);
    if (rc) return rc;
    if (!c->batch) return set_error(c, E3D_ESTATE, "e3d_batch_run: no batch uploaded (e3d_upload_batch)");
    if (c->params.smoothing) return set_error(c, E3D_EINVAL, "e3d_batch_run: residual smoothing is not supported");
    if (max_it < 0) return set_error(c, E3D_EINVAL, "e3d_batch_run: negative iteration count");
    HIP_TRY(c, hipSetDevice(c->device));
    if (!c->started) {  // Solver.cpp:48
        if ((rc = exchange(c, 0))) return rc;
        c->started = true;
    }
    if (!c->clock_started) {
        launch_clock_mark(c->d_clock, c->stream);
        c->clock_started = true;
    }
    int done = 0, n_active = 1;
    if (use_graph) {
        constexpr int kReplays = 8;  // graph launches (two iterations each) between host checks
        if (c->gexec && c->graph_W != c->s.W) drop_graph(c);
        if (!c->gexec && max_it >= 2 && (rc = batch_capture(c))) return rc;
        while (n_active > 0 && max_it - done >= 2) {
            const int k = std::min(kReplays, (max_it - done) / 2);
            for (int r = 0; r < k; ++r) HIP_TRY(c, hipGraphLaunch(c->gexec, c->stream));
            done += 2 * k;
            if ((rc = batch_active(c, &n_active))) return rc;
        }
    }
    while (n_active > 0 && done < max_it) {
        if ((rc = batch_iteration(c))) return rc;
        ++done;
        if ((rc = batch_active(c, &n_active))) return rc;
    }
    if (status) {
        std::vector<int> it(c->n_inst), act(c->n_inst);
        std::vector<long long> clk(c->n_inst + 1);
        HIP_TRY(c, hipMemcpy(it.data(), c->d_iter, sizeof(int) * c->n_inst, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(act.data(), c->d_active, sizeof(int) * c->n_inst, hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(clk.data(), c->d_clock, sizeof(long long) * clk.size(), hipMemcpyDeviceToHost));
        int khz = 0;
        HIP_TRY(c, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
        for (int i = 0; i < c->n_inst; ++i) {
            status[i].iterations = it[i];
            status[i].stopped = act[i] ? 0 : 1;
            status[i].stop_seconds = act[i] || khz <= 0 ? 0.0 : (double)(clk[1 + i] - clk[0]) / (1e3 * khz);
            status[i].last = e3d_monitor{};
            if (it[i] > 0 && it[i] <= c->hist_cap)
                HIP_TRY(c, hipMemcpy(&status[i].last, (MonRow*)c->d_hist + (size_t)i * c->hist_cap + it[i] - 1,
                                     sizeof(MonRow), hipMemcpyDeviceToHost));
        }
    }
    return done;
}

int e3d_batch_history(e3d_ctx* c, int i, e3d_monitor* rows, int max_rows) {
    int rc = check_ready(c, -1);
    if (rc) return rc;
    if (!c->batch || i < 0 || i >= c->n_inst || !rows || max_rows < 0)
        return set_error(c, E3D_EINVAL, "e3d_batch_history: bad argument");
    int it = 0;
    HIP_TRY(c, hipStreamSynchronize(c->stream));  // rows written by work still queued on the stream
    HIP_TRY(c, hipMemcpy(&it, c->d_iter + i, sizeof(int), hipMemcpyDeviceToHost));
    const int n = std::min(std::min(it, c->hist_cap), max_rows);
    if (n > 0)
        HIP_TRY(c, hipMemcpy(rows, (MonRow*)c->d_hist + (size_t)i * c->hist_cap, sizeof(MonRow) * n,
                             hipMemcpyDeviceToHost));
    return n;
}

static_assert(E3D_TILES_AUTO == -1 && E3D_TILES_LINEAR == 0 && E3D_TILES_XCD_CONTIGUOUS == 1 && E3D_TILES_XCD_CHUNKS == 2,
              "tile orders as tile_of numbers them");
static_assert(E3D_ORDER_REFERENCE == ORDER_REFERENCE && E3D_ORDER_RCM == ORDER_RCM &&
                  E3D_ORDER_MORTON == ORDER_MORTON,
              "ordering modes");
int e3d_set_ordering(e3d_ctx* c, int mode) {
    if (!c) return set_error(nullptr, E3D_EINVAL, "null context");
    if (mode != E3D_ORDER_REFERENCE && mode != E3D_ORDER_RCM && mode != E3D_ORDER_MORTON)
        return set_error(c, E3D_EINVAL, "e3d_set_ordering: unknown mode " + std::to_string(mode));
    c->ordering = mode;
    return E3D_OK;
}

int e3d_set_tile_order(e3d_ctx* c, int order) {
    if (!c) return set_error(nullptr, E3D_EINVAL, "null context");
    if (order < E3D_TILES_AUTO || order > E3D_TILES_XCD_CHUNKS)
        return set_error(c, E3D_EINVAL, "e3d_set_tile_order: order must be -1, 0, 1 or 2");
    c->tile_pref = order;
    resolve_tile_order(c);
    drop_graph(c);
    return E3D_OK;
}

int e3d_set_tile_cells(e3d_ctx* c, int cells) {
    if (!c) return set_error(nullptr, E3D_EINVAL, "null context");
    if (cells != 64 && cells != 128 && cells != 256)
        return set_error(c, E3D_EINVAL, "e3d_set_tile_cells: cells must be 64, 128 or 256");
    c->tile_cells = cells;
    drop_graph(c);
    return E3D_OK;
}

int e3d_set_slot_geometry(e3d_ctx* c, int mode) {
    if (!c) return set_error(nullptr, E3D_EINVAL, "null context");
    if (mode != E3D_GEOMETRY_AUTO && mode != E3D_GEOMETRY_STREAMED)
        return set_error(c, E3D_EINVAL, "e3d_set_slot_geometry: mode must be -1 or 0");
    c->slot_geometry = mode;
    drop_graph(c);
    return E3D_OK;
}

int e3d_slot_geometry_nodes(const e3d_ctx* c) {
    if (!c) return E3D_EINVAL;
    return c->slot_geometry != E3D_GEOMETRY_STREAMED && c->have_mesh && c->m.cnode && c->m.cgeo &&
           c->have_params && c->params.gradient == E3D_GRAD_GREEN_GAUSS && c->params.order == 2;
}

// ------------------------------------------------------------- measurement

int e3d_enable_timing(e3d_ctx* c, int on) {
    if (!c) return set_error(nullptr, E3D_EINVAL, "null context");
    c->timing = on != 0;
    for (int i = 0; i < KI_COUNT; ++i) {
        c->t_ms[i] = 0.0;
        c->t_n[i] = 0;
    }
    c->ev_used = 0;
    return E3D_OK;
}

int e3d_kernel_time(e3d_ctx* c, const char* name, double* avg_ms, int* launches) {
    if (!c || !name || !avg_ms || !launches) return set_error(c, E3D_EINVAL, "e3d_kernel_time: null argument");
    for (int i = 0; i < KI_COUNT; ++i)
        if (std::strcmp(name, kKernelNames[i]) == 0) {
            *launches = c->t_n[i];
            *avg_ms = c->t_n[i] ? c->t_ms[i] / c->t_n[i] : 0.0;
            return E3D_OK;
        }
    return set_error(c, E3D_EINVAL, std::string("unknown kernel name ") + name);
}

int e3d_sync(e3d_ctx* c) {
    if (!c) return set_error(nullptr, E3D_EINVAL, "null context");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    ev_collect(c);
    return E3D_OK;
}

}  // extern "C"
