// Batched runs: many independent meshes ("instances", each with its own flow
// conditions) share one global cell space and advance together, every stage
// one launch over all of them. After each iteration one small kernel forms
// every instance's monitor row from its blocks' rows, applies the reference's
// stop test per instance and clears the instance's active flag, so the stage
// kernels skip its tiles from then on; a second kernel copies the final state
// of an instance that just stopped into the other state buffer, so that both
// buffers hold it whatever the buffer parity of later stages. Both run on the
// device, so the iteration can be captured in a hipGraph and replayed without
// host round trips.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kcommon.hpp"

namespace e3d {

// One thread per instance: sums its blocks' rows (rms x5, force/moment sums x5, in
// block order as BlockCommunicator::getGlobal, MPI/src/BlockCommunicator.cpp:341-369),
// the drop against its iteration-0 norms, the stop test (PostProcessing.cpp:59-83;
// a non-finite norm also stops), the history row, and the active/frozen flags.
__global__ void k_batch_reduce(int n_inst, const int* inst_b0, const double* blk_rows, const InstParam* prm,
                               int* iter, double* rms0, int* active, int* frozen, MonRow* hist, int hist_cap,
                               int* n_active, long long* stop_clock) {
    __shared__ int cnt[kWG];
    int mine = 0;
    for (int i = threadIdx.x; i < n_inst; i += blockDim.x) {
        frozen[i] = 0;
        if (!active[i]) continue;
        double sum[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        const InstParam p = prm[i];
        for (int b = inst_b0[i]; b < inst_b0[i + 1]; ++b) {
            const double* r = blk_rows + (size_t)b * 16;
            for (int v = 0; v < 5; ++v) sum[v] += r[v];
            const double cl = r[5], cd = r[6];
            sum[5] += cl * p.cos_a - cd * p.sin_a;
            sum[6] += cl * p.sin_a + cd * p.cos_a;
            sum[7] += r[7];
            sum[8] += r[8];
            sum[9] += r[9];
        }
        const int it = iter[i];
        if (it == 0 && !p.has_ref)
            for (int v = 0; v < 5; ++v) rms0[5 * i + v] = sum[v];
        MonRow row;
        row.iter = it;
        bool finite = true;
        for (int v = 0; v < 5; ++v) {
            row.rms[v] = sum[v];
            row.drop[v] = log10(sum[v]) - log10(rms0[5 * i + v]);
            finite = finite && isfinite(sum[v]);
        }
        row.cl = sum[5];
        row.cd = sum[6];
        row.cmx = sum[7];
        row.cmy = sum[8];
        row.cmz = sum[9];
        const bool stop = (it + 1 == p.max_iter) || (row.drop[0] <= p.criterion) || !finite;
        row.stop = stop ? 1 : 0;
        if (it < hist_cap) hist[(size_t)i * hist_cap + it] = row;
        iter[i] = it + 1;
        if (stop) {
            active[i] = 0;
            frozen[i] = 1;
            stop_clock[i] = wall_clock64();
        } else {
            ++mine;
        }
    }
    cnt[threadIdx.x] = mine;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) cnt[threadIdx.x] += cnt[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *n_active = cnt[0];
}

// Copies W into Wn for every cell (real and ghost) of the instances that stopped in
// this iteration. g_blk: block of each ghost.
__global__ void __launch_bounds__(kWG) k_batch_freeze(DMesh m, const int* frozen, const int* g_blk, const double* W,
                                                      double* Wn) {
    const int x = blockIdx.x * kWG + threadIdx.x;
    if (x >= m.NA) return;
    const int b = x < m.NRp ? m.wg_block[x / kWG] : g_blk[x - m.NRp];
    if (!frozen[m.blk_inst[b]]) return;
#pragma unroll
    for (int v = 0; v < 5; ++v) Wn[(size_t)v * m.NA + x] = W[(size_t)v * m.NA + x];
}

// Device wall clock (constant rate) at the start of a run.
__global__ void k_clock_mark(long long* t) { *t = wall_clock64(); }

void launch_batch_reduce(int n_inst, const int* inst_b0, const double* blk_rows, const void* prm, int* iter,
                         double* rms0, int* active, int* frozen, void* hist, int hist_cap, int* n_active,
                         long long* stop_clock, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_reduce, dim3(1), dim3(kWG), 0, st, n_inst, inst_b0, blk_rows, (const InstParam*)prm,
                       iter, rms0, active, frozen, (MonRow*)hist, hist_cap, n_active, stop_clock);
}

void launch_clock_mark(long long* t, hipStream_t st) { hipLaunchKernelGGL(k_clock_mark, dim3(1), dim3(1), 0, st, t); }

void launch_batch_freeze(const DMesh& m, const int* frozen, const int* g_blk, const double* W, double* Wn,
                         hipStream_t st) {
    hipLaunchKernelGGL(k_batch_freeze, dim3((m.NA + kWG - 1) / kWG), dim3(kWG), 0, st, m, frozen, g_blk, W, Wn);
}


}  // namespace e3d
