// HBM streaming ceiling: how fast this GPU moves bytes for a given mix of
// reads and writes, with the access pattern of the stage kernels at its best
// (every lane a 16-byte load or store, consecutive lanes consecutive
// addresses, nontemporal hints). A launch reads NR planes of n 16-byte
// elements and writes NW planes; each written value depends on every value
// read at the same index, so no load can be dropped. No counterpart in the
// reference: it only calibrates the roofline fractions of the stage kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/euler3d.h"
#include "abi_internal.hpp"

namespace {

typedef double dbl2 __attribute__((ext_vector_type(2)));

// U elements per thread, in a grid of n / (256 U) workgroups: thread t of
// workgroup b handles elements (b * U + u) * 256 + t, u < U (each u a
// coalesced 4 KiB run of the workgroup).
template <int NR, int NW, int U>
__global__ void __launch_bounds__(256) k_stream(const dbl2* __restrict__ in, dbl2* __restrict__ out, long n,
                                                dbl2* __restrict__ sink) {
    const long base = (long)blockIdx.x * U * 256 + threadIdx.x;
    dbl2 acc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) acc[u] = dbl2{0.0, 0.0};
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long i = base + (long)u * 256;
            if (i < n) acc[u] += __builtin_nontemporal_load(in + (long)r * n + i);
        }
    if (NW == 0) {
        dbl2 t = dbl2{0.0, 0.0};
#pragma unroll
        for (int u = 0; u < U; ++u) t += acc[u];
        if (t.x == -1.0) sink[threadIdx.x] = t;  // never true for the all-zero input: keeps the loads
        return;
    }
#pragma unroll
    for (int w = 0; w < NW; ++w)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long i = base + (long)u * 256;
            if (i < n) __builtin_nontemporal_store(acc[u] + dbl2{(double)w, 0.0}, out + (long)w * n + i);
        }
}

template <int NR, int NW>
void launch(int u, const dbl2* in, dbl2* out, long n, dbl2* sink, hipStream_t st) {
    if (u == 1) hipLaunchKernelGGL((k_stream<NR, NW, 1>), dim3((n + 255) / 256), dim3(256), 0, st, in, out, n, sink);
    else hipLaunchKernelGGL((k_stream<NR, NW, 4>), dim3((n + 1023) / 1024), dim3(256), 0, st, in, out, n, sink);
}

using Launch = void (*)(int, const dbl2*, dbl2*, long, dbl2*, hipStream_t);
struct Mix {
    int nr, nw;
    Launch f;
};
const Mix kMixes[] = {{1, 0, launch<1, 0>}, {7, 1, launch<7, 1>}, {3, 1, launch<3, 1>}, {2, 1, launch<2, 1>},
                      {1, 1, launch<1, 1>}, {1, 2, launch<1, 2>}, {0, 1, launch<0, 1>}};

}  // namespace

extern "C" {

int e3d_stream_bandwidth(int device, long bytes, int reads, int writes, int reps, double* gbs) {
    const Mix* mix = nullptr;
    for (const Mix& m : kMixes)
        if (m.nr == reads && m.nw == writes) mix = &m;
    if (!gbs || reps < 1 || bytes < (1L << 20) || !mix)
        return e3d::set_error(nullptr, E3D_EINVAL,
                              "e3d_stream_bandwidth: reads:writes must be 1:0, 7:1, 3:1, 2:1, 1:1, 1:2 or 0:1, "
                              "bytes >= 1 MiB, reps >= 1");
    const long n = bytes / 16 / (reads + writes);  // 16-byte elements per plane
    hipError_t e = hipSetDevice(device);
    hipStream_t st = nullptr;
    dbl2 *in = nullptr, *out = nullptr, *sink = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    double best = 0.0;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&in, 16 * (size_t)n * std::max(reads, 1));
    if (e == hipSuccess) e = hipMalloc(&out, 16 * (size_t)n * std::max(writes, 1));
    if (e == hipSuccess) e = hipMalloc(&sink, 16 * 256);
    if (e == hipSuccess) e = hipMemsetAsync(in, 0, 16 * (size_t)n * std::max(reads, 1), st);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    for (int u : {1, 4}) {  // elements per thread; the faster counts
        if (e != hipSuccess) break;
        for (int w = 0; w < 3; ++w) mix->f(u, in, out, n, sink, st);
        e = hipEventRecord(a, st);
        for (int r = 0; r < reps && e == hipSuccess; ++r) mix->f(u, in, out, n, sink, st);
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(b, st);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        if (e == hipSuccess) best = std::max(best, 16.0 * n * (reads + writes) * reps / (ms * 1e-3) / 1e9);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    if (sink) (void)hipFree(sink);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess) return e3d::set_error(nullptr, E3D_EHIP, std::string("e3d_stream_bandwidth: ") + hipGetErrorString(e));
    *gbs = best;
    return E3D_OK;
}

}  // extern "C"
