// Coordinate MLP of the learned warm start: training and inference on the
// GPU with BF16 operands and FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x32_bf16), all kernels hand-written here.
//
// Network: n_in features (padded to 32) -> depth hidden layers of `width`
// units with SiLU -> n_out outputs (padded to 32), linear. Master weights,
// biases and Adam moments are FP32; the GEMM operands (weights, activations,
// gradients of activations) are BF16, rounded to nearest even; every GEMM
// accumulates in FP32 and every reduction runs in a fixed order, so a run is
// deterministic.
//
// All GEMMs are "NT": C[M x N] = A[M x K] * B[N x K]^T with A and B row-major
// and K contiguous, which is what one MFMA lane reads as 8 consecutive bf16.
// Forward uses W as [out x in] (Wt); the input gradient uses W as [in x out]
// (Wm); weight gradients reduce over the samples using transposed activation
// copies [features x samples] written by the producing epilogue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/euler3d.h"
#include "abi_internal.hpp"

namespace {

typedef uint16_t bf16_t;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int kPad = 32;     // feature / output padding (MFMA reduction step)
constexpr int kRows = 128;   // GEMM rows per workgroup: 4 waves x 32
constexpr int kThreads = 256;

__host__ __device__ inline bf16_t to_bf16(float f) {  // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(bf16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
__device__ inline float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

// EPI_FWD_INFER: the SiLU layer of EPI_FWD_SILU with the same arithmetic,
// storing only the bf16 activation (inference keeps no pre-activations and
// no transposed copy).
enum Epi { EPI_FWD_SILU = 0, EPI_FWD_OUT = 1, EPI_BWD_SILU = 2, EPI_STORE = 3, EPI_FWD_INFER = 4 };

struct EpiArgs {
    const float* bias;  // [N] (forward)
    float* Z;           // [M x N] pre-activations (forward: written; backward: read)
    bf16_t* Y;          // [M x N] bf16 result (forward activation, or activation gradient)
    bf16_t* Yt;         // [N x M] its transpose
    float* F;           // [M x N] fp32 result (output layer; weight-gradient partials, one [M x N] per split)
};

// Epilogue of one 16 x 16 MFMA tile: the lane holds C[mb + q][n], q = 0..3.
template <int EPI>
__device__ __forceinline__ void epilogue(const EpiArgs& e, int M, int N, int mb, int n, const floatx4& acc) {
    bf16_t t[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int m = mb + q;
        const size_t o = (size_t)m * N + n;
        float v = acc[q];
        if (EPI == EPI_FWD_SILU) {
            v += e.bias[n];
            e.Z[o] = v;
            const float y = v * sigmoid(v);
            t[q] = to_bf16(y);
            e.Y[o] = t[q];
        } else if (EPI == EPI_FWD_INFER) {
            v += e.bias[n];
            const float y = v * sigmoid(v);
            e.Y[o] = to_bf16(y);
        } else if (EPI == EPI_FWD_OUT) {
            e.F[o] = v + e.bias[n];
        } else if (EPI == EPI_BWD_SILU) {
            const float z = e.Z[o], s = sigmoid(z);
            const float d = v * (s + z * s * (1.0f - s));
            t[q] = to_bf16(d);
            e.Y[o] = t[q];
        } else {
            e.F[(size_t)blockIdx.z * M * N + o] = v;
        }
    }
    if (EPI == EPI_FWD_SILU || EPI == EPI_BWD_SILU) {
        const uint64_t packed = (uint64_t)t[0] | ((uint64_t)t[1] << 16) | ((uint64_t)t[2] << 32) | ((uint64_t)t[3] << 48);
        *reinterpret_cast<uint64_t*>(e.Yt + (size_t)n * M + mb) = packed;
    }
}

// C[M x N] = A[M x K] * B[N x K]^T over K range [k0, k1) of split blockIdx.z.
// Workgroup: 4 waves, rows [128 x, +128), columns [16*TN y, +16*TN); wave w
// computes rows +32w .. +32w+31 as two 16-row MFMA tiles.
template <int EPI, int TN>
__global__ void __launch_bounds__(kThreads) k_gemm(int M, int N, int ksplit, const bf16_t* __restrict__ A, int lda,
                                                   const bf16_t* __restrict__ B, int ldb, EpiArgs e) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = blockIdx.x * kRows + wave * 32;
    const int n0 = blockIdx.y * 16 * TN;
    const int k0 = blockIdx.z * ksplit, k1 = k0 + ksplit;
    const int r = lane & 15, g = lane >> 4;
    floatx4 acc[2][TN];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k = k0; k < k1; k += 32) {
        bf16x8 a[2], b[TN];
#pragma unroll
        for (int i = 0; i < 2; ++i)
            a[i] = *reinterpret_cast<const bf16x8*>(A + (size_t)(m0 + 16 * i + r) * lda + k + 8 * g);
#pragma unroll
        for (int j = 0; j < TN; ++j)
            b[j] = *reinterpret_cast<const bf16x8*>(B + (size_t)(n0 + 16 * j + r) * ldb + k + 8 * g);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    // lane holds C[m0 + 16 i + 4 g + q][n0 + 16 j + r], q = 0..3
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) epilogue<EPI>(e, M, N, m0 + 16 * i + 4 * g, n0 + 16 * j + r, acc[i][j]);
}

// The same product with the workgroup's operand tiles staged in LDS: a
// TM x T output tile per workgroup (128 x 128 or 64 x 64), a quarter of it
// per wave (4 x 4 or 2 x 2 MFMA tiles), so each 16-byte operand fetched from
// memory feeds up to 16 MFMAs of the workgroup instead of 2-4. A k-step covers KS = 32 or 64 columns of k; the
// next k-step's tiles are loaded into registers while this step's MFMAs run
// and then stored into the second of two LDS buffers, so a k-step needs one
// barrier. Every output element accumulates over k in 32-wide MFMAs in
// ascending order, as in k_gemm, so the two give the same bits for any KS.
// M and N must be multiples of TM and T, the k range of a split a multiple of KS.
template <int EPI, int KS, int TM, int T>
__global__ void __launch_bounds__(kThreads) k_gemm_lds(int M, int N, int ksplit, const bf16_t* __restrict__ A,
                                                       int lda, const bf16_t* __restrict__ B, int ldb, EpiArgs e) {
    constexpr int LD = KS + 8;            // LDS row stride in bf16 (a bank-conflict-free ds_read_b128 pattern)
    constexpr int IM = TM / 32;           // 16-row MFMA tiles per wave (the waves split the TM rows in two)
    constexpr int JN = T / 32;            // 16-column MFMA tiles per wave (and the T columns in two)
    constexpr int CPR = KS / 8;           // 16-byte chunks per tile row and k-step
    constexpr int NCA = TM * CPR / kThreads, NCB = T * CPR / kThreads;  // chunks per thread: A, B
    __shared__ __align__(16) bf16_t smem[2 * (TM + T) * LD];  // A buffers 0, 1, then B buffers 0, 1
    bf16_t* const Bs = smem + 2 * TM * LD;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int mt = blockIdx.x * TM, nt = blockIdx.y * T;
    const int k0 = blockIdx.z * ksplit, k1 = k0 + ksplit;
    const int r = lane & 15, g = lane >> 4;
    // this thread's chunks of each operand tile: chunk c is row (tid + 256 c) / CPR,
    // columns 8 ((tid + 256 c) % CPR) .. + 7 of the k-step
    const int part = (tid % CPR) * 8;
    int lofa[NCA], lofb[NCB];
    const bf16_t* ga_p[NCA];
    const bf16_t* gb_p[NCB];
#pragma unroll
    for (int c = 0; c < NCA; ++c) {
        const int row = (tid + kThreads * c) / CPR;
        lofa[c] = row * LD + part;
        ga_p[c] = A + (size_t)(mt + row) * lda + part;
    }
#pragma unroll
    for (int c = 0; c < NCB; ++c) {
        const int row = (tid + kThreads * c) / CPR;
        lofb[c] = row * LD + part;
        gb_p[c] = B + (size_t)(nt + row) * ldb + part;
    }
    floatx4 acc[IM][JN];
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    bf16x8 ga[NCA], gb[NCB];
#pragma unroll
    for (int c = 0; c < NCA; ++c) ga[c] = *reinterpret_cast<const bf16x8*>(ga_p[c] + k0);
#pragma unroll
    for (int c = 0; c < NCB; ++c) gb[c] = *reinterpret_cast<const bf16x8*>(gb_p[c] + k0);
#pragma unroll
    for (int c = 0; c < NCA; ++c) *reinterpret_cast<bf16x8*>(smem + lofa[c]) = ga[c];
#pragma unroll
    for (int c = 0; c < NCB; ++c) *reinterpret_cast<bf16x8*>(Bs + lofb[c]) = gb[c];
    __syncthreads();
    int cur = 0;
    for (int k = k0; k < k1; k += KS) {
        const bool more = k + KS < k1;  // uniform
        if (more) {  // the next k-step's tiles, in flight while this step's MFMAs run
#pragma unroll
            for (int c = 0; c < NCA; ++c) ga[c] = *reinterpret_cast<const bf16x8*>(ga_p[c] + k + KS);
#pragma unroll
            for (int c = 0; c < NCB; ++c) gb[c] = *reinterpret_cast<const bf16x8*>(gb_p[c] + k + KS);
        }
        const bf16_t* Ac = smem + cur * TM * LD;
        const bf16_t* Bc = Bs + cur * T * LD;
#pragma unroll
        for (int kk = 0; kk < KS; kk += 32) {
            bf16x8 a[IM], b[JN];
#pragma unroll
            for (int i = 0; i < IM; ++i)
                a[i] = *reinterpret_cast<const bf16x8*>(Ac + (wm * (TM / 2) + 16 * i + r) * LD + kk + 8 * g);
#pragma unroll
            for (int j = 0; j < JN; ++j)
                b[j] = *reinterpret_cast<const bf16x8*>(Bc + (wn * (T / 2) + 16 * j + r) * LD + kk + 8 * g);
#pragma unroll
            for (int i = 0; i < IM; ++i)
#pragma unroll
                for (int j = 0; j < JN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {  // into the other buffer, which no wave reads during this step
            bf16_t* An = smem + (cur ^ 1) * TM * LD;
            bf16_t* Bn = Bs + (cur ^ 1) * T * LD;
#pragma unroll
            for (int c = 0; c < NCA; ++c) *reinterpret_cast<bf16x8*>(An + lofa[c]) = ga[c];
#pragma unroll
            for (int c = 0; c < NCB; ++c) *reinterpret_cast<bf16x8*>(Bn + lofb[c]) = gb[c];
        }
        // this step's reads of buffer cur and writes of the other are done before
        // the next step reads the other and (one step later) overwrites cur
        __syncthreads();
        cur ^= 1;
    }
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j)
            epilogue<EPI>(e, M, N, mt + wm * (TM / 2) + 16 * i + 4 * g, nt + wn * (T / 2) + 16 * j + r, acc[i][j]);
}

// Mini-batch assembly: features (padded with zeros to 32) as bf16, row-major and
// transposed, and the fp32 targets, for samples idx[0..B).
// A workgroup takes 64 samples; each of its four waves a quarter of the 32
// padded features of all 64 (so a wave's transposed stores of one feature are
// 64 consecutive samples).
constexpr int kGatherSamples = kThreads / 4;
__global__ void k_gather(const float* X, const float* T, const int* idx, int n_in, int n_out, int B, bf16_t* X0,
                         bf16_t* X0t, float* Tb) {
    const int b = blockIdx.x * kGatherSamples + (threadIdx.x & (kGatherSamples - 1));
    const int quarter = threadIdx.x / kGatherSamples;
    if (b >= B) return;
    const size_t i = (size_t)idx[b];
#pragma unroll
    for (int u = 0; u < kPad / 4; ++u) {
        const int k = quarter * (kPad / 4) + u;
        const bf16_t v = to_bf16(k < n_in ? X[i * n_in + k] : 0.0f);
        X0[(size_t)b * kPad + k] = v;
        X0t[(size_t)k * B + b] = v;
    }
    for (int j = quarter; j < n_out; j += 4) Tb[(size_t)b * n_out + j] = T[i * n_out + j];
}

// Mean squared error over B samples and n_out outputs: per-workgroup partial sums,
// and the output gradient 2 (y - t) / (B n_out) as bf16 (row-major and transposed).
__global__ void k_loss(const float* Y, const float* Tb, int n_out, int B, bf16_t* dY, bf16_t* dYt, float* part) {
    __shared__ float red[kThreads];
    const int b = blockIdx.x * kThreads + threadIdx.x;
    float s = 0.0f;
    if (b < B) {
        const float scale = 2.0f / ((float)B * (float)n_out);
        for (int j = 0; j < kPad; ++j) {
            float d = 0.0f;
            if (j < n_out) {
                const float diff = Y[(size_t)b * kPad + j] - Tb[(size_t)b * n_out + j];
                s += diff * diff;
                d = diff * scale;
            }
            const bf16_t h = to_bf16(d);
            dY[(size_t)b * kPad + j] = h;
            dYt[(size_t)j * B + b] = h;
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Bias gradient: row sums of the transposed activation gradient [N x B], one
// workgroup per row, fixed tree.
__global__ void k_row_sum(const bf16_t* Gt, int B, float* out) {
    __shared__ float red[kThreads];
    const int n = blockIdx.x;
    float s = 0.0f;
    // eight loads in flight, then added in the same order as one at a time
    const bf16_t* row = Gt + (size_t)n * B;
    int b = threadIdx.x;
    for (; b + 7 * kThreads < B; b += 8 * kThreads) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = from_bf16(row[b + u * kThreads]);
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; b < B; b += kThreads) s += from_bf16(row[b]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = red[0];
}

// Sums the split partials of a weight gradient into grad[out x in] (Wt layout);
// `transposed`: the partials are [in x out].
__global__ void k_sum_splits(const float* part, int splits, int n_out, int n_in, int transposed, float* grad) {
    const int x = blockIdx.x * kThreads + threadIdx.x;
    if (x >= n_out * n_in) return;
    const int o = x / n_in, i = x % n_in;
    const size_t src = transposed ? (size_t)i * n_out + o : (size_t)x;
    float s = 0.0f;
    const size_t stride = (size_t)n_out * n_in;
    int z = 0;
    for (; z + 8 <= splits; z += 8) {  // eight loads in flight, added in order
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[(size_t)(z + u) * stride + src];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; z < splits; ++z) s += part[(size_t)z * stride + src];
    grad[x] = s;
}

// Adam step on one layer's weights [out x in] and biases, then the bf16 copies:
// Wt [out x in] and Wm [in x out]. Rows / columns beyond the real sizes stay 0.
__global__ void k_adam(float* W, float* mW, float* vW, const float* gW, float* bvec, float* mb, float* vb,
                       const float* gb, int n_out, int n_in, int real_out, int real_in, float lr, float b1, float b2,
                       float eps, float c1, float c2, bf16_t* Wt, bf16_t* Wm) {
    const int x = blockIdx.x * kThreads + threadIdx.x;
    if (x < n_out * n_in) {
        const int o = x / n_in, i = x % n_in;
        if (o < real_out && i < real_in) {
            const float g = gW[x];
            mW[x] = b1 * mW[x] + (1.0f - b1) * g;
            vW[x] = b2 * vW[x] + (1.0f - b2) * g * g;
            W[x] -= lr * (mW[x] / c1) / (sqrtf(vW[x] / c2) + eps);
        }
        const bf16_t h = to_bf16(W[x]);
        Wt[x] = h;
        Wm[(size_t)i * n_out + o] = h;
    }
    if (x < real_out) {
        const float g = gb[x];
        mb[x] = b1 * mb[x] + (1.0f - b1) * g;
        vb[x] = b2 * vb[x] + (1.0f - b2) * g * g;
        bvec[x] -= lr * (mb[x] / c1) / (sqrtf(vb[x] / c2) + eps);
    }
}

// bf16 copies of the master weights, without an update.
__global__ void k_copy_weights(const float* W, int n_out, int n_in, bf16_t* Wt, bf16_t* Wm) {
    const int x = blockIdx.x * kThreads + threadIdx.x;
    if (x >= n_out * n_in) return;
    const int o = x / n_in, i = x % n_in;
    const bf16_t h = to_bf16(W[x]);
    Wt[x] = h;
    Wm[(size_t)i * n_out + o] = h;
}

// Inference input: rows [0, cnt) of x [cnt x n_in] as bf16 into X0 [R x 32],
// zero features beyond n_in and zero rows beyond cnt (every entry written).
__global__ void k_pack_rows(const float* x, int cnt, int R, int n_in, bf16_t* X0) {
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= R * kPad) return;
    const int b = t / kPad, k = t % kPad;
    X0[t] = to_bf16(b < cnt && k < n_in ? x[(size_t)b * n_in + k] : 0.0f);
}

// Inference output: the first n_out columns of out [cnt x 32] into y [cnt x n_out].
__global__ void k_unpack_rows(const float* out, int cnt, int n_out, float* y) {
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= cnt * n_out) return;
    const int b = t / n_out, j = t % n_out;
    y[t] = out[(size_t)b * kPad + j];
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace


struct e3d_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_in = 0, width = 0, depth = 0, n_out = 0, nl = 0;  // nl = depth + 1 linear layers
    std::vector<int> din, dout, rin, rout;                  // padded / real fan-in and fan-out per layer
    std::vector<float*> W, b, mW, vW, mb, vb, gW, gb;
    std::vector<bf16_t*> Wt, Wm;
    long step = 0;
    // data
    float *X = nullptr, *T = nullptr;
    long n_data = 0;
    // work buffers for batch size cap
    int cap = 0;
    int* idx = nullptr;
    bf16_t *X0 = nullptr, *X0t = nullptr, *dO = nullptr, *dOt = nullptr;
    float *Tb = nullptr, *out = nullptr, *part_loss = nullptr, *part_w = nullptr;
    std::vector<float*> Z;
    std::vector<bf16_t*> Y, Yt, dZ, dZt;
    std::vector<void*> allocs, work;
    std::string err;
};

namespace {

int fail(e3d_mlp* m, int code, const std::string& msg) {
    if (m) m->err = msg;
    return e3d::set_error(nullptr, code, msg);
}

#define MLP_TRY(m, call)                                                                           \
    do {                                                                                           \
        hipError_t _e = (call);                                                                    \
        if (_e != hipSuccess) return fail(m, E3D_EHIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

template <class T>
int alloc(e3d_mlp* m, T** p, size_t n, std::vector<void*>& list) {
    void* q = nullptr;
    MLP_TRY(m, hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    MLP_TRY(m, hipMemsetAsync(q, 0, std::max<size_t>(n, 1) * sizeof(T), m->stream));
    list.push_back(q);
    *p = (T*)q;
    return E3D_OK;
}

constexpr int kSplits = 16;  // weight-gradient reductions: sample range split across workgroups

// Work buffers for mini-batches of up to `cap` samples (a multiple of 128).
int ensure_work(e3d_mlp* m, int cap) {
    if (cap <= m->cap) return E3D_OK;
    for (void* p : m->work) (void)hipFree(p);
    m->work.clear();
    m->Z.assign(m->depth, nullptr);
    m->Y.assign(m->depth, nullptr);
    m->Yt.assign(m->depth, nullptr);
    m->dZ.assign(m->depth, nullptr);
    m->dZt.assign(m->depth, nullptr);
    const int w = m->width;
    int rc;
    if ((rc = alloc(m, &m->idx, cap, m->work)) || (rc = alloc(m, &m->X0, (size_t)cap * kPad, m->work)) ||
        (rc = alloc(m, &m->X0t, (size_t)cap * kPad, m->work)) || (rc = alloc(m, &m->dO, (size_t)cap * kPad, m->work)) ||
        (rc = alloc(m, &m->dOt, (size_t)cap * kPad, m->work)) ||
        (rc = alloc(m, &m->Tb, (size_t)cap * m->n_out, m->work)) ||
        (rc = alloc(m, &m->out, (size_t)cap * kPad, m->work)) || (rc = alloc(m, &m->part_loss, cdiv(cap, kThreads), m->work)) ||
        (rc = alloc(m, &m->part_w, (size_t)kSplits * w * std::max(w, kPad), m->work)))
        return rc;
    for (int l = 0; l < m->depth; ++l)
        if ((rc = alloc(m, &m->Z[l], (size_t)cap * w, m->work)) || (rc = alloc(m, &m->Y[l], (size_t)cap * w, m->work)) ||
            (rc = alloc(m, &m->Yt[l], (size_t)cap * w, m->work)) || (rc = alloc(m, &m->dZ[l], (size_t)cap * w, m->work)) ||
            (rc = alloc(m, &m->dZt[l], (size_t)cap * w, m->work)))
            return rc;
    m->cap = cap;
    return E3D_OK;
}

template <int EPI, int TN>
void gemm(e3d_mlp* m, int M, int N, int K, int splits, const bf16_t* A, int lda, const bf16_t* B, int ldb,
          const EpiArgs& e) {
    hipLaunchKernelGGL((k_gemm<EPI, TN>), dim3(M / kRows, N / (16 * TN), splits), dim3(kThreads), 0, m->stream, M, N,
                       K / splits, A, lda, B, ldb, e);
}

// Wider column tiles (4 MFMA tiles per wave) where N allows: each A operand
// loaded feeds twice as many MFMAs. The accumulation order of every output
// element is the same, so results are unchanged. Tile shapes, measured on the
// 5 x 512 network at 8192 samples (ms per training step, same bits
// throughout): products with an epilogue (forward SiLU, input gradient) in
// 64 x 64 tiles with a 64-wide k-step, so eight workgroups share a CU and one
// workgroup's epilogue overlaps another's main loop (128 x 128 tiles with a
// 32-wide k-step 0.357, 64 x 128 0.327, 128 x 64 0.320, 64 x 64 0.312, with a
// 64-wide k-step 0.300); the weight-gradient products in 128 x 128 tiles with
// a 64-wide k-step (64 x 64 there: 0.310-0.320).
template <int EPI>
void gemm_n(e3d_mlp* m, int M, int N, int K, int splits, const bf16_t* A, int lda, const bf16_t* B, int ldb,
            const EpiArgs& e) {
    const int ks = K / splits;
    if (EPI == EPI_STORE && M % 128 == 0 && N % 128 == 0 && ks % 64 == 0)
        hipLaunchKernelGGL((k_gemm_lds<EPI, 64, 128, 128>), dim3(M / 128, N / 128, splits), dim3(kThreads), 0, m->stream,
                           M, N, ks, A, lda, B, ldb, e);
    else if (EPI != EPI_STORE && M % 64 == 0 && N % 64 == 0 && ks % 64 == 0)
        hipLaunchKernelGGL((k_gemm_lds<EPI, 64, 64, 64>), dim3(M / 64, N / 64, splits), dim3(kThreads), 0, m->stream, M, N,
                           ks, A, lda, B, ldb, e);
    else if (EPI != EPI_STORE && M % 64 == 0 && N % 64 == 0)
        hipLaunchKernelGGL((k_gemm_lds<EPI, 32, 64, 64>), dim3(M / 64, N / 64, splits), dim3(kThreads), 0, m->stream, M, N,
                           ks, A, lda, B, ldb, e);
    else if (N % 64 == 0) gemm<EPI, 4>(m, M, N, K, splits, A, lda, B, ldb, e);
    else gemm<EPI, 2>(m, M, N, K, splits, A, lda, B, ldb, e);
}

// Forward pass of B samples already in X0 (bf16 [B x 32]); output fp32 in m->out [B x 32].
void forward(e3d_mlp* m, int B) {
    const int w = m->width;
    for (int l = 0; l < m->depth; ++l) {
        const bf16_t* in = l == 0 ? m->X0 : m->Y[l - 1];
        const int K = l == 0 ? kPad : w;
        EpiArgs e{m->b[l], m->Z[l], m->Y[l], m->Yt[l], nullptr};
        gemm_n<EPI_FWD_SILU>(m, B, w, K, 1, in, K, m->Wt[l], K, e);
    }
    EpiArgs e{m->b[m->depth], nullptr, nullptr, nullptr, m->out};
    gemm_n<EPI_FWD_OUT>(m, B, kPad, w, 1, m->Y[m->depth - 1], w, m->Wt[m->depth], w, e);
}

// Gradients of the mean squared error over B samples (targets in m->Tb) into gW / gb.
// part_loss: where the loss kernel leaves its per-workgroup partial sums
// (m->part_loss by default).
void backward(e3d_mlp* m, int B, float* part_loss = nullptr) {
    const int w = m->width, L = m->depth;
    hipLaunchKernelGGL(k_loss, dim3(cdiv(B, kThreads)), dim3(kThreads), 0, m->stream, m->out, m->Tb, m->n_out, B, m->dO,
                       m->dOt, part_loss ? part_loss : m->part_loss);
    // output layer: dW [w x 32] (in x out) = Y_{L-1}^T dO; db = column sums of dO
    {
        EpiArgs e{nullptr, nullptr, nullptr, nullptr, m->part_w};
        gemm_n<EPI_STORE>(m, w, kPad, B, kSplits, m->Yt[L - 1], B, m->dOt, B, e);
        hipLaunchKernelGGL(k_sum_splits, dim3(cdiv((long)kPad * w, kThreads)), dim3(kThreads), 0, m->stream, m->part_w,
                           kSplits, kPad, w, 1, m->gW[L]);
        hipLaunchKernelGGL(k_row_sum, dim3(kPad), dim3(kThreads), 0, m->stream, m->dOt, B, m->gb[L]);
        EpiArgs d{nullptr, m->Z[L - 1], m->dZ[L - 1], m->dZt[L - 1], nullptr};
        gemm_n<EPI_BWD_SILU>(m, B, w, kPad, 1, m->dO, kPad, m->Wm[L], kPad, d);
    }
    for (int l = L - 1; l >= 0; --l) {
        const int K = l == 0 ? kPad : w;
        const bf16_t* in_t = l == 0 ? m->X0t : m->Yt[l - 1];
        // dW [w x K] (out x in) = dZ_l^T X_{l-1}
        EpiArgs e{nullptr, nullptr, nullptr, nullptr, m->part_w};
        gemm_n<EPI_STORE>(m, w, K, B, kSplits, m->dZt[l], B, in_t, B, e);
        hipLaunchKernelGGL(k_sum_splits, dim3(cdiv((long)w * K, kThreads)), dim3(kThreads), 0, m->stream, m->part_w,
                           kSplits, w, K, 0, m->gW[l]);
        hipLaunchKernelGGL(k_row_sum, dim3(w), dim3(kThreads), 0, m->stream, m->dZt[l], B, m->gb[l]);
        if (l > 0) {
            EpiArgs d{nullptr, m->Z[l - 1], m->dZ[l - 1], m->dZt[l - 1], nullptr};
            gemm_n<EPI_BWD_SILU>(m, B, w, w, 1, m->dZ[l], w, m->Wm[l], w, d);
        }
    }
}

int check_batch(e3d_mlp* m, long B) {
    if (B <= 0 || B % kRows != 0 || B % (32 * kSplits) != 0 || B > (1L << 24))
        return fail(m, E3D_EINVAL, "mini-batch size must be a positive multiple of 512 (and at most 2^24)");
    return E3D_OK;
}

}  // namespace

// Scratch of e3d_mlp_predict_noalloc: the bf16 input rows, two bf16
// activation buffers and the fp32 output rows of one chunk. One copy per GPU
// for the life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (scratch_address below).
constexpr size_t kScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current: looked up
// once per device, then reused.
int scratch_address(e3d_mlp* m, int device, void** p) {
    if (device < 0 || device >= kMaxDevices) return fail(m, E3D_EINVAL, "e3d_mlp_predict_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* q = nullptr;
        const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q)
            return fail(m, E3D_EHIP, std::string("e3d_mlp_predict_noalloc: hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
        g_scratch_ptr[device] = q;
    }
    *p = g_scratch_ptr[device];
    return E3D_OK;
}

// Rows per chunk of e3d_mlp_predict_noalloc for a network of this width: the
// largest multiple of 512 whose buffers fit the scratch (0 if none does).
long noalloc_rows(const e3d_mlp* m) {
    const size_t per_row = sizeof(bf16_t) * kPad + 2 * sizeof(bf16_t) * (size_t)m->width + sizeof(float) * kPad;
    return (long)(kScratchBytes / per_row) / 512 * 512;
}

// Accepts only memory of the network's device (a host pointer in a kernel
// argument would fault the GPU).
int check_device_ptr(e3d_mlp* m, const void* p, const char* what) {
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(m, E3D_EINVAL, std::string("e3d_mlp_predict_noalloc: ") + what + " is not device memory");
    }
    if ((a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged) || a.device != m->device)
        return fail(m, E3D_EINVAL, std::string("e3d_mlp_predict_noalloc: ") + what +
                                       " must be device memory of the network's GPU " + std::to_string(m->device));
    return E3D_OK;
}

// Inference forward pass of B rows of X0 with the activations alternating
// between Ya and Yb; fp32 outputs [B x 32] into out. The same GEMMs and
// arithmetic as forward(), so the outputs have the same bits.
void forward_infer(e3d_mlp* m, int B, const bf16_t* X0, bf16_t* Ya, bf16_t* Yb, float* out) {
    const int w = m->width;
    const bf16_t* in = X0;
    bf16_t *cur = Ya, *nxt = Yb;
    for (int l = 0; l < m->depth; ++l) {
        const int K = l == 0 ? kPad : w;
        EpiArgs e{m->b[l], nullptr, cur, nullptr, nullptr};
        gemm_n<EPI_FWD_INFER>(m, B, w, K, 1, in, K, m->Wt[l], K, e);
        in = cur;
        std::swap(cur, nxt);
    }
    EpiArgs e{m->b[m->depth], nullptr, nullptr, nullptr, out};
    gemm_n<EPI_FWD_OUT>(m, B, kPad, w, 1, in, w, m->Wt[m->depth], w, e);
}

}  // namespace

extern "C" {

int e3d_mlp_create(int device, int n_in, int width, int depth, int n_out, unsigned long long seed, e3d_mlp** out) {
    if (!out || n_in < 1 || n_in > kPad || n_out < 1 || n_out > kPad || depth < 1 || width < 32 || width % kRows != 0)
        return e3d::set_error(nullptr, E3D_EINVAL,
                              "e3d_mlp_create: need 1 <= n_in, n_out <= 32, depth >= 1, width a multiple of 128");
    e3d_mlp* m = new e3d_mlp();
    m->device = device;
    m->n_in = n_in, m->width = width, m->depth = depth, m->n_out = n_out, m->nl = depth + 1;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete m;
        return e3d::set_error(nullptr, E3D_EHIP, std::string("e3d_mlp_create: ") + hipGetErrorString(e));
    }
    std::mt19937_64 rng(seed);
    for (int l = 0; l < m->nl; ++l) {
        const int pin = l == 0 ? kPad : width, pout = l == m->depth ? kPad : width;
        const int rin = l == 0 ? n_in : width, rout = l == m->depth ? n_out : width;
        m->din.push_back(pin), m->dout.push_back(pout), m->rin.push_back(rin), m->rout.push_back(rout);
        std::vector<float> w((size_t)pout * pin, 0.0f);
        // He-uniform initialisation on the real part; the last layer starts at zero,
        // so an untrained network predicts zero (the freestream start)
        const float a = l == m->depth ? 0.0f : std::sqrt(6.0f / (float)rin);
        std::uniform_real_distribution<float> U(-a, a);
        for (int o = 0; o < rout; ++o)
            for (int i = 0; i < rin; ++i) w[(size_t)o * pin + i] = U(rng);
        float *W, *b, *mW, *vW, *mb, *vb, *gW, *gb;
        bf16_t *Wt, *Wm;
        int rc;
        if ((rc = alloc(m, &W, w.size(), m->allocs)) || (rc = alloc(m, &b, pout, m->allocs)) ||
            (rc = alloc(m, &mW, w.size(), m->allocs)) || (rc = alloc(m, &vW, w.size(), m->allocs)) ||
            (rc = alloc(m, &mb, pout, m->allocs)) || (rc = alloc(m, &vb, pout, m->allocs)) ||
            (rc = alloc(m, &gW, w.size(), m->allocs)) || (rc = alloc(m, &gb, pout, m->allocs)) ||
            (rc = alloc(m, &Wt, w.size(), m->allocs)) || (rc = alloc(m, &Wm, w.size(), m->allocs))) {
            e3d_mlp_destroy(m);
            return rc;
        }
        // on the network's stream, after the zero fill alloc() queued there (the
        // stream does not synchronise with the null stream), and complete before
        // the host vector goes away
        e = hipMemcpyAsync(W, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        if (e != hipSuccess) {
            e3d_mlp_destroy(m);
            return e3d::set_error(nullptr, E3D_EHIP, std::string("e3d_mlp_create: ") + hipGetErrorString(e));
        }
        hipLaunchKernelGGL(k_copy_weights, dim3(cdiv((long)w.size(), kThreads)), dim3(kThreads), 0, m->stream, W, pout, pin,
                           Wt, Wm);
        m->W.push_back(W), m->b.push_back(b), m->mW.push_back(mW), m->vW.push_back(vW), m->mb.push_back(mb);
        m->vb.push_back(vb), m->gW.push_back(gW), m->gb.push_back(gb), m->Wt.push_back(Wt), m->Wm.push_back(Wm);
    }
    e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        e3d_mlp_destroy(m);
        return e3d::set_error(nullptr, E3D_EHIP, std::string("e3d_mlp_create: ") + hipGetErrorString(e));
    }
    *out = m;
    return E3D_OK;
}

void e3d_mlp_destroy(e3d_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* p : m->work) (void)hipFree(p);
    for (void* p : m->allocs) (void)hipFree(p);
    if (m->X) (void)hipFree(m->X);
    if (m->T) (void)hipFree(m->T);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

long e3d_mlp_param_count(const e3d_mlp* m) {
    if (!m) return E3D_EINVAL;
    long n = 0;
    for (int l = 0; l < m->nl; ++l) n += (long)m->rout[l] * m->rin[l] + m->rout[l];
    return n;
}

// Parameters in a flat fp32 array: per layer, W [out x in] row-major then b [out] (real sizes).
int e3d_mlp_get_params(e3d_mlp* m, float* out) {
    if (!m || !out) return e3d::set_error(nullptr, E3D_EINVAL, "e3d_mlp_get_params: null argument");
    MLP_TRY(m, hipSetDevice(m->device));
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    long o = 0;
    for (int l = 0; l < m->nl; ++l) {
        std::vector<float> w((size_t)m->dout[l] * m->din[l]), b(m->dout[l]);
        MLP_TRY(m, hipMemcpy(w.data(), m->W[l], sizeof(float) * w.size(), hipMemcpyDeviceToHost));
        MLP_TRY(m, hipMemcpy(b.data(), m->b[l], sizeof(float) * b.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < m->rout[l]; ++r)
            for (int i = 0; i < m->rin[l]; ++i) out[o++] = w[(size_t)r * m->din[l] + i];
        for (int r = 0; r < m->rout[l]; ++r) out[o++] = b[r];
    }
    return E3D_OK;
}

int e3d_mlp_set_params(e3d_mlp* m, const float* in) {
    if (!m || !in) return e3d::set_error(nullptr, E3D_EINVAL, "e3d_mlp_set_params: null argument");
    MLP_TRY(m, hipSetDevice(m->device));
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    long o = 0;
    for (int l = 0; l < m->nl; ++l) {
        std::vector<float> w((size_t)m->dout[l] * m->din[l], 0.0f), b(m->dout[l], 0.0f);
        for (int r = 0; r < m->rout[l]; ++r)
            for (int i = 0; i < m->rin[l]; ++i) w[(size_t)r * m->din[l] + i] = in[o++];
        for (int r = 0; r < m->rout[l]; ++r) b[r] = in[o++];
        MLP_TRY(m, hipMemcpy(m->W[l], w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice));
        MLP_TRY(m, hipMemcpy(m->b[l], b.data(), sizeof(float) * b.size(), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_copy_weights, dim3(cdiv((long)w.size(), kThreads)), dim3(kThreads), 0, m->stream, m->W[l],
                           m->dout[l], m->din[l], m->Wt[l], m->Wm[l]);
    }
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    return E3D_OK;
}

int e3d_mlp_set_data(e3d_mlp* m, const float* x, const float* t, long n) {
    if (!m || !x || !t || n < 1) return e3d::set_error(nullptr, E3D_EINVAL, "e3d_mlp_set_data: bad argument");
    MLP_TRY(m, hipSetDevice(m->device));
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    if (m->X) (void)hipFree(m->X);
    if (m->T) (void)hipFree(m->T);
    m->X = m->T = nullptr;
    MLP_TRY(m, hipMalloc(&m->X, sizeof(float) * n * m->n_in));
    MLP_TRY(m, hipMalloc(&m->T, sizeof(float) * n * m->n_out));
    MLP_TRY(m, hipMemcpy(m->X, x, sizeof(float) * n * m->n_in, hipMemcpyHostToDevice));
    MLP_TRY(m, hipMemcpy(m->T, t, sizeof(float) * n * m->n_out, hipMemcpyHostToDevice));
    m->n_data = n;
    return E3D_OK;
}

// Loss and gradients (flat, as e3d_mlp_get_params) of one batch of the data set,
// samples idx[0..B); no update. For checking the kernels against a reference.
int e3d_mlp_gradients(e3d_mlp* m, const int* idx, int B, float* loss, float* grads) {
    if (!m || !idx || !grads || !m->X) return e3d::set_error(nullptr, E3D_EINVAL, "e3d_mlp_gradients: bad argument");
    int rc = check_batch(m, B);
    if (rc) return rc;
    MLP_TRY(m, hipSetDevice(m->device));
    if ((rc = ensure_work(m, B))) return rc;
    for (int b = 0; b < B; ++b)
        if (idx[b] < 0 || idx[b] >= m->n_data) return fail(m, E3D_EINVAL, "e3d_mlp_gradients: sample index out of range");
    MLP_TRY(m, hipMemcpyAsync(m->idx, idx, sizeof(int) * B, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_gather, dim3(cdiv(B, kGatherSamples)), dim3(kThreads), 0, m->stream, m->X, m->T, m->idx, m->n_in,
                       m->n_out, B, m->X0, m->X0t, m->Tb);
    forward(m, B);
    backward(m, B);
    MLP_TRY(m, hipGetLastError());
    std::vector<float> part(cdiv(B, kThreads));
    MLP_TRY(m, hipMemcpyAsync(part.data(), m->part_loss, sizeof(float) * part.size(), hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(m, hipStreamSynchronize(m->stream));
    double s = 0.0;
    for (float p : part) s += p;
    if (loss) *loss = (float)(s / ((double)B * m->n_out));
    long o = 0;
    for (int l = 0; l < m->nl; ++l) {
        std::vector<float> g((size_t)m->dout[l] * m->din[l]), gb(m->dout[l]);
        MLP_TRY(m, hipMemcpy(g.data(), m->gW[l], sizeof(float) * g.size(), hipMemcpyDeviceToHost));
        MLP_TRY(m, hipMemcpy(gb.data(), m->gb[l], sizeof(float) * gb.size(), hipMemcpyDeviceToHost));
        for (int r = 0; r < m->rout[l]; ++r)
            for (int i = 0; i < m->rin[l]; ++i) grads[o++] = g[(size_t)r * m->din[l] + i];
        for (int r = 0; r < m->rout[l]; ++r) grads[o++] = gb[r];
    }
    return E3D_OK;
}

// `steps` Adam steps on mini-batches of B samples drawn without replacement from a
// shuffled order of the data set (seeded), learning rate lr (cosine-decayed to lr/20
// over the steps when cosine != 0). loss (may be NULL): one value per step.
int e3d_mlp_train(e3d_mlp* m, int steps, int B, float lr, int cosine, unsigned long long seed, float* loss) {
    if (!m || steps < 0 || !m->X) return e3d::set_error(nullptr, E3D_EINVAL, "e3d_mlp_train: bad argument or no data");
    int rc = check_batch(m, B);
    if (rc) return rc;
    MLP_TRY(m, hipSetDevice(m->device));
    if ((rc = ensure_work(m, B))) return rc;
    // Steps run in chunks: the sample indices of a whole chunk go up in one
    // copy and the loss partials of each step stay on the device until the
    // chunk ends, so the host waits once per chunk, not once per step. The
    // indices, their order and every kernel are those of one step at a time.
    constexpr int kChunk = 32;
    const int nparts = cdiv(B, kThreads);
    int* d_idx = nullptr;
    float* d_part = nullptr;
    MLP_TRY(m, hipMalloc(&d_idx, sizeof(int) * (size_t)kChunk * B));
    hipError_t ea = hipMalloc(&d_part, sizeof(float) * (size_t)kChunk * nparts);
    if (ea != hipSuccess) {
        (void)hipFree(d_idx);
        return fail(m, E3D_EHIP, std::string("e3d_mlp_train: ") + hipGetErrorString(ea));
    }
    std::mt19937_64 rng(seed);
    std::vector<int> order(m->n_data);
    for (long i = 0; i < m->n_data; ++i) order[i] = (int)i;
    long pos = m->n_data;
    std::vector<int> batch((size_t)kChunk * B);
    std::vector<float> part((size_t)kChunk * nparts);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    hipError_t e = hipSuccess;
    for (int s0 = 0; s0 < steps && e == hipSuccess; s0 += kChunk) {
        const int n = std::min(kChunk, steps - s0);
        for (size_t b = 0; b < (size_t)n * B; ++b) {
            if (pos >= m->n_data) {
                std::shuffle(order.begin(), order.end(), rng);
                pos = 0;
            }
            batch[b] = order[pos++];
        }
        e = hipMemcpyAsync(d_idx, batch.data(), sizeof(int) * (size_t)n * B, hipMemcpyHostToDevice, m->stream);
        for (int j = 0; j < n && e == hipSuccess; ++j) {
            const int s = s0 + j;
            hipLaunchKernelGGL(k_gather, dim3(cdiv(B, kGatherSamples)), dim3(kThreads), 0, m->stream, m->X, m->T,
                               d_idx + (size_t)j * B, m->n_in, m->n_out, B, m->X0, m->X0t, m->Tb);
            forward(m, B);
            backward(m, B, d_part + (size_t)j * nparts);
            ++m->step;
            const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
            const float rate =
                cosine ? lr * (0.05f + 0.95f * 0.5f * (1.0f + std::cos(3.14159265f * s / std::max(1, steps)))) : lr;
            for (int l = 0; l < m->nl; ++l)
                hipLaunchKernelGGL(k_adam, dim3(cdiv((long)m->dout[l] * m->din[l], kThreads)), dim3(kThreads), 0,
                                   m->stream, m->W[l], m->mW[l], m->vW[l], m->gW[l], m->b[l], m->mb[l], m->vb[l],
                                   m->gb[l], m->dout[l], m->din[l], m->rout[l], m->rin[l], rate, b1, b2, eps, c1, c2,
                                   m->Wt[l], m->Wm[l]);
            e = hipGetLastError();
        }
        if (e == hipSuccess && loss)
            e = hipMemcpyAsync(part.data(), d_part, sizeof(float) * (size_t)n * nparts, hipMemcpyDeviceToHost,
                               m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);  // before `batch` and `part` are reused
        if (e == hipSuccess && loss)
            for (int j = 0; j < n; ++j) {
                double t = 0.0;
                for (int q = 0; q < nparts; ++q) t += part[(size_t)j * nparts + q];
                loss[s0 + j] = (float)(t / ((double)B * m->n_out));
            }
    }
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(d_idx);
    (void)hipFree(d_part);
    if (e != hipSuccess) return fail(m, E3D_EHIP, std::string("e3d_mlp_train: ") + hipGetErrorString(e));
    return E3D_OK;
}

// Network outputs for n feature rows (host arrays, fp32), in chunks.
int e3d_mlp_predict(e3d_mlp* m, const float* x, long n, float* y) {
    if (!m || !x || !y || n < 0) return e3d::set_error(nullptr, E3D_EINVAL, "e3d_mlp_predict: bad argument");
    MLP_TRY(m, hipSetDevice(m->device));
    const int chunk = 1 << 16;
    int rc;
    if ((rc = ensure_work(m, chunk))) return rc;
    float* dx = nullptr;
    MLP_TRY(m, hipMalloc(&dx, sizeof(float) * (size_t)chunk * m->n_in));
    std::vector<int> idx(chunk);
    std::vector<float> out((size_t)chunk * kPad);
    for (long s = 0; s < n; s += chunk) {
        const int cnt = (int)std::min<long>(chunk, n - s);
        const int B = (cnt + 511) / 512 * 512;
        for (int b = 0; b < B; ++b) idx[b] = b < cnt ? b : 0;
        hipError_t e = hipMemcpyAsync(dx, x + s * m->n_in, sizeof(float) * cnt * m->n_in, hipMemcpyHostToDevice, m->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(m->idx, idx.data(), sizeof(int) * B, hipMemcpyHostToDevice, m->stream);
        if (e != hipSuccess) {
            (void)hipFree(dx);
            return fail(m, E3D_EHIP, std::string("e3d_mlp_predict: ") + hipGetErrorString(e));
        }
        hipLaunchKernelGGL(k_gather, dim3(cdiv(B, kGatherSamples)), dim3(kThreads), 0, m->stream, dx, m->T ? m->T : dx, m->idx,
                           m->n_in, 0, B, m->X0, m->X0t, m->Tb);
        forward(m, B);
        e = hipMemcpyAsync(out.data(), m->out, sizeof(float) * (size_t)B * kPad, hipMemcpyDeviceToHost, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
        if (e != hipSuccess) {
            (void)hipFree(dx);
            return fail(m, E3D_EHIP, std::string("e3d_mlp_predict: ") + hipGetErrorString(e));
        }
        for (int b = 0; b < cnt; ++b)
            for (int j = 0; j < m->n_out; ++j) y[(s + b) * m->n_out + j] = out[(size_t)b * kPad + j];
    }
    (void)hipFree(dx);
    return E3D_OK;
}

// e3d_mlp_predict on device buffers without allocating: each chunk of rows
// goes through the static scratch (packed inputs, activations, outputs).
// Every scratch entry a chunk reads is written earlier in the same chunk, so
// nothing needs clearing between calls; calls that overlap on one GPU share
// the scratch and race.
int e3d_mlp_predict_noalloc(e3d_mlp* m, const float* x, long n, float* y) {
    if (!m || !x || !y || n < 0) return e3d::set_error(nullptr, E3D_EINVAL, "e3d_mlp_predict_noalloc: bad argument");
    MLP_TRY(m, hipSetDevice(m->device));
    const long R = noalloc_rows(m);
    if (R < 512)
        return fail(m, E3D_EINVAL, "e3d_mlp_predict_noalloc: width " + std::to_string(m->width) +
                                       " does not fit the " + std::to_string(kScratchBytes >> 20) +
                                       " MiB scratch (512 rows need " +
                                       std::to_string(512 * (2 * kPad + 4 * (size_t)m->width + 4 * kPad)) + " bytes)");
    int rc;
    if ((rc = check_device_ptr(m, x, "x")) || (rc = check_device_ptr(m, y, "y"))) return rc;
    void* base = nullptr;
    if ((rc = scratch_address(m, m->device, &base))) return rc;
    bf16_t* X0 = static_cast<bf16_t*>(base);
    bf16_t* Ya = X0 + (size_t)R * kPad;
    bf16_t* Yb = Ya + (size_t)R * m->width;
    float* out = reinterpret_cast<float*>(Yb + (size_t)R * m->width);
    for (long s = 0; s < n; s += R) {
        const int cnt = (int)std::min<long>(R, n - s);
        const int B = (cnt + 511) / 512 * 512;
        hipLaunchKernelGGL(k_pack_rows, dim3(cdiv((long)B * kPad, kThreads)), dim3(kThreads), 0, m->stream,
                           x + (size_t)s * m->n_in, cnt, B, m->n_in, X0);
        forward_infer(m, B, X0, Ya, Yb, out);
        hipLaunchKernelGGL(k_unpack_rows, dim3(cdiv((long)cnt * m->n_out, kThreads)), dim3(kThreads), 0, m->stream, out,
                           cnt, m->n_out, y + (size_t)s * m->n_out);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) return fail(m, E3D_EHIP, std::string("e3d_mlp_predict_noalloc: ") + hipGetErrorString(e));
    return E3D_OK;
}

// Forward + backward of one mini-batch, timed with HIP events: milliseconds per
// training step and the step's matrix-core flops (for the roofline).
int e3d_mlp_time_step(e3d_mlp* m, int B, int reps, double* ms, double* flops) {
    if (!m || !ms || !flops || reps < 1 || !m->X) return e3d::set_error(nullptr, E3D_EINVAL, "e3d_mlp_time_step: bad argument");
    int rc = check_batch(m, B);
    if (rc) return rc;
    MLP_TRY(m, hipSetDevice(m->device));
    if ((rc = ensure_work(m, B))) return rc;
    std::vector<int> idx(B);
    for (int b = 0; b < B; ++b) idx[b] = (int)(b % m->n_data);
    MLP_TRY(m, hipMemcpyAsync(m->idx, idx.data(), sizeof(int) * B, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_gather, dim3(cdiv(B, kGatherSamples)), dim3(kThreads), 0, m->stream, m->X, m->T, m->idx, m->n_in,
                       m->n_out, B, m->X0, m->X0t, m->Tb);
    hipEvent_t a, b;
    MLP_TRY(m, hipEventCreate(&a));
    MLP_TRY(m, hipEventCreate(&b));
    forward(m, B);
    backward(m, B);
    MLP_TRY(m, hipEventRecord(a, m->stream));
    for (int r = 0; r < reps; ++r) {
        forward(m, B);
        backward(m, B);
    }
    MLP_TRY(m, hipEventRecord(b, m->stream));
    MLP_TRY(m, hipEventSynchronize(b));
    float t = 0.0f;
    MLP_TRY(m, hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t / reps;
    double f = 0.0;  // 2 flops per multiply-add, padded sizes as computed
    for (int l = 0; l < m->nl; ++l) {
        const double mk = (double)m->din[l] * m->dout[l] * B;
        f += 2.0 * mk * (l == 0 ? 2.0 : 3.0);  // forward, weight gradient, input gradient (not for layer 0)
    }
    *flops = f;
    return E3D_OK;
}

}  // extern "C"
