// kt_learn.hip -- the learned count-kernel picker: phase features and a
// small FP32 MLP (16 -> 64 -> 64 -> C) with forward, backward and Adam
// kernels.  The network is tiny (about 6K parameters), so every kernel maps
// one sample to one wave of 64 lanes (lane j owns hidden unit j); gradients
// are reduced deterministically (per-workgroup partials summed in a fixed
// order).  No MFMA: these contractions are 64 wide and latency-bound.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kt_internal.h"
#include "kt_learn.h"

namespace kt {
namespace {

constexpr int kIn = KT_MLP_IN;
constexpr int kHid = KT_MLP_HIDDEN;

// Parameter layout (floats): W1[kHid][kIn] b1[kHid] W2[kHid][kHid] b2[kHid] W3[C][kHid] b3[C]
struct Layout {
    int w1, b1, w2, b2, w3, b3, total;
};
__host__ __device__ inline Layout layout(int C) {
    Layout L;
    L.w1 = 0;
    L.b1 = L.w1 + kHid * kIn;
    L.w2 = L.b1 + kHid;
    L.b2 = L.w2 + kHid * kHid;
    L.w3 = L.b2 + kHid;
    L.b3 = L.w3 + C * kHid;
    L.total = L.b3 + C;
    return L;
}

// ---------------------------------------------------------------------------
// Phase features.  Every workgroup of the two feature kernels (kFeatBlocks of
// them, grid-stride) writes its own partials; k_feat_finish sums them in a
// fixed order (no atomics, deterministic).  Partials (u64) per workgroup:
// 0 rows, 1 sum len, 2 sum len^2, 3 max len, 4 entries in rows > 256,
// 5 entries in rows > 1024 (stored rows), 6..10 support histogram bins,
// 11 live entries.
constexpr uint32_t kFeatBlocks = KT_FEAT_BLOCKS;
constexpr uint32_t kFeatThreads = 1024;  // 16 waves per CU: the scan is latency-bound

template <int N>
__device__ void block_sum_store(unsigned long long (&v)[N], bool is_max_3,
                                unsigned long long* out) {
    __shared__ unsigned long long s[kFeatThreads / 64][N];
    for (int j = 0; j < N; ++j)
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(v[j], off);
            v[j] = (is_max_3 && j == 3) ? (o > v[j] ? o : v[j]) : v[j] + o;
        }
    if ((threadIdx.x & 63) == 0)
        for (int j = 0; j < N; ++j) s[threadIdx.x / 64][j] = v[j];
    __syncthreads();
    if (threadIdx.x < (unsigned)N) {
        const int j = threadIdx.x;
        unsigned long long r = s[0][j];
        for (uint32_t w = 1; w < kFeatThreads / 64; ++w)
            r = (is_max_3 && j == 3) ? (s[w][j] > r ? s[w][j] : r) : r + s[w][j];
        out[j] = r;
    }
}

// Entries are sampled in 64-entry chunks, one chunk every `stride` (every
// entry up to kFeatSample entries; the histogram is normalised and the live
// count rescaled), so the features cost a bounded, coalesced pass however
// large the graph.  One kernel: the rows (live lengths when given, else the
// stored ones) and the sampled entries (dst, sup).
constexpr uint32_t kFeatSample = 1u << 22;
constexpr uint32_t kRowItems = 8;    // rows per thread whose loads are in flight together
constexpr uint32_t kChunkItems = 8;  // sampled chunks per wave whose loads are in flight together

__global__ __launch_bounds__(kFeatThreads) void k_feat_scan(const uint32_t* __restrict__ ptr,
                                                           const uint32_t* __restrict__ rowlen, uint32_t n,
                                                           const uint32_t* __restrict__ dst,
                                                           const uint32_t* __restrict__ sup, uint32_t E,
                                                           uint32_t stride, uint32_t k, int sup_valid,
                                                           unsigned long long* part) {
    unsigned long long v[6] = {0, 0, 0, 0, 0, 0};  // rows, s1, s2, max, big, huge
    // kRowItems rows per thread with their loads in flight together (rows past
    // the end read row n - 1 and count as empty)
    const uint32_t nth = gridDim.x * blockDim.x;
    for (uint32_t r0 = blockIdx.x * blockDim.x + threadIdx.x; r0 < n; r0 += kRowItems * nth) {
        uint32_t len[kRowItems];
#pragma unroll
        for (uint32_t q = 0; q < kRowItems; ++q) {
            const uint32_t r = r0 + q * nth, rr = r < n ? r : n - 1;
            const uint32_t l = rowlen ? rowlen[rr] : ptr[rr + 1] - ptr[rr];
            len[q] = r < n ? l : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < kRowItems; ++q) {
            const uint32_t l = len[q];
            if (!l) continue;
            v[0] += 1;
            v[1] += l;
            v[2] += (unsigned long long)l * l;
            v[3] = l > v[3] ? l : v[3];
            if (l > 256) v[4] += l;
            if (l > 1024) v[5] += l;
        }
    }
    block_sum_store<6>(v, true, part + (size_t)blockIdx.x * KT_FEAT_ACC);
    __syncthreads();  // block_sum_store's LDS is reused below
    unsigned long long w[6] = {0, 0, 0, 0, 0, 0};  // 5 bins, live entries (sampled)
    const float denom = (float)(k > 2 ? k - 2 : 1);
    const uint32_t lane = threadIdx.x & 63;
    // kChunkItems sampled chunks per round, their loads in flight together
    // (indices past the end read entry E - 1 and are not counted)
    const uint64_t cstep = (uint64_t)gridDim.x * (blockDim.x / 64) * stride * 64;
    for (uint64_t c = ((uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) * stride * 64; c < E;
         c += kChunkItems * cstep) {
        uint32_t dv[kChunkItems], sv[kChunkItems];
        bool in[kChunkItems];
#pragma unroll
        for (uint32_t q = 0; q < kChunkItems; ++q) {
            const uint64_t e = c + q * cstep + lane;
            in[q] = e < E;
            const uint64_t ee = in[q] ? e : E - 1;
            dv[q] = dst[ee];
            sv[q] = sup[ee];
        }
#pragma unroll
        for (uint32_t q = 0; q < kChunkItems; ++q) {
            if (!in[q] || (dv[q] & (KT_DEAD | KT_DYING))) continue;
            w[5] += 1;
            if (sup_valid) {
                const float r = (float)sv[q] / denom;
                const int b = r < 1.f ? 0 : r < 2.f ? 1 : r < 4.f ? 2 : r < 8.f ? 3 : 4;
                w[b] += 1;
            }
        }
    }
    block_sum_store<6>(w, false, part + (size_t)blockIdx.x * KT_FEAT_ACC + 6);
}

__device__ float log2p1(float x) { return log2f(1.f + x); }

// ---------------------------------------------------------------------------
// MLP forward for one sample in one wave; activations land in LDS.
__device__ void mlp_forward_wave(const float* __restrict__ P, int C, const float* x, float* s_h1,
                                 float* s_h2, float* out, int lane) {
    const Layout L = layout(C);
    float a = P[L.b1 + lane];
    for (int i = 0; i < kIn; ++i) a += P[L.w1 + lane * kIn + i] * x[i];
    s_h1[lane] = a > 0.f ? a : 0.f;
    __syncthreads();
    float b = P[L.b2 + lane];
    for (int i = 0; i < kHid; ++i) b += P[L.w2 + lane * kHid + i] * s_h1[i];
    s_h2[lane] = b > 0.f ? b : 0.f;
    __syncthreads();
    if (lane < C) {
        float o = P[L.b3 + lane];
        for (int i = 0; i < kHid; ++i) o += P[L.w3 + lane * kHid + i] * s_h2[i];
        out[lane] = o;
    }
    __syncthreads();
}

// Inference inputs are clamped to the model's input range R = [lo[16], hi[16]]
// (the range of its training features), so the network never extrapolates.
__device__ __forceinline__ float clamp_in(float x, const float* R, int i) {
    return fminf(fmaxf(x, R[i]), R[kIn + i]);
}

__global__ __launch_bounds__(64) void k_mlp_forward(const float* __restrict__ P, int C,
                                                    const float* __restrict__ R,
                                                    const float* __restrict__ X, uint32_t batch,
                                                    float* __restrict__ Y) {
    __shared__ float s_x[kIn], s_h1[kHid], s_h2[kHid], s_o[KT_MLP_MAX_OUT];
    const int lane = threadIdx.x;
    for (uint32_t s = blockIdx.x; s < batch; s += gridDim.x) {
        if (lane < kIn) s_x[lane] = clamp_in(X[(size_t)s * kIn + lane], R, lane);
        __syncthreads();
        mlp_forward_wave(P, C, s_x, s_h1, s_h2, s_o, lane);
        if (lane < C) Y[(size_t)s * C + lane] = s_o[lane];
        __syncthreads();
    }
}

// Loss: sum_s w_s * sum_c mask_sc (y_sc - t_sc)^2 / sum_s w_s.  Each
// workgroup (one wave) accumulates its samples' gradients into a partial row
// part[blockIdx.x][P] and its loss into part_loss[blockIdx.x].
__global__ __launch_bounds__(64) void k_mlp_backward(const float* __restrict__ P, int C,
                                                     const float* __restrict__ X,
                                                     const float* __restrict__ T,
                                                     const float* __restrict__ M,
                                                     const float* __restrict__ Wt, uint32_t batch,
                                                     float inv_wsum, float* __restrict__ part,
                                                     float* __restrict__ part_loss) {
    extern __shared__ float g[];  // Layout(C).total floats of gradient
    __shared__ float s_x[kIn], s_h1[kHid], s_h2[kHid], s_o[KT_MLP_MAX_OUT], s_do[KT_MLP_MAX_OUT];
    __shared__ float s_d2[kHid], s_d1[kHid];
    const Layout L = layout(C);
    const int lane = threadIdx.x;
    for (int p = lane; p < L.total; p += 64) g[p] = 0.f;
    float loss = 0.f;
    __syncthreads();
    for (uint32_t s = blockIdx.x; s < batch; s += gridDim.x) {
        if (lane < kIn) s_x[lane] = X[(size_t)s * kIn + lane];
        __syncthreads();
        mlp_forward_wave(P, C, s_x, s_h1, s_h2, s_o, lane);
        const float w = Wt ? Wt[s] : 1.f;
        if (lane < C) {
            const float m = M ? M[(size_t)s * C + lane] : 1.f;
            const float diff = (s_o[lane] - T[(size_t)s * C + lane]) * m;
            loss += w * diff * diff;
            s_do[lane] = 2.f * w * diff * inv_wsum;
        }
        __syncthreads();
        // layer 3: dW3[c][lane] += do[c] * h2[lane]; dh2[lane]
        float dh2 = 0.f;
        for (int c = 0; c < C; ++c) {
            g[L.w3 + c * kHid + lane] += s_do[c] * s_h2[lane];
            dh2 += P[L.w3 + c * kHid + lane] * s_do[c];
        }
        if (lane < C) g[L.b3 + lane] += s_do[lane];
        s_d2[lane] = s_h2[lane] > 0.f ? dh2 : 0.f;
        __syncthreads();
        // layer 2: dW2[j][i] += d2[j] * h1[i] (lane = i); dh1[i] = sum_j W2[j][i] d2[j]
        float dh1 = 0.f;
        for (int j = 0; j < kHid; ++j) {
            g[L.w2 + j * kHid + lane] += s_d2[j] * s_h1[lane];
            dh1 += P[L.w2 + j * kHid + lane] * s_d2[j];
        }
        g[L.b2 + lane] += s_d2[lane];
        s_d1[lane] = s_h1[lane] > 0.f ? dh1 : 0.f;
        __syncthreads();
        // layer 1: dW1[lane][i] += d1[lane] * x[i]
        for (int i = 0; i < kIn; ++i) g[L.w1 + lane * kIn + i] += s_d1[lane] * s_x[i];
        g[L.b1 + lane] += s_d1[lane];
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    for (int p = lane; p < L.total; p += 64) part[(size_t)blockIdx.x * L.total + p] = g[p];
    if (lane == 0) part_loss[blockIdx.x] = loss * inv_wsum;
}

// Fixed-order reduction of the partial rows.
__global__ void k_reduce_parts(const float* __restrict__ part, uint32_t nparts, uint32_t total,
                               float* __restrict__ grad) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (uint32_t b = 0; b < nparts; ++b) s += part[(size_t)b * total + p];
        grad[p] = s;
    }
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ m1,
                       float* __restrict__ m2, uint32_t total, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        const float g = G[p];
        const float a = b1 * m1[p] + (1.f - b1) * g;
        const float b = b2 * m2[p] + (1.f - b2) * g * g;
        m1[p] = a;
        m2[p] = b;
        P[p] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    }
}

// The phase features from the workgroup partials (one wave); with P, the
// model's pick on them as well (one launch per phase).
__global__ __launch_bounds__(64) void k_feat_finish(const unsigned long long* part, uint32_t E0,
                                                    uint32_t stride, uint32_t k, int sup_valid,
                                                    float* feat, const float* __restrict__ P, int C,
                                                    const float* __restrict__ R, unsigned int* choice,
                                                    uint32_t* host_choice) {
    __shared__ unsigned long long acc[KT_FEAT_ACC];
    __shared__ float s_x[kIn], s_h1[kHid], s_h2[kHid], s_o[KT_MLP_MAX_OUT];
    const int lane = threadIdx.x;
    // every lane folds kFeatBlocks / 64 workgroups (all of their partials
    // loaded before the first wait), then the wave reduces
    static_assert(kFeatBlocks % 64 == 0, "whole waves of partials");
    unsigned long long x[KT_FEAT_ACC][kFeatBlocks / 64];
#pragma unroll
    for (int j = 0; j < KT_FEAT_ACC; ++j)
#pragma unroll
        for (uint32_t q = 0; q < kFeatBlocks / 64; ++q) x[j][q] = part[(size_t)(lane + 64 * q) * KT_FEAT_ACC + j];
#pragma unroll
    for (int j = 0; j < KT_FEAT_ACC; ++j) {
        unsigned long long r = 0;
#pragma unroll
        for (uint32_t q = 0; q < kFeatBlocks / 64; ++q) r = j == 3 ? (x[j][q] > r ? x[j][q] : r) : r + x[j][q];
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long o = __shfl_xor(r, off);
            r = j == 3 ? (o > r ? o : r) : r + o;
        }
        if (lane == 0) acc[j] = r;
    }
    __syncthreads();
    if (lane == 0) {
        float f[kIn];
        const float rows = (float)acc[0], s1 = (float)acc[1], s2 = (float)acc[2];
        const float mean = rows > 0 ? s1 / rows : 0.f;
        f[0] = log2p1(s1) / 32.f;
        f[1] = log2p1(rows) / 32.f;
        f[2] = mean / 256.f;
        f[3] = log2p1((float)acc[3]) / 16.f;
        f[4] = mean > 0 ? sqrtf(s2 / rows) / mean / 8.f : 0.f;
        f[5] = s1 > 0 ? (float)acc[4] / s1 : 0.f;
        f[6] = s1 > 0 ? (float)acc[5] / s1 : 0.f;
        f[7] = E0 ? fminf(1.f, (float)acc[11] * (float)stride / (float)E0) : 0.f;  // live fraction
        f[8] = log2f((float)k) / 10.f;
        f[9] = s1 > 0 ? s2 / s1 / 1024.f : 0.f;
        float tot = 0;
        for (int b = 0; b < 5; ++b) tot += (float)acc[6 + b];
        for (int b = 0; b < 5; ++b) f[10 + b] = (sup_valid && tot > 0) ? (float)acc[6 + b] / tot : 0.f;
        f[15] = sup_valid ? 1.f : 0.f;
        for (int i = 0; i < kIn; ++i) {
            feat[i] = f[i];
            s_x[i] = P ? clamp_in(f[i], R, i) : 0.f;
        }
    }
    __syncthreads();
    if (!P) return;  // uniform
    mlp_forward_wave(P, C, s_x, s_h1, s_h2, s_o, lane);
    if (lane == 0) {
        int best = 0;
        for (int c = 1; c < C; ++c)
            if (s_o[c] < s_o[best]) best = c;
        *choice = (unsigned int)best;
        if (host_choice) __hip_atomic_store(host_choice, (uint32_t)best, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

}  // namespace

namespace dev {

uint32_t mlp_num_params(uint32_t C) { return (uint32_t)layout((int)C).total; }

hipError_t phase_features(const uint32_t* ptr, const uint32_t* rowlen, const uint32_t* dst,
                          const uint32_t* sup, uint32_t n, uint32_t E, uint32_t E0, uint32_t k,
                          bool sup_valid, unsigned long long* acc, float* feat, const float* P,
                          uint32_t C, const float* R, unsigned int* choice, hipStream_t s,
                          uint32_t* host_choice) {
    const uint32_t stride = E > kFeatSample ? (E + kFeatSample - 1) / kFeatSample : 1u;
    hipLaunchKernelGGL(k_feat_scan, dim3(kFeatBlocks), dim3(kFeatThreads), 0, s, ptr, rowlen, n, dst, sup, E,
                       stride, k, sup_valid ? 1 : 0, acc);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_feat_finish, dim3(1), dim3(64), 0, s, acc, E0, stride, k, sup_valid ? 1 : 0,
                       feat, P, (int)C, R, choice, host_choice);
    return hipGetLastError();
}

hipError_t mlp_forward(const float* P, uint32_t C, const float* R, const float* X, uint32_t batch,
                       float* Y, hipStream_t s) {
    if (!batch) return hipSuccess;
    const uint32_t gr = std::min<uint32_t>(batch, 4096);
    hipLaunchKernelGGL(k_mlp_forward, dim3(gr), dim3(64), 0, s, P, (int)C, R, X, batch, Y);
    return hipGetLastError();
}

hipError_t mlp_loss_grad(const float* P, uint32_t C, const float* X, const float* T, const float* M,
                         const float* W, uint32_t batch, float inv_wsum, float* part,
                         float* part_loss, uint32_t nparts, float* grad, hipStream_t s) {
    const uint32_t total = mlp_num_params(C);
    hipLaunchKernelGGL(k_mlp_backward, dim3(nparts), dim3(64), total * sizeof(float), s, P, (int)C,
                       X, T, M, W, batch, inv_wsum, part, part_loss);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_reduce_parts, dim3((total + 255) / 256), dim3(256), 0, s, part, nparts,
                       total, grad);
    return hipGetLastError();
}

hipError_t adam_step(float* P, const float* G, float* m1, float* m2, uint32_t total, float lr,
                     uint32_t t, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((total + 255) / 256), dim3(256), 0, s, P, G, m1, m2, total, lr,
                       b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace dev
}  // namespace kt
