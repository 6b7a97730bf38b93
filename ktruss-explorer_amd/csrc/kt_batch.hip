// kt_batch.hip -- gfx950 kernels of the batched decomposition.
//
// A batch of G graphs is stored as one disjoint union of oriented COO+CSR
// layouts: graph g owns the entries [gbeg[g], gend[g]) and a vertex range of
// its own, so the union's rows never cross graphs and the support count of
// the single-graph engine runs on it unchanged.  What differs per graph is
// the peel state (k, live entries, phase), kept in device arrays and advanced
// by a one-workgroup control kernel.  One peel step of every unfinished graph
// is
//     mark (active list) -> control -> expand -> zero (dirty list) -> count
// and every kernel reads its sizes and task lists from device memory, so the
// host captures a run of steps once as a HIP graph and replays it, reading
// one counter block per replay to learn whether graphs are left.
//
// Per graph the steps follow the single-graph schedule (kernel_ddc.cu:116-214
// with the JUMP / REUSE schedule of the engine): the mark uses supports that
// are exact for the graph's current edge set, so each graph deletes exactly
// the edges it deletes when decomposed alone and truss numbers are identical.
#include <hip/hip_runtime.h>

#include "kt_batch.h"

namespace kt {
namespace {

constexpr uint32_t kWave = 64;

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(x, off);
        x = o < x ? o : x;
    }
    return x;
}

__device__ __forceinline__ uint32_t gwave() {
    return __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
}
__device__ __forceinline__ uint32_t nwaves() { return gridDim.x * blockDim.x / kWave; }

// clearIterInfoOnDevice's support reset (info.cu:75) for the graphs about to
// be recounted.
__global__ void k_b_zero(TaskList tl, const uint32_t* __restrict__ dst, uint32_t* __restrict__ sup) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t n = *tl.ntasks;
    for (uint32_t t = gwave(); t < n; t += nwaves()) {
        const uint4 d = tl.desc[t];
        for (uint32_t i = d.x + lane; i < d.y; i += kWave)
            if (!(dst[i] & KT_DEAD)) sup[i] = 0;
    }
}

// mark_deleted_edges_kernel (kernel_ddc.cu:56-77) with the graph's own k; the
// minimum support of the entries that stay alive is folded per graph, which
// is the JUMP target whenever the graph's phase ends with this step.  DELTA:
// a deleted entry is flagged KT_DYING and listed in dl (count *dn) for the
// support update that follows.  Each wave lists its deletions in its own LDS
// slice and the workgroup reserves room in dl with one atomic at the end (a
// wave whose slice fills reserves for it alone): a reservation per 64-entry
// chunk would put one returning atomic per chunk on a single address.
// A deleted entry's support is KT_SUP_GONE from its DELTA mark on, so an
// entry below k - 2 is live: its layout word is loaded with the other words
// its flagging needs (a flag found there is counted in *bad, never expected),
// and flags are stored, not read, modified and written (a live entry's slots
// hold plain ids: sa[i], in src[i]'s list, holds dst[i]; sb[i] holds src[i]).
// The previous step's deletions are finalized by kBFinBlocks workgroups of
// their own beside the scan, kBFinItems entries per thread in flight (their
// supports are KT_SUP_GONE: the scan never reads their flags).
constexpr uint32_t kBMarkCap = 1024;  // LDS list entries per wave
constexpr uint32_t kMarkGroup = 8;    // 64-entry chunks whose loads are in flight together
constexpr uint32_t kBFinBlocks = 64;
constexpr uint32_t kBFinItems = 8;

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool DELTA>
__global__ __launch_bounds__(256) void k_b_mark(TaskList tl, uint32_t* __restrict__ dst,
                                                uint32_t* __restrict__ sup,
                                                uint32_t* __restrict__ tau, const uint32_t* __restrict__ gk,
                                                uint32_t* gndel, uint32_t* gmin, uint32_t* __restrict__ dl,
                                                unsigned int* dn, uint32_t* __restrict__ snbr,
                                                const uint32_t* __restrict__ sa,
                                                const uint32_t* __restrict__ sb,
                                                const uint32_t* __restrict__ fin_dl,
                                                const unsigned int* __restrict__ fin_nd,
                                                const uint32_t* __restrict__ src, unsigned int* bad,
                                                uint32_t mark_blocks) {
    __shared__ uint32_t s_list[DELTA ? (256 / kWave) * kBMarkCap : 1];
    if (DELTA && blockIdx.x >= mark_blocks) {
        // the previous step's deletions: KT_DEAD, entry and both slots
        if (!fin_dl) return;
        const uint32_t nf = *fin_nd;
        const uint32_t stride = (gridDim.x - mark_blocks) * blockDim.x;
        for (uint32_t t0 = (blockIdx.x - mark_blocks) * blockDim.x + threadIdx.x; t0 < nf;
             t0 += kBFinItems * stride) {
            uint32_t f[kBFinItems], fu[kBFinItems], fv[kBFinItems], fa[kBFinItems], fb[kBFinItems];
#pragma unroll
            for (uint32_t r = 0; r < kBFinItems; ++r) {
                const uint32_t t = t0 + r * stride;
                f[r] = fin_dl[t < nf ? t : nf - 1];
            }
#pragma unroll
            for (uint32_t r = 0; r < kBFinItems; ++r) {
                fu[r] = src[f[r]];
                fv[r] = dst[f[r]] & KT_IDMASK;
                fa[r] = sa[f[r]];
                fb[r] = sb[f[r]];
            }
#pragma unroll
            for (uint32_t r = 0; r < kBFinItems; ++r) {
                if (t0 + r * stride < nf) {
                    dst[f[r]] = fv[r] | KT_DEAD;
                    snbr[fa[r]] = fv[r] | KT_DEAD;
                    snbr[fb[r]] = fu[r] | KT_DEAD;
                }
            }
        }
        return;
    }
    __shared__ uint32_t s_n[256 / kWave], s_base;
    const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t* my = s_list + (DELTA ? wid * kBMarkCap : 0);
    uint32_t cnt = 0;  // wave-uniform length of the wave's LDS list
    const uint32_t n = *tl.ntasks;
    const uint32_t nw = mark_blocks * (256 / kWave);
    for (uint32_t t = gwave(); t < n; t += nw) {
        const uint4 d = tl.desc[t];
        const uint32_t k = gk[d.z];
        uint32_t ndel = 0, smin = 0xFFFFFFFFu;
        // the deletions of one 64-entry chunk: LDS list (DELTA), counts
        auto take = [&](uint32_t i, bool del) {
            const uint64_t m = ballot(del);
            if (DELTA && m) {
                if (cnt + kWave > kBMarkCap) {  // the slice is full: reserve for it alone
                    uint32_t b = 0;
                    if (lane == 0) b = atomicAdd(dn, cnt);
                    b = __shfl(b, 0);
                    wave_sync_lds();
                    for (uint32_t q = lane; q < cnt; q += kWave) dl[b + q] = my[q];
                    wave_sync_lds();
                    cnt = 0;
                }
                if (del) {
                    my[cnt + (uint32_t)__popcll(m & below)] = i;
                    const uint32_t dv = dst[i], a = sa[i], b = sb[i], u = src[i];
                    const uint32_t vid = dv & KT_IDMASK;
                    if (dv & (KT_DEAD | KT_DYING)) atomicAdd(bad, 1u);
                    dst[i] = vid | KT_DYING;  // the entry and its two adjacency slots
                    snbr[a] = vid | KT_DYING;
                    snbr[b] = u | KT_DYING;
                }
                cnt += (uint32_t)__popcll(m);
            }
            ndel += (uint32_t)__popcll(m);
        };
        if (DELTA) {
            // supports only (a deleted entry's support is KT_SUP_GONE), eight
            // chunks in flight
            const uint32_t last = d.y - 1;
            for (uint32_t base = d.x; base < d.y; base += kMarkGroup * kWave) {
                uint32_t sp[kMarkGroup];
#pragma unroll
                for (uint32_t j = 0; j < kMarkGroup; ++j) {
                    const uint32_t i = base + j * kWave + lane;
                    const uint32_t x = sup[i < last ? i : last];
                    sp[j] = i < d.y ? x : KT_SUP_GONE;
                }
#pragma unroll
                for (uint32_t j = 0; j < kMarkGroup; ++j) {
                    const uint32_t i = base + j * kWave + lane;
                    const bool del = sp[j] < k - 2;
                    if (del) {
                        sup[i] = KT_SUP_GONE;
                        tau[i] = k - 1;
                    } else if (sp[j] >= k - 2 && sp[j] < smin) {
                        smin = sp[j];
                    }
                    take(i, del);
                }
            }
        } else {
            for (uint32_t base = d.x; base < d.y; base += kWave) {
                const uint32_t i = base + lane;
                bool del = false;
                if (i < d.y) {
                    const uint32_t v = dst[i];
                    if (!(v & KT_DEAD)) {
                        const uint32_t s = sup[i];
                        if (s < k - 2) {
                            del = true;
                            dst[i] = v | KT_DEAD;
                            tau[i] = k - 1;
                        } else if (s < smin) {
                            smin = s;
                        }
                    }
                }
                take(i, del);
            }
        }
        smin = wave_min(smin);
        if (lane == 0) {
            if (ndel) atomicAdd(gndel + d.z, ndel);
            if (smin != 0xFFFFFFFFu) atomicMin(gmin + d.z, smin);
        }
    }
    if (DELTA) {
        if (lane == 0) s_n[wid] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (uint32_t w = 0; w < 256 / kWave; ++w) {
                const uint32_t c = s_n[w];
                s_n[w] = tot;
                tot += c;
            }
            s_base = tot ? atomicAdd(dn, tot) : 0u;
        }
        __syncthreads();
        const uint32_t b = s_base + s_n[wid];
        for (uint32_t q = lane; q < cnt; q += kWave) dl[b + q] = my[q];
    }
}

// Minimum live support per graph (the first phase's JUMP, after the first
// full count).
__global__ void k_b_min(TaskList tl, const uint32_t* __restrict__ dst,
                        const uint32_t* __restrict__ sup, uint32_t* gmin) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t n = *tl.ntasks;
    for (uint32_t t = gwave(); t < n; t += nwaves()) {
        const uint4 d = tl.desc[t];
        uint32_t smin = 0xFFFFFFFFu;
        for (uint32_t i = d.x + lane; i < d.y; i += kWave)
            if (!(dst[i] & KT_DEAD) && sup[i] < smin) smin = sup[i];
        smin = wave_min(smin);
        if (lane == 0 && smin != 0xFFFFFFFFu) atomicMin(gmin + d.z, smin);
    }
}

// Block-wide exclusive scan of one value per thread (1024 threads).
__device__ uint32_t block_exscan(uint32_t x, uint32_t* s_wave, uint32_t& total) {
    const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    uint32_t inc = x;
    for (int off = 1; off < (int)kWave; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off);
        if ((int)lane >= off) inc += y;
    }
    if (lane == kWave - 1) s_wave[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        const uint32_t nw = blockDim.x / kWave;
        uint32_t w = lane < nw ? s_wave[lane] : 0;
        uint32_t wi = w;
        for (int off = 1; off < (int)kWave; off <<= 1) {
            const uint32_t y = __shfl_up(wi, off);
            if ((int)lane >= off) wi += y;
        }
        if (lane < nw) s_wave[lane] = wi - w;
        if (lane == nw - 1) s_wave[32] = wi;
    }
    __syncthreads();
    const uint32_t r = s_wave[wid] + inc - x;
    total = s_wave[32];
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint32_t ntasks_of(uint32_t len, uint32_t chunk) {
    return (len + chunk - 1) / chunk;
}

// The per-graph state machine and the next step's task lists.  One
// workgroup of 1024 threads; thread i owns the contiguous graph slice
// [i * per, (i + 1) * per), so the lists stay in graph order.
//   mode 0 (init) : k = 3, every non-empty graph dirty and active
//   mode 1 (jump) : after the first full count: k = max(3, min support + 3)
//   mode 2 (step) : after a mark pass: close the phase (no deletion, or the
//                   k == 3 pass, kernel_ddc.cu:183-186) with k = max(k + 1,
//                   min surviving support + 3), or schedule a recount
__global__ __launch_bounds__(1024) void k_b_control(BatchDev b, uint32_t mode) {
    __shared__ uint32_t s_wave[33];
    const uint32_t G = b.G;
    const uint32_t prev_active = b.bc->nactive;  // graphs the mark pass just served
    const uint32_t per = (G + blockDim.x - 1) / blockDim.x;
    const uint32_t g0 = min(G, threadIdx.x * per), g1 = min(G, g0 + per);
    const bool jump = (b.schedule & KT_SCHED_JUMP) != 0;
    const bool reuse = (b.schedule & KT_SCHED_REUSE) != 0;
    uint32_t nd = 0, na = 0, td = 0, ta = 0;
    unsigned long long live_counted = 0;
    for (uint32_t g = g0; g < g1; ++g) {
        uint32_t f = b.gflag[g];
        if (mode == 0) {
            const uint32_t len = b.gend[g] - b.gbeg[g];
            b.gk[g] = 3;
            b.glive[g] = len;
            b.gndel[g] = 0;
            b.gmin[g] = 0xFFFFFFFFu;
            b.giter[g] = 0;
            b.gphase[g] = 0;
            b.gkmax[g] = 0;
            b.gpass[g] = 0;
            f = len ? KT_BF_DIRTY : KT_BF_DONE;
        } else if (mode == 1) {
            if (!(f & KT_BF_DONE)) {
                const uint32_t m = b.gmin[g];
                if (jump && m != 0xFFFFFFFFu && m + 3 > b.gk[g]) b.gk[g] = m + 3;
                b.gmin[g] = 0xFFFFFFFFu;
                f = 0;  // supports exact: the next step marks
            }
        } else if (!(f & KT_BF_DONE)) {
            const uint32_t k = b.gk[g];
            const uint32_t ndel = b.gndel[g];
            const uint32_t live = b.glive[g] - ndel;
            b.glive[g] = live;
            b.giter[g] += 1;
            if (ndel) b.gkmax[g] = k - 1;
            if (live == 0) {
                f = KT_BF_DONE;
                b.gphase[g] += 1;
            } else if (ndel == 0 || k == 3) {
                const uint32_t m = b.gmin[g];
                uint32_t kn = k + 1;
                if (jump && m + 3 > kn) kn = m + 3;
                b.gk[g] = kn;
                b.gphase[g] += 1;
                f = reuse ? 0u : KT_BF_DIRTY;
            } else {
                f = b.delta ? 0u : KT_BF_DIRTY;  // DELTA keeps the supports exact
            }
            b.gndel[g] = 0;
            b.gmin[g] = 0xFFFFFFFFu;
        }
        b.gflag[g] = f;
        if (!(f & KT_BF_DONE)) {
            const uint32_t len = b.gend[g] - b.gbeg[g];
            ++na;
            ta += ntasks_of(len, b.mchunk);
            if ((f & KT_BF_DIRTY) && mode != 1) {
                ++nd;
                td += ntasks_of(len, b.cchunk);
                live_counted += b.glive[g];
                b.gpass[g] += 1;
            }
        }
    }
    uint32_t tot_nd, tot_na, tot_td, tot_ta;
    uint32_t o_nd = block_exscan(nd, s_wave, tot_nd);
    uint32_t o_na = block_exscan(na, s_wave, tot_na);
    uint32_t o_td = block_exscan(td, s_wave, tot_td);
    uint32_t o_ta = block_exscan(ta, s_wave, tot_ta);
    for (uint32_t g = g0; g < g1; ++g) {
        const uint32_t f = b.gflag[g];
        if (f & KT_BF_DONE) continue;
        const uint32_t len = b.gend[g] - b.gbeg[g];
        b.alist[o_na] = g;
        b.ascan[o_na++] = o_ta;
        o_ta += ntasks_of(len, b.mchunk);
        if ((f & KT_BF_DIRTY) && mode != 1) {
            b.dlist[o_nd] = g;
            b.dscan[o_nd++] = o_td;
            o_td += ntasks_of(len, b.cchunk);
        }
    }
    // live entries of the graphs recounted next, for the algorithmic bytes
    for (int off = 32; off > 0; off >>= 1) live_counted += __shfl_xor(live_counted, off);
    if ((threadIdx.x & (kWave - 1)) == 0 && live_counted)
        atomicAdd(&b.bc->live_in_passes, live_counted);
    if (threadIdx.x == 0) {
        if (b.zero_cnt) *b.zero_cnt = 0;  // the DELTA list of the next mark pass starts empty
        b.bc->ndirty = mode == 1 ? 0u : tot_nd;
        b.bc->nactive = tot_na;
        b.bc->dtasks = mode == 1 ? 0u : tot_td;
        b.bc->atasks = tot_ta;
        if (mode == 2 && prev_active) b.bc->steps += 1;
    }
}

// Task descriptors of both lists: one wave per listed graph writes that
// graph's tasks.
__global__ void k_b_expand(BatchDev b) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t nd = b.bc->ndirty, na = b.bc->nactive;
    for (uint32_t j = gwave(); j < nd + na; j += nwaves()) {
        const bool dirty = j < nd;
        const uint32_t jj = dirty ? j : j - nd;
        const uint32_t g = dirty ? b.dlist[jj] : b.alist[jj];
        const uint32_t base = dirty ? b.dscan[jj] : b.ascan[jj];
        const uint32_t chunk = dirty ? b.cchunk : b.mchunk;
        uint4* out = dirty ? b.ddesc : b.adesc;
        const uint32_t beg = b.gbeg[g], end = b.gend[g];
        const uint32_t nt = ntasks_of(end - beg, chunk);
        for (uint32_t i = lane; i < nt; i += kWave) {
            const uint32_t x = beg + i * chunk;
            out[base + i] = make_uint4(x, end - x < chunk ? end : x + chunk, g, 0u);
        }
    }
}

constexpr uint32_t kStreamBlock = 256;

}  // namespace

namespace dev {

hipError_t batch_control(const BatchDev& b, uint32_t mode, hipStream_t s) {
    hipLaunchKernelGGL(k_b_control, dim3(1), dim3(1024), 0, s, b, mode);
    return hipGetLastError();
}

hipError_t batch_expand(const BatchDev& b, uint32_t blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_b_expand, dim3(blocks), dim3(kStreamBlock), 0, s, b);
    return hipGetLastError();
}

hipError_t batch_zero(const TaskList& tl, const uint32_t* dst, uint32_t* sup, uint32_t blocks,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_b_zero, dim3(blocks), dim3(kStreamBlock), 0, s, tl, dst, sup);
    return hipGetLastError();
}

hipError_t batch_mark(const TaskList& tl, uint32_t* dst, uint32_t* sup, uint32_t* tau,
                      const uint32_t* gk, uint32_t* gndel, uint32_t* gmin, uint32_t* dl,
                      unsigned int* dn, uint32_t* snbr, const uint32_t* sa, const uint32_t* sb,
                      const uint32_t* fin_dl, const unsigned int* fin_nd, const uint32_t* src,
                      unsigned int* bad, uint32_t blocks, hipStream_t s) {
    if (dl)
        hipLaunchKernelGGL(k_b_mark<true>, dim3(blocks + (fin_dl ? kBFinBlocks : 0u)), dim3(kStreamBlock), 0,
                           s, tl, dst, sup, tau, gk, gndel, gmin, dl, dn, snbr, sa, sb, fin_dl, fin_nd, src,
                           bad, blocks);
    else
        hipLaunchKernelGGL(k_b_mark<false>, dim3(blocks), dim3(kStreamBlock), 0, s, tl, dst, sup, tau,
                           gk, gndel, gmin, dl, dn, snbr, sa, sb, nullptr, nullptr, nullptr, nullptr, blocks);
    return hipGetLastError();
}

hipError_t batch_min(const TaskList& tl, const uint32_t* dst, const uint32_t* sup, uint32_t* gmin,
                     uint32_t blocks, hipStream_t s) {
    hipLaunchKernelGGL(k_b_min, dim3(blocks), dim3(kStreamBlock), 0, s, tl, dst, sup, gmin);
    return hipGetLastError();
}

}  // namespace dev
}  // namespace kt
