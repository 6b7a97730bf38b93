// kt_delta.hip -- gfx950 kernels of DELTA support maintenance.
//
// Directed COO+CSR only (every undirected edge is one oriented entry e, and
// its support sup[e] is the exact number of triangles through it).  After a
// mark pass the deleted entries carry KT_DYING and sit in the list dl; the
// supports are brought to those of the surviving graph by
//   k_dprep     per deleted entry d = (u, v): the shorter and the longer of
//               the endpoints' full adjacency lists, and the tasks of 64
//               neighbours of the shorter list, numbered by an exclusive scan
//               (per tile of deletions, then over the tiles by the last
//               workgroup to finish);
//   k_delta     waves take equal ranges of tasks; each neighbour w of a task
//               is binary-searched (uniform trip count) in the longer list.
//               Work is split by list length, so a deletion between two hubs
//               is spread over many waves instead of serialising one.  A hit is a
//               triangle {u, v, w} that was intact before the pass.  It is
//               handled by the deleted edge of lowest entry index among its
//               three edges, which decrements the supports of the triangle's
//               edges that survive -- so every destroyed triangle is removed
//               exactly once from every surviving edge (k_delta<true> also
//               lists the entries whose support falls below k - 2: the next
//               deletions, without a mark pass);
//   finalize    the deleted entries become KT_DEAD in the oriented layout and
//               in both of their adjacency-list slots: in workgroups of the
//               next mark's launch, in the next task list after a light
//               iteration, or in k_finalize.
// Between its flagging and its finalization a deleted entry and both of its
// adjacency slots carry KT_DYING (bit 30; ids stay below 2^30).
// Adjacency layout (symmetric CSR over the relabelled ids):
//   sptr[n+1]   list starts
//   snbr[2E]    neighbour id, KT_DEAD set once the edge is gone; a vertex's
//               list is its in-neighbours then its out-neighbours, both
//               ascending, so the list is sorted
//   seid[2E]    oriented entry of the edge
//   sa[E], sb[E] the two list slots of each oriented entry (in u's list, in v's)
#include <hip/hip_runtime.h>

#include "kt_internal.h"

namespace kt {
namespace {

constexpr uint32_t kWave = 64;
constexpr uint32_t kBlock = 256;

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t gwave() {
    return __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
}
__device__ __forceinline__ uint32_t nwaves() { return gridDim.x * blockDim.x / kWave; }

// mark_deleted_edges_kernel (kernel_ddc.cu:56-77), deferred: a deleted entry
// is flagged KT_DYING and listed; it becomes KT_DEAD in k_finalize.
// The deleted entry's two adjacency slots get KT_DYING as well.
// The scan reads supports only: a deleted entry's support becomes
// KT_SUP_GONE (above every threshold), so an entry below k - 2 is live; its
// layout word is loaded with the other words its flagging needs (and a flag
// found there is counted in *bad, never expected).
// A workgroup walks tiles of kMarkTile entries (every lane loads all of its
// supports before the first wait) and lists its deletions in
// LDS; the global list is reserved with one atomic per workgroup when the
// LDS list is full and at the end.  A reservation per wave would put one
// returning atomic per wave on a single address, served one at a time.
// A live entry's layout word and adjacency slots hold plain ids (sa[e]: the
// slot in src[e]'s list, which holds dst[e]; sb[e]: the one holding src[e]),
// so flags are stored, not read, modified and written.
// The finalization of a previous list runs in workgroups of its own beside
// the scan (kFinItems entries per thread in flight): its entries' supports
// are KT_SUP_GONE, so the scan never reads their flags.
// A deleted entry whose update is done: KT_DEAD in the oriented layout and
// in both of its adjacency slots.
__device__ __forceinline__ void finalize_one(uint32_t d, uint32_t* dst, uint32_t* snbr,
                                             const uint32_t* sa, const uint32_t* sb) {
    dst[d] = (dst[d] & ~KT_DYING) | KT_DEAD;
    const uint32_t a = sa[d], b = sb[d];
    snbr[a] = (snbr[a] & ~KT_DYING) | KT_DEAD;
    snbr[b] = (snbr[b] & ~KT_DYING) | KT_DEAD;
}

constexpr uint32_t kMarkItems = 8;
// a listed deletion that lies in no triangle (support 0): flag bit of its
// LDS list word (entries stay below 2^31: DELTA plans have fewer)
constexpr uint32_t kMarkNoTri = 0x80000000u;
constexpr uint32_t kMarkTile = kBlock * kMarkItems;
constexpr uint32_t kMarkBuf = 2 * kMarkTile;
constexpr uint32_t kMarkGrid = 1024;
constexpr uint32_t kFinItems = 8;
constexpr uint32_t kFinGrid = 64;  // most finalizing workgroups of a mark launch

// Finalizes fin_dl[0, *fin_nd) with nb workgroups (this one is fb of them).
__device__ __forceinline__ void finalize_list(const uint32_t* __restrict__ fin_dl,
                                              const uint32_t* __restrict__ fin_nd, uint32_t fb, uint32_t nb,
                                              const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                              uint32_t* __restrict__ snbr, const uint32_t* __restrict__ sa,
                                              const uint32_t* __restrict__ sb) {
    const uint32_t nf = *fin_nd;
    const uint32_t stride = nb * kBlock;
    for (uint32_t t0 = fb * kBlock + threadIdx.x; t0 < nf; t0 += kFinItems * stride) {
        uint32_t f[kFinItems], fu[kFinItems], fv[kFinItems], fa[kFinItems], fc[kFinItems];
#pragma unroll
        for (uint32_t r = 0; r < kFinItems; ++r) {
            const uint32_t t = t0 + r * stride;
            f[r] = fin_dl[t < nf ? t : nf - 1];
        }
#pragma unroll
        for (uint32_t r = 0; r < kFinItems; ++r) {
            fu[r] = src[f[r]];
            fv[r] = dst[f[r]] & KT_IDMASK;
            fa[r] = sa[f[r]];
            fc[r] = sb[f[r]];
        }
#pragma unroll
        for (uint32_t r = 0; r < kFinItems; ++r) {
            if (t0 + r * stride < nf) {
                dst[f[r]] = fv[r] | KT_DEAD;
                snbr[fa[r]] = fv[r] | KT_DEAD;
                snbr[fc[r]] = fu[r] | KT_DEAD;
            }
        }
    }
}

__device__ __forceinline__ void mark_flush(const uint32_t* s_out, uint32_t* s_cnt, uint32_t* s_base,
                                           uint32_t* dl, unsigned int* ndel) {
    const uint32_t c = *s_cnt;  // every thread reads it after a barrier and before anyone changes it
    if (!c) return;
    if (threadIdx.x == 0) *s_base = atomicAdd(ndel, c);
    __syncthreads();
    const uint32_t base = *s_base;
    for (uint32_t i = threadIdx.x; i < c; i += kBlock) dl[base + i] = s_out[i] & ~kMarkNoTri;
    __syncthreads();
    if (threadIdx.x == 0) *s_cnt = 0;
    __syncthreads();
}

// The flush of a listing mark: the staged task counts of the c buffered
// deletions are scanned (thread t owns slots [16t, 16t + 16)), one 64-bit
// atomic reserves the list slots and the task range, then the deletions,
// their first tasks and their descriptors are written out.
constexpr uint32_t kFlushPer = kMarkBuf / kBlock;
__device__ __forceinline__ void mark_flush_tasks(const uint32_t* s_out, uint32_t* s_cnt, uint32_t* s_res,
                                                 uint32_t* s_w, uint32_t* dl, const dev::MarkTasks& mt) {
    const uint32_t c = *s_cnt;  // every thread reads it after a barrier and before anyone changes it
    if (!c) return;
    const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const uint32_t stage = blockIdx.x * kMarkBuf;
    const uint32_t i0 = threadIdx.x * kFlushPer;
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t r = 0; r < kFlushPer; ++r) sum += i0 + r < c ? mt.ws.ttmp[stage + i0 + r] : 0u;
    uint32_t inc = sum;  // inclusive scan over the workgroup
    for (int off = 1; off < (int)kWave; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off);
        if ((int)lane >= off) inc += y;
    }
    if (lane == kWave - 1) s_w[wid] = inc;
    __syncthreads();
    uint32_t run = inc - sum, total = 0;
    for (uint32_t w = 0; w < kBlock / kWave; ++w) {
        if (w < wid) run += s_w[w];
        total += s_w[w];
    }
    if (threadIdx.x == 0) {
        const unsigned long long old = atomicAdd(mt.pair, ((unsigned long long)total << 32) | c);
        s_res[0] = (uint32_t)old;
        s_res[1] = (uint32_t)(old >> 32);
    }
    __syncthreads();
    const uint32_t base = s_res[0];
    uint32_t ft = s_res[1] + run;
    for (uint32_t r = 0; r < kFlushPer && i0 + r < c; ++r) {  // the counts again (cached)
        dl[base + i0 + r] = s_out[i0 + r] & ~kMarkNoTri;
        mt.ws.ntask[base + i0 + r] = ft;
        ft += mt.ws.ttmp[stage + i0 + r];
    }
    for (uint32_t i = threadIdx.x; i < c; i += kBlock) mt.ws.meta[base + i] = mt.ws.mtmp[stage + i];
    __syncthreads();
    if (threadIdx.x == 0) *s_cnt = 0;
    __syncthreads();
}

__global__ __launch_bounds__(kBlock, 8) void k_mark_dying(uint32_t* __restrict__ dst,
                                                       const uint32_t* __restrict__ eid,
                                                       uint32_t* __restrict__ sup,
                                                       uint32_t* __restrict__ tau,
                                                       uint32_t* __restrict__ dl, uint32_t n, uint32_t k,
                                                       uint32_t* __restrict__ snbr,
                                                       const uint32_t* __restrict__ sa,
                                                       const uint32_t* __restrict__ sb, unsigned int* ndel,
                                                       const uint32_t* __restrict__ fin_dl,
                                                       const uint32_t* __restrict__ fin_nd,
                                                       const uint32_t* __restrict__ src,
                                                       uint32_t* __restrict__ ldeg, unsigned int* bad,
                                                       uint32_t mark_blocks, dev::MarkTasks mt) {
    __shared__ uint32_t s_out[kMarkBuf];
    __shared__ uint32_t s_cnt, s_base;
    __shared__ uint32_t s_res[2], s_w[kBlock / kWave];
    const bool listing = mt.pair != nullptr;
    if (blockIdx.x >= mark_blocks) {  // the previous deletions, whose update is done, become KT_DEAD
        finalize_list(fin_dl, fin_nd, blockIdx.x - mark_blocks, gridDim.x - mark_blocks, src, dst, snbr, sa,
                      sb);
        return;
    }
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t ntile = (n + kMarkTile - 1) / kMarkTile;
    uint32_t c_prev = 0;  // s_cnt before the tile (uniform)
    for (uint32_t tile = blockIdx.x; tile < ntile; tile += mark_blocks) {
        const uint32_t e0 = tile * kMarkTile + threadIdx.x;
        uint32_t sp[kMarkItems];
#pragma unroll
        for (uint32_t j = 0; j < kMarkItems; ++j) {
            const uint32_t e = e0 + j * kBlock;
            const uint32_t x = sup[e < n ? e : n - 1];  // unconditional: the loads stay in flight together
            sp[j] = e < n ? x : KT_SUP_GONE;
        }
        // the tile's deletions into the LDS list first (entry, and whether it
        // lies in no triangle), then flagged from the list one per thread: a
        // deletion's reads (its words, then its endpoints' list starts) are
        // two round trips per tile instead of two per 64-entry chunk with one
#pragma unroll
        for (uint32_t j = 0; j < kMarkItems; ++j) {
            const uint32_t e = e0 + j * kBlock;
            const bool del = sp[j] < k - 2;
            const uint64_t m = ballot(del);
            if (!m) continue;
            uint32_t b = 0;
            if (lane == 0) b = atomicAdd(&s_cnt, (uint32_t)__popcll(m));  // LDS
            b = __shfl(b, 0);
            if (del) s_out[b + (uint32_t)__popcll(m & below)] = e | (sp[j] ? 0u : kMarkNoTri);
        }
        // every thread reads the count between two barriers: a wave that
        // went on to the next tile could otherwise add to it before a slower
        // wave has read it, and the two would disagree on flushing
        __syncthreads();
        const uint32_t c = s_cnt;
        __syncthreads();
        for (uint32_t slot = c_prev + threadIdx.x; slot < c; slot += kBlock) {
            const uint32_t w = s_out[slot];
            const uint32_t e = w & ~kMarkNoTri;
            const bool tri = !(w & kMarkNoTri);
            const uint32_t dv = dst[e], ei = eid[e], a = sa[e], cc = sb[e], u = src[e];
            const uint32_t vid = dv & KT_IDMASK;
            if (dv & (KT_DEAD | KT_DYING)) atomicAdd(bad, 1u);
            sup[e] = tri ? KT_SUP_GONE : KT_SUP_NONE;  // NONE: in no triangle, nothing to update
            dst[e] = vid | KT_DYING;
            tau[ei] = k - 1;
            snbr[a] = vid | KT_DYING;
            snbr[cc] = u | KT_DYING;
            if (ldeg) atomicSub(ldeg + u, 1u);
            if (listing) {
                // the deletion's shorter and longer list and its tasks of
                // 64 candidates (none for a deletion in no triangle)
                const uint32_t ub = mt.sptr[u], ue = mt.sptr[u + 1], vb = mt.sptr[vid], ve = mt.sptr[vid + 1];
                const bool us = ue - ub <= ve - vb;
                const uint4 md = us ? make_uint4(ub, ue, vb, ve - vb) : make_uint4(vb, ve, ub, ue - ub);
                mt.ws.mtmp[blockIdx.x * kMarkBuf + slot] = md;
                mt.ws.ttmp[blockIdx.x * kMarkBuf + slot] = tri ? (md.y - md.x + kWave - 1) / kWave : 0u;
            }
        }
        c_prev = c;
        if (c > kMarkBuf - kMarkTile) {
            // the flush reads every thread's staged descriptors and task
            // counts (global) and list words
            __syncthreads();
            if (listing) mark_flush_tasks(s_out, &s_cnt, s_res, s_w, dl, mt);
            else mark_flush(s_out, &s_cnt, &s_base, dl, ndel);
            c_prev = 0;
        }
    }
    __syncthreads();
    if (listing) mark_flush_tasks(s_out, &s_cnt, s_res, s_w, dl, mt);
    else mark_flush(s_out, &s_cnt, &s_base, dl, ndel);
}

// Minimum live support (the JUMP target); one atomic per workgroup.
constexpr uint32_t kMinThreads = 1024;
// Its end -- one atomic and one ticket per workgroup on single addresses --
// costs more than the scan: 128 workgroups of 1024 threads against 256 of
// 256 took R-MAT 20 / 22 from 91.9 / 366.0 to 91.0 / 364.2 ms; 1024 and 2048
// workgroups of 256, 94.4 / 371.5 and 99.2 / 378.8.
constexpr uint32_t kMinGrid = 128;
__global__ __launch_bounds__(kMinThreads) void k_min_list(uint32_t n, const uint32_t* __restrict__ sup,
                                                          DevCounters* ctr, uint32_t* host_min) {
    __shared__ uint32_t s_min[kMinThreads / kWave];
    __shared__ uint32_t s_last;
    uint32_t m = 0xFFFFFFFFu;
    // eight supports per thread in flight (indices past the end read entry
    // n - 1, which is then not counted); dead and dying entries hold
    // KT_SUP_GONE, above every live support
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t e0 = blockIdx.x * blockDim.x + threadIdx.x; e0 < n; e0 += 8 * stride) {
        uint32_t x[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            const uint32_t e = e0 + j * stride;
            x[j] = sup[e < n ? e : n - 1];
        }
#pragma unroll
        for (uint32_t j = 0; j < 8; ++j) {
            const uint32_t e = e0 + j * stride;
            if (e < n && x[j] < m && x[j] < KT_SUP_NONE) m = x[j];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(m, off);
        m = o < m ? o : m;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) s_min[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / kWave; ++w) m = s_min[w] < m ? s_min[w] : m;
        if (m != 0xFFFFFFFFu) atomicMin(&ctr->min_sup, m);
    }
    if (!host_min) return;  // uniform
    // the last workgroup to finish (a ticket taken after a device-scope
    // fence) publishes the minimum
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&ctr->mticket, 1u) == gridDim.x - 1 ? 1u : 0u;
    }
    __syncthreads();
    if (s_last && threadIdx.x == 0) {
        __threadfence();
        const uint32_t v = __hip_atomic_load(&ctr->min_sup, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ctr->mticket = 0;
        __hip_atomic_store(host_min, v == 0xFFFFFFFFu ? dev::kNoLiveSupport : v, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ void k_phase_start(unsigned int* from, DevCounters* ctr, int zero_ndel, int reset_min) {
    if (from) {  // a list's (count, task total) pair
        ctr->void_del |= from[0];
        from[0] = 0;
        from[1] = 0;
    }
    if (zero_ndel) {
        ctr->ndel = 0;
        ctr->ntask = 0;
    }
    if (reset_min) ctr->min_sup = 0xFFFFFFFFu;
}

// Task list of the pending deletions, one kernel.  Per deletion i (d = dl[i]
// = (u, v)): the shorter and the longer of the two adjacency lists and the
// number of 64-candidate tasks over the shorter one (at least 1: the list
// holds the edge itself; none for a deletion that lies in no triangle, whose
// support the mark left KT_SUP_NONE).  Deletions form tiles of kTileD; a workgroup scans
// its tile (lofs[i] = first task of deletion i counted from the tile's
// first) and stores the tile total; the last workgroup to finish (a ticket
// taken after a device-scope fence) scans the tile totals in place, so the
// first task of deletion i is tsum[i / kTileD] + lofs[i] and the task total
// is tsum[ntiles] = ctr->dtasks.
// A thread takes the deletions i = tile + r * kBlock + threadIdx.x (r <
// kPerD) and issues every load of a step for all of them before the first
// wait (indices past the end repeat the last deletion; only their stores are
// skipped), so a tile costs three dependent memory round trips; the counts
// go through LDS to be scanned in deletion order.  (With each thread taking
// kPerD consecutive deletions under a lane condition, the compiler waits for
// each one's loads before the next: 3 x kPerD round trips per tile.)
// lm.k: the pending list came from an update and is flagged here, its words
// loaded beside the task list's; lm.fin_dl: finalized by the whole grid first.
constexpr uint32_t kTileD = 2048;
constexpr uint32_t kPerD = kTileD / kBlock;  // deletions per thread of a tile

__global__ __launch_bounds__(kBlock) void k_dprep(const uint32_t* __restrict__ dl,
                                                  const uint32_t* __restrict__ nd_dev,
                                                  const uint32_t* __restrict__ src, const uint32_t* dst,
                                                  const uint32_t* __restrict__ sptr,
                                                  uint4* __restrict__ meta, uint32_t* __restrict__ lofs,
                                                  uint32_t* tsum, DevCounters* ctr, unsigned int* zero_cnt,
                                                  uint32_t* host_cnt, dev::ListMark lm,
                                                  const uint32_t* __restrict__ sup) {
    __shared__ uint32_t s_w[kBlock / kWave];
    __shared__ uint32_t s_c[kTileD];
    __shared__ uint32_t s_last;
    const uint32_t nd = *nd_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (zero_cnt) {  // a list's (count, task total) pair
            zero_cnt[0] = 0;
            zero_cnt[1] = 0;
        }
        if (host_cnt) __hip_atomic_store(host_cnt, nd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    if (lm.fin_dl) finalize_list(lm.fin_dl, lm.fin_nd, blockIdx.x, gridDim.x, src, lm.dst, lm.snbr, lm.sa, lm.sb);
    const uint32_t ntiles = (nd + kTileD - 1) / kTileD;
    const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (ntiles == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            tsum[0] = 0;
            ctr->dtasks = 0;
        }
        return;
    }
    if (blockIdx.x >= ntiles) return;  // whole workgroup
    const uint32_t last = nd - 1;
    uint32_t tile_total = 0;  // the last tile's total (thread kBlock - 1)
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t base = t * kTileD + threadIdx.x;
        uint32_t d[kPerD], u[kPerD], dv[kPerD];
#pragma unroll
        for (uint32_t r = 0; r < kPerD; ++r) {
            const uint32_t i = base + r * kBlock;
            d[r] = dl[i < last ? i : last];
        }
        bool none[kPerD];  // the deletion lies in no triangle
#pragma unroll
        for (uint32_t r = 0; r < kPerD; ++r) {
            u[r] = src[d[r]];
            dv[r] = dst[d[r]];
            none[r] = sup && !lm.k && sup[d[r]] == KT_SUP_NONE;
        }
        if (lm.k) {
            // a list the update made: its entries fell to k - 3 or below, so
            // at k = 3 they lie in no triangle
#pragma unroll
            for (uint32_t r = 0; r < kPerD; ++r) none[r] = sup && lm.k == 3;
            uint32_t e[kPerD], a[kPerD], b[kPerD];
#pragma unroll
            for (uint32_t r = 0; r < kPerD; ++r) {
                e[r] = lm.eid[d[r]];
                a[r] = lm.sa[d[r]];
                b[r] = lm.sb[d[r]];
            }
#pragma unroll
            for (uint32_t r = 0; r < kPerD; ++r) {
                if (base + r * kBlock < nd) {
                    const uint32_t v = dv[r] & KT_IDMASK;
                    lm.dst[d[r]] = v | KT_DYING;
                    lm.sup[d[r]] = KT_SUP_GONE;
                    lm.tau[e[r]] = lm.k - 1;
                    lm.snbr[a[r]] = v | KT_DYING;
                    lm.snbr[b[r]] = u[r] | KT_DYING;
                    if (lm.ldeg) atomicSub(lm.ldeg + u[r], 1u);
                }
            }
        }
        uint32_t ub[kPerD], ue[kPerD], vb[kPerD], ve[kPerD];
#pragma unroll
        for (uint32_t r = 0; r < kPerD; ++r) {
            const uint32_t v = dv[r] & KT_IDMASK;
            ub[r] = sptr[u[r]];
            ue[r] = sptr[u[r] + 1];
            vb[r] = sptr[v];
            ve[r] = sptr[v + 1];
        }
#pragma unroll
        for (uint32_t r = 0; r < kPerD; ++r) {
            const uint32_t i = base + r * kBlock;
            const bool us = ue[r] - ub[r] <= ve[r] - vb[r];
            const uint4 m = us ? make_uint4(ub[r], ue[r], vb[r], ve[r] - vb[r])
                               : make_uint4(vb[r], ve[r], ub[r], ue[r] - ub[r]);
            uint32_t c = 0;
            if (i < nd) {
                meta[i] = m;
                c = none[r] ? 0u : (m.y - m.x + kWave - 1) / kWave;
            }
            s_c[r * kBlock + threadIdx.x] = c;
        }
        __syncthreads();
        // the scan runs in deletion order: thread x holds deletions
        // tile + x * kPerD + [0, kPerD)
        const uint32_t i0 = t * kTileD + threadIdx.x * kPerD;
        uint32_t cnt[kPerD];
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t r = 0; r < kPerD; ++r) {
            cnt[r] = s_c[threadIdx.x * kPerD + r];
            sum += cnt[r];
        }
        uint32_t inc = sum;  // inclusive scan over the workgroup
        for (int off = 1; off < (int)kWave; off <<= 1) {
            const uint32_t y = __shfl_up(inc, off);
            if ((int)lane >= off) inc += y;
        }
        if (lane == kWave - 1) s_w[wid] = inc;
        __syncthreads();
        uint32_t run = inc - sum;
        for (uint32_t w = 0; w < wid; ++w) run += s_w[w];
#pragma unroll
        for (uint32_t r = 0; r < kPerD; ++r) {
            if (i0 + r < nd) lofs[i0 + r] = run;
            run += cnt[r];
        }
        if (threadIdx.x == kBlock - 1) {
            tsum[t] = run;  // the tile total
            tile_total = run;
        }
        __syncthreads();
    }
    if (ntiles == 1) {  // one workgroup did everything: no ticket
        if (threadIdx.x == kBlock - 1) {
            tsum[0] = 0;
            tsum[1] = tile_total;
            ctr->dtasks = tile_total;
        }
        return;
    }
    // the last workgroup to finish scans the tile totals
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t parts = ntiles < gridDim.x ? ntiles : gridDim.x;
        s_last = atomicAdd(&ctr->dticket, 1u) == parts - 1 ? 1u : 0u;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    const uint32_t per = (ntiles + kBlock - 1) / kBlock;
    const uint32_t b = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t i = b; i < b + per && i < ntiles; ++i) sum += tsum[i];
    uint32_t inc = sum;
    for (int off = 1; off < (int)kWave; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off);
        if ((int)lane >= off) inc += y;
    }
    if (lane == kWave - 1) s_w[wid] = inc;
    __syncthreads();
    uint32_t run = inc - sum;
    for (uint32_t w = 0; w < wid; ++w) run += s_w[w];
    for (uint32_t i = b; i < b + per && i < ntiles; ++i) {
        const uint32_t c = tsum[i];
        tsum[i] = run;
        run += c;
    }
    if (threadIdx.x == kBlock - 1) {
        tsum[ntiles] = run;
        ctr->dtasks = run;
        ctr->dticket = 0;
    }
}

// First task of deletion i: from the task-list kernel's per-tile offsets, or
// (tsum null) as a listing mark stored it.
__device__ __forceinline__ uint32_t first_task(const uint32_t* lofs, const uint32_t* tsum, uint32_t i) {
    return tsum ? tsum[i / kTileD] + lofs[i] : lofs[i];
}

// The update proper.  Task t = (deletion i, chunk c) holds the candidates
// short[64c, 64c + 64) of deletion i; wave w takes the task range
// [T w / W, T (w + 1) / W), finds its first deletion by a binary search over
// the task offsets and walks forward, so the tasks of one deletion follow
// each other.  A task whose
// candidates are all dead is skipped.  Every
// candidate is binary searched in LDS, never by dependent probes of HBM:
//   longer list of at most kStage entries: staged once per deletion in the
//     wave's LDS slice (one coalesced read);
//   longer list of at most 64 kSample entries: the first id of every
//     64-entry block is staged once per deletion (one strided read); two
//     searches there bound the window of the list that the task's candidate
//     range [xmin, xmax] can hit, and the window (at most kStage entries, or
//     searched in place beyond) is staged for the task;
//   longer still (hub lists): the window is bounded by two interleaved
//     wave-uniform searches of the list itself.
// Adjacency slots carry KT_DYING for the edges the current pass deleted, so
// the ownership test reads no flags of the oriented layout: a triangle whose
// other two edges survive (the common case) costs two slot->entry reads and
// two atomic decrements.
// kCross: the decrements return the old support and the entries taken from
// thr = k - 2 to k - 3 are listed in `next`, first in the wave's LDS list
// (written out with one reservation when it fills), the rest with one
// reservation per workgroup at the end; a task's returned supports are
// examined during the next task, after its loads are issued.
constexpr uint32_t kDBlock = 256;  // k_delta workgroup (512 and 1024 measured the same)
constexpr uint32_t kCrossCap = 256;  // per-wave LDS list of the next deletions
constexpr uint32_t kStage = 512;
constexpr uint32_t kDeltaGrid = 2048;  // most k_delta workgroups per launch
constexpr size_t kStatBytes = (size_t)kDeltaGrid * 4 * 8;
static_assert(8192 / (kDBlock / kWave) <= kDeltaGrid, "k_delta grids of 8192 waves need a slot per workgroup");
constexpr uint32_t kSample = 512;
constexpr uint32_t kGather = 8;  // blocks gathered per round trip (two 16-byte loads per lane)
static_assert(kGather * kWave <= kStage, "gathered blocks fit the slice");

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Staging of up to N entries src[q * stride] (q < n) into a wave's LDS slice.
// Every lane issues its N / 64 loads before the first one is waited for, so a
// staging costs one memory round trip, not one per 64 entries.  The loads are
// unconditional (indices past the end repeat the last entry): a load under a
// lane condition makes the compiler wait for it inside the condition.
template <uint32_t N>
struct Staged {
    uint32_t r[N / kWave];
    __device__ __forceinline__ void load(const uint32_t* src, uint32_t n, uint32_t stride,
                                         uint32_t lane) {
        if (n == 0) return;
        const uint32_t last = n - 1;
#pragma unroll
        for (uint32_t k = 0; k < N / kWave; ++k) {
            const uint32_t q = lane + k * kWave;
            r[k] = src[(size_t)(q < last ? q : last) * stride];
        }
    }
    __device__ __forceinline__ void store(uint32_t* buf, uint32_t n, uint32_t mask,
                                          uint32_t lane) const {
        wave_lds_sync();
#pragma unroll
        for (uint32_t k = 0; k < N / kWave; ++k) {
            const uint32_t q = lane + k * kWave;
            if (q < n) buf[q] = r[k] & mask;
        }
        wave_lds_sync();
    }
};

// First index in [lo, lo + len) whose id (flags masked) is >= x; the trip
// count depends on len only.
__device__ __forceinline__ uint32_t lower_bound_u(const uint32_t* a, uint32_t lo, uint32_t len,
                                                  uint32_t x) {
    if (len == 0) return lo;
    while (len > 1) {
        const uint32_t half = len >> 1;
        if ((a[lo + half] & KT_IDMASK) < x) lo += half;
        len -= half;
    }
    return lo + ((a[lo] & KT_IDMASK) < x ? 1u : 0u);
}

// Lower bounds of x0 and x1 in [lo, lo + len), interleaved (same trip count).
__device__ __forceinline__ void lower_bounds_2(const uint32_t* a, uint32_t lo, uint32_t len,
                                               uint32_t x0, uint32_t x1, uint32_t& r0,
                                               uint32_t& r1) {
    uint32_t b0 = lo, b1 = lo;
    while (len > 1) {
        const uint32_t half = len >> 1;
        const uint32_t v0 = a[b0 + half] & KT_IDMASK, v1 = a[b1 + half] & KT_IDMASK;
        if (v0 < x0) b0 += half;
        if (v1 < x1) b1 += half;
        len -= half;
    }
    r0 = b0 + ((a[b0] & KT_IDMASK) < x0 ? 1u : 0u);
    r1 = b1 + ((a[b1] & KT_IDMASK) < x1 ? 1u : 0u);
}

// Number of samples < x in s[0, n) (n >= 1), uniform trip count.
__device__ __forceinline__ uint32_t count_less_lds(const uint32_t* s, uint32_t n, uint32_t x) {
    uint32_t b = 0;
    while (n > 1) {
        const uint32_t half = n >> 1;
        if (s[b + half] < x) b += half;
        n -= half;
    }
    return b + (s[b] < x ? 1u : 0u);
}

// Numbers of samples below x0 and below x1 in s[0, n) (n <= kSample, x0 and
// x1 the same in every lane), counted lane-parallel: lane l compares
// s[4l, 4l + 4) and s[256 + 4l, 256 + 4l + 4) (two 16-byte reads, lanes on
// consecutive 16-byte chunks, so no bank conflicts) and the ballots are
// summed -- one LDS round trip instead of two dependent binary searches.
static_assert(kSample == 8 * kWave, "two 16-byte reads per lane cover the samples");
__device__ __forceinline__ void count_less2_lds(const uint32_t* s, uint32_t n, uint32_t x0, uint32_t x1,
                                                uint32_t lane, uint32_t& c0, uint32_t& c1) {
    const uint32_t b = 4u * lane;
    const uint4 lo = *(const uint4*)(s + b);
    const uint4 hi = *(const uint4*)(s + 4 * kWave + b);
    const uint32_t v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    c0 = c1 = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        const bool in = (j < 4 ? b + j : 4 * kWave + b + (j - 4)) < n;
        c0 += (uint32_t)__popcll(ballot(in && v[j] < x0));
        c1 += (uint32_t)__popcll(ballot(in && v[j] < x1));
    }
}

// Searches xv in buf[0, n) (staged list entries) and returns the slot offset
// and the entry (KT_DEAD when past the end).
__device__ __forceinline__ uint32_t search_staged(const uint32_t* buf, uint32_t n, uint32_t xv,
                                                  uint32_t& y) {
    uint32_t b = 0, m = n;
    while (m > 1) {
        const uint32_t half = m >> 1;
        if ((buf[b + half] & KT_IDMASK) < xv) b += half;
        m -= half;
    }
    const uint32_t q = b + ((buf[b] & KT_IDMASK) < xv ? 1u : 0u);
    y = q < n ? buf[q] : KT_DEAD;
    return q;
}

template <bool kCross>
__global__ __launch_bounds__(kDBlock, kCross ? 7 : 8) void k_delta(const uint32_t* __restrict__ dl,
                                                  const uint32_t* __restrict__ nd_dev,
                                                  const uint4* __restrict__ meta,
                                                  const uint32_t* __restrict__ lofs,
                                                  const uint32_t* tsum,
                                                  const uint32_t* __restrict__ snbr,
                                                  const uint32_t* __restrict__ seid,
                                                  uint32_t* sup, DevCounters* ctr,
                                                  unsigned long long* __restrict__ bstat, uint32_t thr,
                                                  uint32_t* __restrict__ next, unsigned int* next_cnt,
                                                  const uint32_t* __restrict__ gsmp, dev::DirectTasks dt) {
    __shared__ __attribute__((aligned(16))) uint32_t s_buf[(kDBlock / kWave) * (kStage + kSample + kWave)];
    __shared__ uint32_t s_cross[kCross ? (kDBlock / kWave) * kCrossCap : 1];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    uint32_t* buf = s_buf + (threadIdx.x / kWave) * (kStage + kSample + kWave);
    uint32_t* smp = buf + kStage;
    uint32_t* gl = smp + kSample;  // distinct blocks of a gathered task
    uint32_t* cl = s_cross + (kCross ? (threadIdx.x / kWave) * kCrossCap : 0u);
    uint32_t nc = 0;  // entries in cl (wave-uniform)
    // the last task's decremented entries and their old supports (0: none;
    // thr >= 2)
    uint32_t pe1 = 0, pe2 = 0, po1 = 0, po2 = 0;
    auto cross = [&]() {
        const uint64_t below = (1ull << lane) - 1ull;
        const bool c1 = po1 == thr, c2 = po2 == thr;
        po1 = po2 = 0;
        const uint64_t m1 = ballot(c1), m2 = ballot(c2);
        if (!(m1 | m2)) return;
        if (c1) cl[nc + (uint32_t)__popcll(m1 & below)] = pe1;
        nc += (uint32_t)__popcll(m1);
        if (c2) cl[nc + (uint32_t)__popcll(m2 & below)] = pe2;
        nc += (uint32_t)__popcll(m2);
        if (nc > kCrossCap - 2 * kWave) {  // room for one more task's entries
            uint32_t at = 0;
            if (lane == 0) at = atomicAdd(next_cnt, nc);
            at = __shfl(at, 0);
            wave_lds_sync();
            for (uint32_t q = lane; q < nc; q += kWave) next[at + q] = cl[q];
            wave_lds_sync();
            nc = 0;
        }
    };
    const uint32_t nd = *nd_dev;
    const uint32_t T = dt.ntask ? *dt.ntask : ctr->dtasks;
    if (dt.ntask) tsum = nullptr;  // the list's first tasks are absolute
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // the task-list kernel's other duties
        if (dt.zero_pair) *dt.zero_pair = 0;
        if (dt.host_cnt) __hip_atomic_store(dt.host_cnt, nd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    const uint32_t W = nwaves(), w = gwave();
    const uint32_t t0 = (uint32_t)((uint64_t)T * w / W);
    const uint32_t t1 = nd ? (uint32_t)((uint64_t)T * (w + 1) / W) : t0;  // no early return:
                                                                          // the block reduces below
    // last deletion whose first task is <= t0: a 64-ary search, the lanes
    // probing 64 evenly spaced deletions per step (three dependent loads for
    // 2^18 deletions instead of eighteen); first_task(0) = 0 <= t0
    uint32_t i = 0;
    if (t0 < t1) {
        uint32_t len = nd;
        while (len > 1) {
            const uint32_t step = (len + kWave - 1) / kWave, end = i + len;
            const uint32_t idx = i + lane * step;
            const bool ok = idx < end && first_task(lofs, tsum, idx) <= t0;
            const uint64_t m = ballot(ok);  // lane 0 always holds
            i += (63u - (uint32_t)__builtin_clzll(m)) * step;
            len = min(step, end - i);
        }
    }
    uint32_t full = 0xFFFFFFFFu;     // deletion whose whole longer list is in buf
    uint32_t sampled = 0xFFFFFFFFu;  // deletion whose block samples are in smp
    // a wave's statistics of one launch stay far below 2^32
    uint32_t acc_elems = 0, acc_tri = 0, acc_win = 0, acc_live = 0;
    // deletion i: its descriptor, its first task and the next deletion's (a
    // deletion without tasks shares its first task with the next one, so the
    // search above lands on the task's owner and the walk below steps over
    // it); the next task's candidates are loaded before this
    // task's search when they belong to the same deletion, so their round trip
    // overlaps this task's slot->entry reads
    uint4 m = make_uint4(0, 0, 0, 0);
    uint32_t dcur = 0;  // deletion i's entry (the ownership test's)
    uint32_t fi = 0, fnext = 0xFFFFFFFFu;
    uint32_t xpre = KT_DEAD;
    bool pre = false;  // xpre holds task t's candidates (wave-uniform)
    if (t0 < t1) {
        fi = first_task(lofs, tsum, i);
        fnext = i + 1 < nd ? first_task(lofs, tsum, i + 1) : 0xFFFFFFFFu;
        m = meta[i];
        dcur = dl[i];
    }
    for (uint32_t t = t0; t < t1;) {
        if (t >= fnext) {
            while (t >= fnext) {  // fnext < ~0 means i + 1 < nd
                ++i;
                fi = fnext;
                fnext = i + 1 < nd ? first_task(lofs, tsum, i + 1) : 0xFFFFFFFFu;
            }
            m = meta[i];
            dcur = dl[i];
            pre = false;
        }
        const uint32_t tc = t++;  // this task; the body may continue below
        const uint32_t j0 = m.x + (tc - fi) * kWave;
        const uint32_t j = j0 + lane;
        acc_elems += min(kWave, m.y - j0);
        // a longer list that fits the slice is staged with the candidates'
        // load in flight
        // a longer list that fits the slice is staged, or the first ids of its
        // 64-entry blocks are, with the candidates' load in flight
        // samples of the longer list: with gsmp, its first id and the first
        // ids of the 64-slot blocks that start inside it (gsmp[b] = id of
        // slot 64 b, contiguous); else the first id of every 64-entry block
        // of the list (one line per sample)
        const uint32_t b0 = m.z >> 6;
        const uint32_t nb = gsmp ? 1u + ((m.z + m.w - 1) >> 6) - b0 : (m.w + kWave - 1) / kWave;
        const bool fresh = m.w <= kStage && full != i;
        const bool fresh_smp = m.w > kStage && nb <= kSample && sampled != i;
        // One staging load site for the three cases (loads from two sites
        // would meet in register copies, each waiting for its load), and the
        // gsmp case's first id loaded beside it, stored as sample 0 (merged
        // into the staged registers here, its select would wait at once).
        Staged<kStage> st;
        uint32_t f0 = 0;
        if (fresh || fresh_smp) {
            const uint32_t* from = fresh ? snbr + m.z : (gsmp ? gsmp + b0 : snbr + m.z);
            st.load(from, fresh ? m.w : nb, fresh || gsmp ? 1u : kWave, lane);
            f0 = snbr[m.z];
        }
        // This task's candidates (prefetched by the task before when it has
        // the same deletion).  Loads are unconditional and clamped to the
        // list, the lanes past it masked here: a load under a lane condition,
        // or a select on its value, is waited for at once.
        const uint32_t x = pre ? xpre : snbr[min(j, m.y - 1)];
        if (kCross) cross();
        const bool cand = j < m.y && !(x & KT_DEAD);
        const uint64_t cm = ballot(cand);
        acc_live += (uint32_t)__popcll(cm);
        // the next task's candidates, issued once this task's are in use (a
        // load issued before that use would be waited for with them)
        pre = t < t1 && t < fnext;
        xpre = snbr[min(j + kWave, m.y - 1)];
        if (fresh) {
            st.store(buf, m.w, 0xFFFFFFFFu, lane);
            full = i;
            acc_win += m.w;
        } else if (fresh_smp) {
            st.store(smp, nb, KT_IDMASK, lane);
            if (gsmp) {  // block b0 may start before the list: sample 0 is its first id
                if (lane == 0) smp[0] = f0 & KT_IDMASK;
                wave_lds_sync();
            }
            sampled = i;
            acc_win += nb;
        }
        if (!cm) continue;
        const uint32_t xv = x & KT_IDMASK;
        uint32_t p, y;
        if (m.w <= kStage) {
            p = m.z + search_staged(buf, m.w, xv, y);
        } else {
            // candidates ascend with the lane
            const uint32_t xmin = __builtin_amdgcn_readlane(xv, (uint32_t)__ffsll((long long)cm) - 1);
            const uint32_t xmax = __builtin_amdgcn_readlane(xv, 63u - (uint32_t)__builtin_clzll(cm));
            uint32_t lo, hi;
            if (nb <= kSample) {
                // blocks whose first id is below xmin / xmax + 1: the bounds lie
                // in the last such block (or at the list start)
                // sample q lies at slot m.z (q = 0) or at the start of block
                // b0 + q (gsmp), else at slot m.z + 64 q
                uint32_t ca, cb;
                count_less2_lds(smp, nb, xmin, xmax + 1, lane, ca, cb);
                if (gsmp) {
                    lo = ca > 1 ? (b0 + ca - 1) << 6 : m.z;
                    hi = cb == 0 ? m.z : (cb < nb ? (b0 + cb) << 6 : m.z + m.w);
                } else {
                    lo = m.z + (ca ? (ca - 1) * kWave : 0u);
                    hi = m.z + min(m.w, cb * kWave);
                }
            } else {
                lower_bounds_2(snbr, m.z, m.w, xmin, xmax + 1, lo, hi);
            }
            const uint32_t wl = hi - lo;
            if (wl == 0) continue;
            if (wl <= kStage) {
                full = 0xFFFFFFFFu;  // buf now holds a window
                acc_win += wl;
                Staged<kStage> sw;
                sw.load(snbr + lo, wl, 1u, lane);
                sw.store(buf, wl, 0xFFFFFFFFu, lane);
                p = lo + search_staged(buf, wl, xv, y);
            } else if (gsmp && nb <= kSample) {
                // a wide window: each candidate's 64-slot block (the last
                // sample at or below it) is gathered instead, eight distinct
                // blocks of the task per round trip, and searched in the slice
                full = 0xFFFFFFFFu;  // buf now holds gathered blocks
                // slot of sample q (q = nb: the list end)
                auto spos = [&](uint32_t q) -> uint32_t {
                    return q >= nb ? m.z + m.w : (q ? (b0 + q) << 6 : m.z);
                };
                const int32_t blk = cand ? (int32_t)count_less_lds(smp, nb, xv + 1) - 1 : -1;
                // blocks ascend over the candidate lanes: a block's first lane
                // differs from the running maximum of the lanes before it
                int32_t run = blk;
                for (uint32_t off = 1; off < kWave; off <<= 1) {
                    const int32_t o = __shfl_up(run, off);
                    if (lane >= off) run = o > run ? o : run;
                }
                int32_t before = __shfl_up(run, 1);
                if (lane == 0) before = -1;
                const bool first = blk >= 0 && blk != before;
                const uint64_t firsts = ballot(first);
                const uint32_t grp = (uint32_t)__popcll(firsts & ((2ull << lane) - 1ull)) - 1u;
                const uint32_t D = (uint32_t)__popcll(firsts);
                wave_lds_sync();
                if (first) gl[grp] = (uint32_t)blk;  // the task's distinct blocks, ascending
                wave_lds_sync();
                p = 0;
                y = KT_DEAD;
                for (uint32_t g0 = 0; g0 < D; g0 += kGather) {
                    // blocks g0 .. g0 + 7 from their 64-aligned starts (the
                    // lists are padded to the last): lane l loads slots
                    // 4 (l & 15) .. + 3 of block g0 + (l >> 4) and of block
                    // g0 + 4 + (l >> 4), so each 16-byte store of the wave
                    // covers 1 KB of the slice without bank conflicts
                    const uint32_t q = lane >> 4, o = 4u * (lane & 15u);
                    const uint32_t b1 = g0 + q < D ? gl[g0 + q] : gl[g0];
                    const uint32_t b2 = g0 + 4 + q < D ? gl[g0 + 4 + q] : gl[g0];
                    const uint4 lo4 = *(const uint4*)(snbr + ((b0 + b1) << 6) + o);
                    const uint4 hi4 = *(const uint4*)(snbr + ((b0 + b2) << 6) + o);
                    acc_win += min(D - g0, kGather) * kWave;
                    wave_lds_sync();
                    *(uint4*)(buf + 4u * lane) = lo4;
                    *(uint4*)(buf + 4u * kWave + 4u * lane) = hi4;
                    wave_lds_sync();
                    if (blk >= 0 && grp >= g0 && grp < g0 + kGather) {
                        // the block's slots of the list: [s0, s1) of its 64
                        const uint32_t base = ((b0 + (uint32_t)blk) << 6);
                        const uint32_t s0 = spos((uint32_t)blk) - base, s1 = spos((uint32_t)blk + 1) - base;
                        uint32_t yy;
                        const uint32_t r = search_staged(buf + (grp - g0) * kWave + s0, s1 - s0, xv, yy);
                        p = base + s0 + r;
                        y = yy;
                    }
                }
            } else {
                acc_win += (uint32_t)__popcll(cm) * (32u - __builtin_clz(wl));
                p = lower_bound_u(snbr, lo, wl, xv);
                y = p < hi ? snbr[p] : KT_DEAD;
            }
        }
        const bool hit = cand && (y & (KT_DEAD | KT_IDMASK)) == xv;
        bool tri = false;
        if (hit) {
            const uint32_t e1 = seid[j], e2 = seid[p];
            const bool y1 = (x & KT_DYING) != 0, y2 = (y & KT_DYING) != 0;
            // the deleted edge of lowest index owns the triangle; both
            // entries are compared in every hit lane, so both loads are
            // consumed before the decrements (a load left unconsumed where a
            // decrement is skipped is waited for later, together with the
            // decrements issued after it)
            const bool own = !((y1 & (e1 < dcur)) | (y2 & (e2 < dcur)));
            if (own) {
                if (kCross) {
                    pe1 = e1;
                    pe2 = e2;
                    if (!y1) po1 = atomicSub(sup + e1, 1u);
                    if (!y2) po2 = atomicSub(sup + e2, 1u);
                } else {
                    if (!y1) atomicSub(sup + e1, 1u);
                    if (!y2) atomicSub(sup + e2, 1u);
                }
                tri = true;
            }
        }
        acc_tri += (uint32_t)__popcll(ballot(tri));
    }
    if (kCross) {  // the last task's entries, then the waves' remaining ones:
                   // one reservation per workgroup
        cross();
        __shared__ uint32_t s_nc[kDBlock / kWave];
        __shared__ uint32_t s_at;
        const uint32_t wid = threadIdx.x / kWave;
        if (lane == 0) s_nc[wid] = nc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (uint32_t q = 0; q < kDBlock / kWave; ++q) tot += s_nc[q];
            s_at = tot ? atomicAdd(next_cnt, tot) : 0u;
        }
        __syncthreads();
        uint32_t at = s_at;
        for (uint32_t q = 0; q < wid; ++q) at += s_nc[q];
        for (uint32_t q = lane; q < nc; q += kWave) next[at + q] = cl[q];
    }
    // statistics: one slot per workgroup, updated in place (a launch has at
    // most kDeltaGrid workgroups and launches on one stream do not overlap);
    // counters every wave bumps with an atomic serialize on their one address
    // and hold back the whole launch
    __shared__ unsigned long long s_stat[kDBlock / kWave][4];
    if (lane == 0) {
        unsigned long long* q = s_stat[threadIdx.x / kWave];
        q[0] = acc_elems;
        q[1] = acc_win;
        q[2] = acc_live;
        q[3] = acc_tri;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long v = 0;
        for (uint32_t k = 0; k < kDBlock / kWave; ++k) v += s_stat[k][threadIdx.x];
        if (v) bstat[blockIdx.x * 4 + threadIdx.x] += v;
    }
}

// Adds the per-workgroup statistics to the counters and clears the slots.
__global__ __launch_bounds__(256) void k_delta_fold(unsigned long long* __restrict__ bstat,
                                                    DevCounters* ctr) {
    __shared__ unsigned long long s[256];
    const uint32_t c = threadIdx.x & 3;
    unsigned long long v = 0;
    for (uint32_t b = threadIdx.x >> 2; b < kDeltaGrid; b += 64) {
        v += bstat[b * 4 + c];
        bstat[b * 4 + c] = 0;
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 128; off >= 4; off >>= 1) {  // offsets keep the stat index t & 3
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctr->delta_elems += s[0];
        ctr->delta_win += s[1];
        ctr->delta_live += s[2];
        ctr->delta_tris += s[3];
    }
}

__global__ void k_finalize(const uint32_t* __restrict__ dl, const uint32_t* __restrict__ nd_dev,
                           uint32_t* __restrict__ dst, uint32_t* __restrict__ snbr,
                           const uint32_t* __restrict__ sa, const uint32_t* __restrict__ sb) {
    const uint32_t nd = *nd_dev;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nd; t += gridDim.x * blockDim.x)
        finalize_one(dl[t], dst, snbr, sa, sb);
}

uint32_t blocks_for(uint64_t items, uint32_t per_block) {
    uint64_t b = (items + per_block - 1) / per_block;
    if (b > 8192) b = 8192;
    return b ? (uint32_t)b : 1u;
}

// ---------------------------------------------------------------------------
// Compaction of the adjacency lists (stable, tiles of kTileC slots, one
// 256-thread workgroup per tile: count -> one-workgroup scan of the tile
// counts -> scatter with in-tile ballot prefixes).  Slots whose edge is dead
// leave the lists; the slot -> entry map is rewritten through the oriented
// layout's own compaction map (the entries move too); the list starts are
// rebuilt from the slot prefix and every live entry's two slots (sa, sb) are
// found again from the kept slots.
constexpr uint32_t kTileC = 2048;

__global__ __launch_bounds__(kBlock) void k_cmp_count(uint32_t n, const uint32_t* __restrict__ snbr,
                                                      uint32_t* __restrict__ tsum) {
    __shared__ uint32_t s_w[kBlock / kWave];
    const uint32_t t = blockIdx.x;
    uint32_t c = 0;
    for (uint32_t i = t * kTileC + threadIdx.x; i < (t + 1) * kTileC && i < n; i += kBlock)
        c += (snbr[i] & KT_DEAD) ? 0u : 1u;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (uint32_t w = 0; w < kBlock / kWave; ++w) sum += s_w[w];
        tsum[t] = sum;
    }
}

// Exclusive scan of nt tile counts in place by one workgroup; total at tsum[nt].
__global__ __launch_bounds__(1024) void k_cmp_scan(uint32_t* tsum, uint32_t nt) {
    __shared__ uint32_t s_part[1024];
    const uint32_t per = (nt + 1023) / 1024;
    const uint32_t b = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t i = b; i < b + per && i < nt; ++i) sum += tsum[i];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t x = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0;
    for (uint32_t i = b; i < b + per && i < nt; ++i) {
        const uint32_t c = tsum[i];
        tsum[i] = run;
        run += c;
    }
    if (threadIdx.x == 1023) tsum[nt] = s_part[1023];
}

// Scatter of the kept slots: snbr2 / seid2 receive the slot's neighbour and
// its entry's new position remap[seid], and xs[i] = kept slots before slot i
// (for every slot; xs[n] is set by the last tile).
__global__ __launch_bounds__(kBlock) void k_cmp_scatter(uint32_t n, const uint32_t* __restrict__ snbr,
                                                        const uint32_t* __restrict__ seid,
                                                        const uint32_t* __restrict__ remap,
                                                        const uint32_t* __restrict__ tsum,
                                                        uint32_t* __restrict__ snbr2,
                                                        uint32_t* __restrict__ seid2,
                                                        uint32_t* __restrict__ xs) {
    __shared__ uint32_t s_w[kBlock / kWave];
    const uint32_t t = blockIdx.x;
    const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    uint32_t run = tsum[t];
    for (uint32_t r = 0; r < kTileC / kBlock; ++r) {
        const uint32_t i = t * kTileC + r * kBlock + threadIdx.x;
        const uint32_t y = i < n ? snbr[i] : KT_DEAD;
        const bool keep = !(y & KT_DEAD);
        const uint64_t m = ballot(keep);
        if (lane == 0) s_w[wid] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = run, round = 0;
        for (uint32_t w = 0; w < kBlock / kWave; ++w) {
            if (w < wid) before += s_w[w];
            round += s_w[w];
        }
        const uint32_t pos = before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (i < n) xs[i] = pos;
        if (keep) {
            snbr2[pos] = y;
            seid2[pos] = remap[seid[i]];
        }
        run += round;
        __syncthreads();
    }
    if (t == gridDim.x - 1 && threadIdx.x == 0) xs[n] = tsum[gridDim.x];
}

__global__ void k_cmp_sptr(const uint32_t* __restrict__ sptr, const uint32_t* __restrict__ xs,
                           uint32_t* __restrict__ sptr2, uint32_t n) {
    for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x <= n; x += gridDim.x * blockDim.x)
        sptr2[x] = xs[sptr[x]];
}

// A kept slot j holds neighbour y of edge e = (u, v): it lies in u's list
// when y == v.
__global__ void k_cmp_slots(const uint32_t* __restrict__ snbr2, const uint32_t* __restrict__ seid2,
                            uint32_t S2, const uint32_t* __restrict__ dst, uint32_t* __restrict__ sa,
                            uint32_t* __restrict__ sb) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < S2; j += gridDim.x * blockDim.x) {
        const uint32_t e = seid2[j];
        if ((snbr2[j] & KT_IDMASK) == (dst[e] & KT_IDMASK)) sa[e] = j;
        else sb[e] = j;
    }
}

// gsmp[b] = id of slot 64 b (flags masked) for the S slots of the lists.
__global__ void k_block_samples(const uint32_t* __restrict__ snbr, uint32_t S, uint32_t* __restrict__ gsmp) {
    const uint32_t nb = (S + kWave - 1) / kWave;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x)
        gsmp[b] = snbr[(size_t)b * kWave] & KT_IDMASK;
}

__global__ void k_row_lengths(const uint32_t* __restrict__ ptr, uint32_t nv, uint32_t* __restrict__ out) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += gridDim.x * blockDim.x)
        out[v] = ptr[v + 1] - ptr[v];
}

}  // namespace

namespace dev {

hipError_t mark_dying(uint32_t* dst, const uint32_t* eid, uint32_t* sup, uint32_t* tau,
                      uint32_t* dl, uint32_t n, uint32_t k, uint32_t* snbr, const uint32_t* sa,
                      const uint32_t* sb, unsigned int* ndel, const uint32_t* fin_dl,
                      const uint32_t* fin_nd, uint32_t fin_max, const uint32_t* src, uint32_t* ldeg,
                      unsigned int* bad, hipStream_t s, const MarkTasks* mt) {
    if (!fin_max) fin_dl = nullptr;
    if (mt && (!mt->pair || !mt->sptr || !mt->ws.meta || !mt->ws.ntask || !mt->ws.mtmp || !mt->ws.ttmp ||
               (uint64_t)n > mt->ws.cap))
        return hipErrorInvalidValue;
    if (!n) {
        if (fin_dl) return finalize_deleted(fin_dl, fin_nd, fin_max, dst, snbr, sa, sb, s);
        return hipSuccess;
    }
    const uint32_t nt = (n + kMarkTile - 1) / kMarkTile;
    const uint32_t mb = nt < kMarkGrid ? nt : kMarkGrid;
    // finalizing workgroups: fin_max is a bound (the grid strides over the true count)
    const uint32_t fw = (uint32_t)std::min<uint64_t>(((uint64_t)fin_max + kBlock - 1) / kBlock, kFinGrid);
    hipLaunchKernelGGL(k_mark_dying, dim3(mb + (fin_dl ? fw : 0u)), dim3(kBlock), 0, s, dst, eid, sup, tau, dl,
                       n, k, snbr, sa, sb, ndel, fin_dl, fin_nd, src, ldeg, bad, mb, mt ? *mt : MarkTasks{});
    return hipGetLastError();
}

hipError_t phase_start(unsigned int* from, DevCounters* ctr, bool zero_ndel, bool reset_min,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_phase_start, dim3(1), dim3(1), 0, s, from, ctr, zero_ndel ? 1 : 0, reset_min ? 1 : 0);
    return hipGetLastError();
}

hipError_t min_live_delta(uint32_t n, const uint32_t* sup, DevCounters* ctr, hipStream_t s,
                          uint32_t* host_min) {
    if (!n) return host_min ? hipErrorInvalidValue : hipSuccess;
    const uint32_t b = blocks_for(n, kMinThreads);
    hipLaunchKernelGGL(k_min_list, dim3(b < kMinGrid ? b : kMinGrid), dim3(kMinThreads), 0, s, n, sup, ctr,
                       host_min);
    return hipGetLastError();
}

// a listing mark's staging: one buffer of kMarkBuf slots per mark workgroup
// (a mark of at most cap entries launches at most this many)
static size_t mark_stage_slots(uint32_t cap) {
    const size_t tiles = ((size_t)(cap ? cap : 1) + kMarkTile - 1) / kMarkTile;
    return (tiles < kMarkGrid ? tiles : kMarkGrid) * kMarkBuf;
}

size_t delta_scratch_bytes(uint32_t cap) {
    const size_t c = cap ? cap : 1;
    const size_t tail = (20 * c + 4 * (c / kTileD + 2) + 15) & ~(size_t)15;
    return kStatBytes + tail + 20 * mark_stage_slots(cap) + 256;
}

void delta_scratch_carve(void* base, uint32_t cap, DeltaScratch& s) {
    const size_t c = cap ? cap : 1;
    s.bstat = (unsigned long long*)base;
    char* p = (char*)base + kStatBytes;
    s.meta = (uint4*)p;
    s.ntask = (uint32_t*)(p + 16 * c);
    s.tsum = (uint32_t*)(p + 20 * c);
    char* q = p + ((20 * c + 4 * (c / kTileD + 2) + 15) & ~(size_t)15);
    s.mtmp = (uint4*)q;
    s.ttmp = (uint32_t*)(q + 16 * mark_stage_slots(cap));
    s.cap = cap;
}

hipError_t delta_prepare(const uint32_t* dl, const uint32_t* nd_dev, uint32_t nd_max,
                         const uint32_t* src, const uint32_t* dst, const uint32_t* sptr,
                         const DeltaScratch& ws, DevCounters* ctr, hipStream_t s,
                         unsigned int* zero_cnt, uint32_t* host_cnt, const ListMark* lm, const uint32_t* sup) {
    if (!nd_max && !zero_cnt && !host_cnt && !lm) return hipSuccess;
    ListMark m{};
    if (lm) {
        m = *lm;
        if ((m.k && (m.k < 3 || !m.dst || !m.eid || !m.sup || !m.tau || !m.snbr || !m.sa || !m.sb)) ||
            (m.fin_dl && (!m.dst || !m.fin_nd || !m.snbr || !m.sa || !m.sb)))
            return hipErrorInvalidValue;
    }
    const uint32_t tiles = (nd_max + kTileD - 1) / kTileD;
    const uint32_t tb = tiles < 256 ? (tiles ? tiles : 1u) : 256u;
    hipLaunchKernelGGL(k_dprep, dim3(tb), dim3(kBlock), 0, s, dl, nd_dev, src, dst, sptr, ws.meta,
                       ws.ntask, ws.tsum, ctr, zero_cnt, host_cnt, m, sup);
    return hipGetLastError();
}

hipError_t delta_apply(const uint32_t* dl, const uint32_t* nd_dev, uint32_t nd_max,
                       const uint32_t* snbr, const uint32_t* seid, uint32_t* sup,
                       const DeltaScratch& ws, DevCounters* ctr, hipStream_t s, uint32_t thr,
                       uint32_t* next, unsigned int* next_cnt, const uint32_t* gsmp,
                       const DirectTasks* dt) {
    if (dt && !dt->ntask) return hipErrorInvalidValue;
    // with no deletions the launch still clears and publishes the counts
    if (!nd_max && !(dt && (dt->zero_pair || dt->host_cnt))) return hipSuccess;
    const DirectTasks d = dt ? *dt : DirectTasks{};
    if (next && (!next_cnt || thr < 2)) return hipErrorInvalidValue;
    // four waves per deletion, 256 to 8192 waves
    const uint32_t wpb = kDBlock / kWave;
    const uint64_t want = (uint64_t)nd_max * 4;
    const uint32_t w = want < 256 ? 256u : (want > 8192 ? 8192u : (uint32_t)want);
    const uint32_t db = (w + wpb - 1) / wpb;
    if (next)
        hipLaunchKernelGGL(k_delta<true>, dim3(db), dim3(kDBlock), 0, s, dl, nd_dev, ws.meta, ws.ntask,
                           ws.tsum, snbr, seid, sup, ctr, ws.bstat, thr, next, next_cnt, gsmp, d);
    else
        hipLaunchKernelGGL(k_delta<false>, dim3(db), dim3(kDBlock), 0, s, dl, nd_dev, ws.meta, ws.ntask,
                           ws.tsum, snbr, seid, sup, ctr, ws.bstat, 0u, nullptr, nullptr, gsmp, d);
    return hipGetLastError();
}

hipError_t block_samples(const uint32_t* snbr, uint32_t S, uint32_t* gsmp, hipStream_t s) {
    if (!S) return hipSuccess;
    const uint32_t b = blocks_for((S + kWave - 1) / kWave, kBlock);
    hipLaunchKernelGGL(k_block_samples, dim3(b < 2048 ? b : 2048), dim3(kBlock), 0, s, snbr, S, gsmp);
    return hipGetLastError();
}

hipError_t row_lengths(const uint32_t* ptr, uint32_t nv, uint32_t* out, hipStream_t s) {
    if (!nv) return hipSuccess;
    const uint32_t b = blocks_for(nv, kBlock);
    hipLaunchKernelGGL(k_row_lengths, dim3(b < 2048 ? b : 2048), dim3(kBlock), 0, s, ptr, nv, out);
    return hipGetLastError();
}

hipError_t delta_stats_reset(const DeltaScratch& ws, hipStream_t s) {
    return fill_u32((uint32_t*)ws.bstat, 0, kStatBytes / 4, s);
}

hipError_t delta_stats_fold(const DeltaScratch& ws, DevCounters* ctr, hipStream_t s) {
    hipLaunchKernelGGL(k_delta_fold, dim3(1), dim3(256), 0, s, ws.bstat, ctr);
    return hipGetLastError();
}

hipError_t finalize_deleted(const uint32_t* dl, const uint32_t* nd_dev, uint32_t nd_max,
                            uint32_t* dst, uint32_t* snbr, const uint32_t* sa, const uint32_t* sb,
                            hipStream_t s) {
    if (!nd_max) return hipSuccess;
    const uint32_t fb = blocks_for(nd_max, kBlock);  // grid-stride: nd_max may be a loose bound
    hipLaunchKernelGGL(k_finalize, dim3(fb < 1024 ? fb : 1024), dim3(kBlock), 0, s, dl, nd_dev, dst,
                       snbr, sa, sb);
    return hipGetLastError();
}

uint32_t compact_tiles_for(uint32_t n) { return (n + kTileC - 1) / kTileC; }

hipError_t compact_adjacency(uint32_t nv, uint32_t S, const uint32_t* sptr, const uint32_t* snbr,
                             const uint32_t* seid, const uint32_t* remap, uint32_t* sptr2,
                             uint32_t* snbr2, uint32_t* seid2, uint32_t S2, const uint32_t* dst,
                             uint32_t* sa, uint32_t* sb, uint32_t* tsum, uint32_t* xs, hipStream_t s) {
    const uint32_t nt = compact_tiles_for(S);
    if (!nt) return hipSuccess;
    hipLaunchKernelGGL(k_cmp_count, dim3(nt), dim3(kBlock), 0, s, S, snbr, tsum);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cmp_scan, dim3(1), dim3(1024), 0, s, tsum, nt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_cmp_scatter, dim3(nt), dim3(kBlock), 0, s, S, snbr, seid, remap, tsum, snbr2,
                       seid2, xs);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_cmp_sptr, dim3(blocks_for((uint64_t)nv + 1, kBlock)), dim3(kBlock), 0, s,
                       sptr, xs, sptr2, nv);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (S2)
        hipLaunchKernelGGL(k_cmp_slots, dim3(blocks_for(S2, kBlock)), dim3(kBlock), 0, s, snbr2, seid2,
                           S2, dst, sa, sb);
    return hipGetLastError();
}

}  // namespace dev
}  // namespace kt
