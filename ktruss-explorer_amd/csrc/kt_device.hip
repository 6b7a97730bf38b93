// kt_device.hip -- gfx950 kernels of the k-truss peel.
//
// Wave64 throughout.  The support count is the hot kernel; everything else
// streams the entry arrays once (coalesced, 4 entries per lane).
//
// Support count (replaces count_triangles_kernel, kernel_ddc.cu:11-54,
// kernel_ddt.cu:11-63, kernel_duc.cu:11-54, kernel_dut.cu:11-65).  A wave owns
// a task of `chunk` consecutive worklist items.  Lanes load the metadata of 64
// items at once; the wave then walks the items in order.  For item e = (u,v)
// the row of u is staged once in the wave's LDS slice (reused by every
// following item of the same row) and the row of v is streamed from HBM 64
// entries at a time; each lane binary-searches its entry in the staged row and
// the wave counts hits with ballot + popcount.  A row longer than the LDS slice
// is searched in place in HBM instead; when u's candidate range is much
// shorter than v's row, the roles flip and u's entries are searched in v's row.
//   directed   : triangle u<v<w found on (u,v) credits (u,v) and (u,w) in LDS
//                (flushed with one atomic per entry) and (v,w) with one global
//                atomic -- the three credits of kernel_ddc.cu:42-50.
//   undirected : |N(u) n N(v)| capped at k-2 with early exit, plain store
//                (kernel_duc.cu:27-50); tiled: capped partial per x-tile class,
//                atomically summed (kernel_dut.cu:57-61).
// Blocks are remapped so each XCD gets a contiguous range of tasks (and so of
// rows), keeping a row's data and its support counters in one XCD's L2.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "kt_internal.h"

namespace kt {
namespace {

constexpr uint32_t kWave = 64;
constexpr uint32_t kNumXcd = 8;

__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nb) {
    // Workgroup bid runs on XCD bid % 8; hand XCD x a contiguous block range.
    const uint32_t q = nb / kNumXcd, r = nb % kNumXcd;
    const uint32_t x = bid % kNumXcd, i = bid / kNumXcd;
    return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t readlane(uint32_t x, uint32_t l) {
    return __builtin_amdgcn_readlane(x, l);
}

// LDS operations of one wave complete in issue order; this keeps the
// compiler from moving LDS accesses across the point where lanes start
// reading what other lanes wrote.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// First index in [lo, hi) whose value (dead bit masked) is >= x.
__device__ __forceinline__ uint32_t lower_bound_masked(const uint32_t* a, uint32_t lo,
                                                       uint32_t hi, uint32_t x) {
    uint32_t len = hi - lo;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if ((a[lo + half] & KT_VMASK) < x) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

// lower_bound_masked for a range [lo, lo + len) that is the same in every lane
// of the wave: the trip count is uniform (a scalar loop) and the body is
// branch-free, so no exec-mask bookkeeping per step.
__device__ __forceinline__ uint32_t lower_bound_uniform(const uint32_t* a, uint32_t lo,
                                                        uint32_t len, uint32_t x) {
    if (len == 0) return lo;
    uint32_t base = lo;
    while (len > 1) {
        const uint32_t half = len >> 1;
        base = (a[base + half] & KT_VMASK) < x ? base + half : base;
        len -= half;
    }
    return base + ((a[base] & KT_VMASK) < x ? 1u : 0u);
}

// Wave-uniform lower bound in an HBM row; lanes split the range 64-way per
// step so the search takes log64(len) dependent loads.
__device__ __forceinline__ uint32_t wave_lower_bound(const uint32_t* a, uint32_t lo, uint32_t hi,
                                                     uint32_t x, uint32_t lane) {
    while (hi - lo > kWave) {
        const uint32_t step = (hi - lo + kWave - 1) / kWave;
        const uint32_t idx = lo + lane * step;
        const bool below = idx < hi && (a[idx] & KT_VMASK) < x;
        const uint32_t nb = (uint32_t)__popcll(ballot(below));  // lanes 0..nb-1 are below
        const uint32_t nlo = nb == 0 ? lo : lo + (nb - 1) * step + 1;
        const uint32_t nhi = lo + nb * step < hi ? lo + nb * step : hi;
        lo = nlo;
        hi = nhi;
    }
    const uint32_t idx = lo + lane;
    const bool below = idx < hi && (a[idx] & KT_VMASK) < x;
    return lo + (uint32_t)__popcll(ballot(below));
}

struct RowState {
    uint32_t u;      // staged row (0xFFFFFFFF none)
    uint32_t ub;     // row start in HBM
    uint32_t ulen;   // stored length
    uint32_t umax;   // largest value in the row (dead bit masked)
    bool in_lds;
};

template <bool DIR>
__device__ void flush_row(const RowState& r, const uint32_t* s_cnt, uint32_t* sup, uint32_t lane) {
    if (!DIR || !r.in_lds || r.u == 0xFFFFFFFFu) return;
    wave_lds_sync();
    for (uint32_t j = lane; j < r.ulen; j += kWave) {
        const uint32_t c = s_cnt[j];
        if (c) atomicAdd(sup + r.ub + j, c);
    }
}

template <bool DIR>
__device__ void stage_row(RowState& r, uint32_t u, const uint32_t* ptr, const uint32_t* dst,
                          uint32_t* s_val, uint32_t* s_cnt, uint32_t cap_lds, uint32_t lane) {
    r.u = u;
    r.ub = ptr[u];
    r.ulen = ptr[u + 1] - r.ub;
    r.in_lds = r.ulen <= cap_lds;
    if (r.in_lds) {
        wave_lds_sync();  // previous readers of the slice are done
        for (uint32_t j = lane; j < r.ulen; j += kWave) {
            s_val[j] = dst[r.ub + j];
            if (DIR) s_cnt[j] = 0;
        }
        wave_lds_sync();
        r.umax = r.ulen ? (s_val[r.ulen - 1] & KT_VMASK) : 0;
    } else {
        r.umax = r.ulen ? (dst[r.ub + r.ulen - 1] & KT_VMASK) : 0;
    }
}

// Directed entry e = (u, v): credits of every triangle u < v < w.
// u's candidates are row positions [lo, ulen) (the entries after v); v's row
// is [vb, vb + vlen).  Tiled formats pass block rows (the class-y parts of
// both rows), so the same code counts one x-tile class.
__device__ uint32_t count_directed_item(const CountArgs& a, const RowState& r, uint32_t lo, uint32_t vb,
                                        uint32_t vlen, uint32_t* s_val, uint32_t* s_cnt, uint32_t lane) {
    const uint32_t* dst = a.dst;
    uint32_t* sup = a.sup;
    const uint32_t ulen = r.ulen;
    if (lo >= ulen || vlen == 0) return 0;
    const uint32_t* urow = r.in_lds ? s_val : dst + r.ub;
    const uint32_t umin = urow[lo] & KT_VMASK;
    uint32_t c = 0;

    if ((ulen - lo) * a.query_factor < vlen) {
        // Query mode: u's candidates searched in v's row (HBM)
        const uint32_t vmax = dst[vb + vlen - 1] & KT_VMASK;
        for (uint32_t p0 = lo; p0 < ulen; p0 += kWave) {
            const uint32_t p = p0 + lane;
            const uint32_t x = p < ulen ? urow[p] : KT_DEAD;
            bool hit = false;
            uint32_t j = 0;
            if (!(x & KT_DEAD) && x <= vmax) {
                j = lower_bound_masked(dst + vb, 0, vlen, x);
                hit = j < vlen && dst[vb + j] == x;
            }
            if (hit) {
                if (r.in_lds) atomicAdd(s_cnt + p, 1u);
                else atomicAdd(sup + r.ub + p, 1u);
                atomicAdd(sup + vb + j, 1u);
            }
            c += (uint32_t)__popcll(ballot(hit));
            if (ballot(p < ulen && (x & KT_VMASK) > vmax)) break;
        }
        return c;
    }

    // Stream mode: v's row streamed, each entry searched among u's candidates.
    const uint32_t j0 = vlen > 2 * kWave ? wave_lower_bound(dst + vb, 0, vlen, umin, lane) : 0u;
    for (uint32_t jb = j0; jb < vlen; jb += kWave) {
        const uint32_t j = jb + lane;
        const bool inr = j < vlen;
        const uint32_t w = inr ? dst[vb + j] : KT_VMASK;
        const uint32_t wm = w & KT_VMASK;
        bool hit = false;
        uint32_t p = 0;
        if (inr && !(w & KT_DEAD) && wm <= r.umax) {
            p = lower_bound_masked(urow, lo, ulen, wm);
            hit = p < ulen && urow[p] == wm;
        }
        if (hit) {
            if (r.in_lds) atomicAdd(s_cnt + p, 1u);
            else atomicAdd(sup + r.ub + p, 1u);
            atomicAdd(sup + vb + j, 1u);
        }
        c += (uint32_t)__popcll(ballot(hit));
        if (ballot(inr && wm > r.umax)) break;
    }
    return c;
}

// Undirected entry e = (u, v): |N(u) n N(v)|, stopping at cap (tiled: the
// class-y block rows of both, one x-tile class's partial, and cap what the
// item still lacks of k-2).
__device__ uint32_t count_undirected_item(const CountArgs& a, const RowState& r, uint32_t vb, uint32_t vlen,
                                          const uint32_t* s_val, uint32_t lane, uint32_t cap) {
    const uint32_t* dst = a.dst;
    const uint32_t ulen = r.ulen;
    if (ulen == 0 || vlen == 0 || cap == 0) return 0;
    const uint32_t* urow = r.in_lds ? s_val : dst + r.ub;
    uint32_t c = 0;

    if (ulen * a.query_factor < vlen) {
        const uint32_t vmax = dst[vb + vlen - 1] & KT_VMASK;
        for (uint32_t p0 = 0; p0 < ulen; p0 += kWave) {
            const uint32_t p = p0 + lane;
            const uint32_t x = p < ulen ? urow[p] : KT_DEAD;
            bool hit = false;
            if (!(x & KT_DEAD) && x <= vmax) {
                const uint32_t j = lower_bound_masked(dst + vb, 0, vlen, x);
                hit = j < vlen && dst[vb + j] == x;
            }
            c += (uint32_t)__popcll(ballot(hit));
            if (c >= cap) return cap;
            if (ballot(p < ulen && (x & KT_VMASK) > vmax)) break;
        }
        return c;
    }

    const uint32_t umin = urow[0] & KT_VMASK;
    const uint32_t j0 = vlen > 2 * kWave ? wave_lower_bound(dst + vb, 0, vlen, umin, lane) : 0u;
    for (uint32_t jb = j0; jb < vlen; jb += kWave) {
        const uint32_t j = jb + lane;
        const bool inr = j < vlen;
        const uint32_t w = inr ? dst[vb + j] : KT_VMASK;
        const uint32_t wm = w & KT_VMASK;
        bool hit = false;
        if (inr && !(w & KT_DEAD) && wm <= r.umax) {
            const uint32_t p = lower_bound_masked(urow, 0, ulen, wm);
            hit = p < ulen && urow[p] == wm;
        }
        c += (uint32_t)__popcll(ballot(hit));
        if (c >= cap) return cap;
        if (ballot(inr && wm > r.umax)) break;
    }
    return c < cap ? c : cap;
}


// Adds a wave's statistics to its stripe line (one line per wave index modulo
// kStripes, so the atomics of a launch are spread over kStripes addresses).
__device__ __forceinline__ void stripe_stats(const CountArgs& a, unsigned long long merge,
                                             unsigned long long rec) {
    const uint32_t w = ((blockIdx.y * gridDim.x + blockIdx.x) * (blockDim.x / kWave) +
                        threadIdx.x / kWave) % kStripes;
    unsigned long long* q = a.stripes + (size_t)w * kStripeWords;
    atomicAdd(q, merge);
    atomicAdd(q + 1, rec);
}

// Items of a tiled count whose two class-y block rows hold at most this many
// entries together are merged by one lane (fine tiles leave most block rows
// a few entries long, too little for a wave).
constexpr uint32_t kLaneMerge = 48;

// One lane's item e = (u, v) of a tiled count: u's block row [ub, ub + ulen)
// merged with v's [vb, vb + vlen).  Directed: the triangles u < v < w of
// this class -- u's entries after v (from e itself when e lies in the block
// row: own), credits on (u,v), (u,w), (v,w) by atomics.  Undirected: the
// common neighbours up to cap, added to e's support.
template <bool DIR>
__device__ void lane_merge_item(const CountArgs& a, uint32_t e, uint32_t v, uint32_t ub, uint32_t ulen, uint32_t vb,
                                uint32_t vlen, bool own, uint32_t cap) {
    const uint32_t* dst = a.dst;
    uint32_t i = ub, iend = ub + ulen, j = vb, jend = vb + vlen, c = 0;
    if (DIR) i = own ? e + 1 : ub + lower_bound_masked(dst + ub, 0, ulen, v + 1);
    if (DIR) cap = 0xFFFFFFFFu;
    while (i < iend && j < jend && c < cap) {
        const uint32_t x = dst[i], w = dst[j];
        if (x & KT_DEAD) { ++i; continue; }
        if (w & KT_DEAD) { ++j; continue; }
        if (x < w) {
            ++i;
        } else if (x > w) {
            ++j;
        } else {
            if (DIR) {
                atomicAdd(a.sup + i, 1u);
                atomicAdd(a.sup + j, 1u);
            }
            ++c;
            ++i;
            ++j;
        }
    }
    if (c) {
        if (DIR || a.ytiles > 1) atomicAdd(a.sup + e, c);
        else a.sup[e] = c;
    }
}

// A lane's item whose block rows both hold at most kRegRow entries: both are
// loaded into registers with independent loads (one memory round trip) and
// compared all against all; the same credits as lane_merge_item.
constexpr uint32_t kRegRow = 16;

template <bool DIR>
__device__ void lane_regs_item(const CountArgs& a, uint32_t e, uint32_t v, uint32_t ub, uint32_t ulen, uint32_t vb,
                               uint32_t vlen, uint32_t cap) {
    const uint32_t* dst = a.dst;
    uint32_t uu[kRegRow], vv[kRegRow];
#pragma unroll
    for (uint32_t t = 0; t < kRegRow; ++t) uu[t] = t < ulen ? dst[ub + t] : KT_DEAD;
#pragma unroll
    for (uint32_t t = 0; t < kRegRow; ++t) vv[t] = t < vlen ? dst[vb + t] : KT_DEAD;
    uint32_t c = 0;
#pragma unroll
    for (uint32_t i = 0; i < kRegRow; ++i) {
        const uint32_t x = uu[i];
        // directed: u's candidates are the entries after v (rows ascend)
        if (!(x & KT_DEAD) && (!DIR || x > v)) {
#pragma unroll
            for (uint32_t j = 0; j < kRegRow; ++j) {
                if (vv[j] == x) {  // a dead entry of v's row has bit 31 set: never equal
                    ++c;
                    if (DIR) {
                        atomicAdd(a.sup + ub + i, 1u);
                        atomicAdd(a.sup + vb + j, 1u);
                    }
                }
            }
        }
    }
    if (!DIR && c > cap) c = cap;
    if (c) {
        if (DIR || a.ytiles > 1) atomicAdd(a.sup + e, c);
        else a.sup[e] = c;
    }
}

// Support count per item e = (u, v) with u's row staged in the wave's LDS
// slice (undirected COO+CSR; both tiled formats).  Tiled formats: y-block y
// (blockIdx.y) counts the triangles whose third vertex w lies in a tile of
// class y -- the x-tile segments of kernel_ddt.cu:25-34 / kernel_dut.cu:25-34,
// which the block layout stores contiguously as u's and v's block rows of
// class y.  u's block row is staged (in stored order an item's row changes
// once per block); summed over the classes (directed: credits added by
// atomics; undirected: per-class partials capped at k-2 and summed
// atomically, kernel_dut.cu:57-61) the supports equal the untiled count's.
template <bool DIR, bool TILED>
__global__ __launch_bounds__(1024) void k_count(CountArgs a) {
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t nw = blockDim.x / kWave;
    uint32_t* s_val = lds + wid * (DIR ? 2u : 1u) * a.cap_lds;
    uint32_t* s_cnt = s_val + a.cap_lds;
    const uint32_t S = a.tile_size, Y = TILED ? a.ytiles : 1u;
    // tiled: blockIdx.y is the x-tile class (the y-blocks of kernel_ddt.cu:28
    // and kernel_dut.cu:28); a class's workgroups run together and the classes
    // one after another, so no two workgroups credit one item's support at once
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t y0 = TILED ? blockIdx.y : 0u;
    const uint32_t task = lb * nw + wid;
    const uint64_t first = (uint64_t)task * a.chunk;
    if (first >= a.nitems) return;
    const uint32_t i0 = (uint32_t)first;
    const uint32_t i1 = a.nitems - i0 < a.chunk ? a.nitems : i0 + a.chunk;
    const uint32_t* ptr = a.ptr;
    const uint32_t* dst = a.dst;

    RowState r;
    r.u = 0xFFFFFFFFu;
    r.ub = r.ulen = r.umax = 0;
    r.in_lds = false;
    unsigned long long acc_merge = 0, acc_rec = 0;

    for (uint32_t base = i0; base < i1; base += kWave) {
        const uint32_t my = base + lane;
        uint32_t e = 0, u = 0, v = KT_DEAD;
        if (my < i1) {
            e = a.wl ? a.wl[my] : my;
            v = dst[e];
            if (!(v & KT_DEAD)) u = a.src[e];
        }
        const bool live = my < i1 && !(v & KT_DEAD);
        if (live && y0 == 0) acc_rec += 1;
        // tiled: this workgroup's x-tile class, a y-block's share of
        // kernel_ddt.cu:28 / kernel_dut.cu:28; untiled: the one pass
        for (uint32_t y = y0; y < y0 + 1; ++y) {
            uint32_t vb = 0, vlen = 0, icap = a.cap;
            bool work = live;
            if (live) {
                if (TILED) {
                    const uint32_t vr = tiled_row(v, y, S, Y), ur = tiled_row(u, y, S, Y);
                    vb = ptr[vr];
                    vlen = ptr[vr + 1] - vb;
                    const uint32_t ub = ptr[ur], ulen = ptr[ur + 1] - ub;
                    // class-y block rows are short: the lane settles the item's
                    // statistics and drops it when a block row is empty
                    // (directed: or holds nothing after v; undirected: or a
                    // whole stored row is shorter than k-2, the exact
                    // pre-filter of kernel_duc.cu:27)
                    uint32_t lu = ulen;
                    if (DIR && ulen) lu -= lower_bound_masked(dst + ub, 0, ulen, (v / S) * S);
                    acc_merge += (unsigned long long)lu + vlen;
                    work = ulen && vlen && (!DIR || (dst[ub + ulen - 1] & KT_VMASK) > v);
                    if (!DIR && work) {
                        if (a.rowlen) work = a.rowlen[u] >= a.cap && a.rowlen[v] >= a.cap;
                        // Only "support >= k-2" is decided: a class counts what
                        // the item still lacks after the classes that added to
                        // its support already.  If one class reaches its cap,
                        // the sum reaches k-2 (the others add what it saw); if
                        // none does, every partial is whole: the decision is
                        // the full count's either way.
                        const uint32_t have = a.sup[e];
                        icap = have < a.cap ? a.cap - have : 0u;
                        work = work && icap;
                    }
                    if (work && ulen <= kRegRow && vlen <= kRegRow) {
                        lane_regs_item<DIR>(a, e, v, ub, ulen, vb, vlen, icap);
                        work = false;
                    } else if (work && ulen + vlen <= kLaneMerge) {
                        // both block rows short: this lane merges them alone
                        lane_merge_item<DIR>(a, e, v, ub, ulen, vb, vlen, (v / S) % Y == y, icap);
                        work = false;
                    }
                } else {
                    vb = ptr[v];
                    vlen = ptr[v + 1] - vb;
                }
            }
            uint64_t todo = ballot(work);
            while (todo) {
                const uint32_t l = (uint32_t)__ffsll((long long)todo) - 1;
                todo &= todo - 1;
                const uint32_t eu = readlane(u, l), ee = readlane(e, l), ev = readlane(v, l);
                const uint32_t evb = readlane(vb, l), evlen = readlane(vlen, l), ecap = readlane(icap, l);
                const uint32_t ur = TILED ? tiled_row(eu, y, S, Y) : eu;  // u's row (of class y)
                if (ur != r.u) {
                    flush_row<DIR>(r, s_cnt, a.sup, lane);
                    stage_row<DIR>(r, ur, ptr, dst, s_val, s_cnt, a.cap_lds, lane);
                }
                const uint32_t* urow = r.in_lds ? s_val : dst + r.ub;
                // e itself lies in the staged row (always untiled; tiled when
                // v's tile is of class y)
                const bool own = !TILED || (ev / S) % Y == y;
                if (!TILED && lane == 0) acc_merge += (unsigned long long)r.ulen + evlen;  // summed over lanes below
                if (DIR) {
                    // u's candidates: the entries after v
                    uint32_t lo = ee - r.ub + 1;
                    if (!own)
                        lo = r.in_lds ? lower_bound_masked(urow, 0, r.ulen, ev + 1)
                                      : wave_lower_bound(urow, 0, r.ulen, ev + 1, lane);
                    const uint32_t c = count_directed_item(a, r, lo, evb, evlen, s_val, s_cnt, lane);
                    if (c && lane == 0) {
                        if (own && r.in_lds) atomicAdd(s_cnt + (ee - r.ub), c);
                        else atomicAdd(a.sup + ee, c);
                    }
                } else {
                    uint32_t c = 0;
                    if (!a.prefilter || (r.ulen >= a.cap && evlen >= a.cap))
                        c = count_undirected_item(a, r, evb, evlen, s_val, lane, ecap);
                    if (lane == 0) {
                        if (TILED && Y > 1) {
                            if (c) atomicAdd(a.sup + ee, c);
                        } else {
                            a.sup[ee] = c;
                        }
                    }
                }
            }
        }
    }
    flush_row<DIR>(r, s_cnt, a.sup, lane);
    for (int off = 32; off > 0; off >>= 1) {  // statistics kept per lane
        acc_merge += __shfl_xor(acc_merge, off);
        acc_rec += __shfl_xor(acc_rec, off);
    }
    if (lane == 0 && (acc_rec || acc_merge)) stripe_stats(a, acc_merge, acc_rec);
}

__global__ void k_tiled_row_totals(const uint32_t* __restrict__ ptr, uint32_t n, uint32_t S, uint32_t Y,
                                   uint32_t* __restrict__ out) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        uint32_t t = 0;
        for (uint32_t y = 0; y < Y; ++y) {
            const uint32_t row = tiled_row(u, y, S, Y);
            t += ptr[row + 1] - ptr[row];
        }
        out[u] = t;
    }
}

// ---------------------------------------------------------------------------
// Directed support count on the COO+CSR layout, flattened.
//
// A wave takes blocks of 64 consecutive items; lane i holds item i, the entry
// (u_i, v_i) at position q_i of u's row.  Consecutive items of one row with
// increasing positions form a run, and each run's row tail [q_head, ulen) is
// staged once in the wave's LDS slice (a run longer than the slice is searched
// in HBM).  Lane i narrows v_i's row to the window of values that u's
// candidates can match -- [value at q_i+1, last value of u's row] -- by two
// binary searches.  The 64 windows are concatenated (exclusive scan) and
// streamed 64 entries per round, one independent load per lane, so no lane
// idles on a short row and no round waits on another item's latency.  Every
// streamed entry w is binary-searched among its item's candidates in LDS; a
// hit credits (u,v) and (u,w) through 16-bit LDS counters (flushed with one
// atomic per staged entry) and (v,w) with one global atomic.

struct FlatItem {     // per-item LDS record (32 bytes)
    uint32_t vb;      // v's row start
    uint32_t lo;      // window start within v's row
    uint32_t pre;     // exclusive prefix of window lengths
    uint32_t eslot;   // LDS slot of the item's own entry, or its HBM position
    uint32_t slo;     // candidate range [slo, shi): LDS slots or HBM positions
    uint32_t shi;
    uint32_t hbm;     // 1: candidates searched in HBM
    uint32_t pad;
};

constexpr uint32_t kFlatFixedWords = 68 + 68 + 64 * 8;  // s_roff, s_pre, items

__host__ __device__ __forceinline__ uint32_t flat_words(uint32_t cap) {
    return cap + cap / 2 + kFlatFixedWords;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, uint32_t lane) {
    for (uint32_t off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    return x;
}

// First index in [0, n) with a[idx] > t (a non-decreasing).
__device__ __forceinline__ uint32_t upper_bound_u32(const uint32_t* a, uint32_t n, uint32_t t) {
    uint32_t lo = 0, len = n;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (a[lo + half] <= t) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}

__device__ __forceinline__ void lds_credit16(uint32_t* s_cnt, uint32_t slot) {
    atomicAdd(s_cnt + (slot >> 1), 1u << ((slot & 1) << 4));
}

__device__ __forceinline__ void lds_credit16_n(uint32_t* s_cnt, uint32_t slot, uint32_t n) {
    atomicAdd(s_cnt + (slot >> 1), n << ((slot & 1) << 4));
}

// MERGE = false: every streamed entry is binary-searched among its item's
// candidates (round-robin lanes, coalesced loads).  MERGE = true: each lane
// takes a contiguous share of the concatenated windows and merges it against
// the staged candidates, which are sorted like the window, so a lane advances
// its LDS cursor instead of searching (the merge of kernel_hdc.cu:22-44 with
// u's row in LDS and the work split evenly over the wave).
template <bool MERGE>
__global__ __launch_bounds__(1024) void k_count_flat(CountArgs a) {
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = threadIdx.x / kWave;
    const uint32_t nw = blockDim.x / kWave;
    const uint32_t cap = a.cap_lds;
    uint32_t* s_val = lds + wid * flat_words(cap);
    uint32_t* s_cnt = s_val + cap;
    uint32_t* s_roff = s_cnt + cap / 2;
    uint32_t* s_pre = s_roff + 68;
    FlatItem* s_item = reinterpret_cast<FlatItem*>(s_pre + 68);
    const uint32_t task = xcd_block(blockIdx.x, gridDim.x) * nw + wid;
    const uint64_t first = (uint64_t)task * a.chunk;
    if (first >= a.nitems) return;
    const uint32_t i0 = (uint32_t)first;
    const uint32_t i1 = a.nitems - i0 < a.chunk ? a.nitems : i0 + a.chunk;
    const uint32_t* ptr = a.ptr;
    const uint32_t* dst = a.dst;
    uint32_t* sup = a.sup;
    unsigned long long acc_merge = 0, acc_rec = 0;

    for (uint32_t base = i0; base < i1; base += kWave) {
        const uint32_t my = base + lane;
        const bool in = my < i1;
        uint32_t e = 0, u = 0xFFFFFFFFu, v = KT_DEAD, ub = 0, ulen = 0, vb = 0, vlen = 0;
        if (in) {
            e = a.wl ? a.wl[my] : my;
            v = dst[e];
            u = a.src[e];
            ub = ptr[u];
            ulen = ptr[u + 1] - ub;
            if (!(v & KT_DEAD)) {
                vb = ptr[v];
                vlen = ptr[v + 1] - vb;
            }
        }
        const bool live = in && !(v & KT_DEAD);
        const uint32_t q = e - ub;
        // runs: same row, increasing positions; lane 0 is always in range
        const uint32_t pu = __shfl_up(u, 1), pe = __shfl_up(e, 1);
        const bool head = in && (lane == 0 || u != pu || e < pe);
        const uint64_t heads = ballot(head);
        const uint32_t h = 63u - (uint32_t)__clzll(heads & ((2ull << lane) - 1ull));
        const uint32_t rsize = ulen - q;  // staged length if this lane heads a run
        const uint32_t gsrc = ub + q;     // HBM position of a run's first staged entry

        uint64_t remaining = heads;
        while (remaining) {
            // A pass takes the longest prefix of remaining runs that fits the slice.
            const bool rem_head = (remaining >> lane) & 1ull;
            const uint32_t ssz = (rem_head && rsize <= cap) ? rsize : 0u;
            const uint32_t cum = wave_incl_scan(ssz, lane);
            const bool take = rem_head && cum <= cap;
            const uint64_t pass = ballot(take);
            remaining &= ~pass;
            const uint32_t psz = take ? ssz : 0u;
            const uint32_t pcum = wave_incl_scan(psz, lane);
            const uint32_t poff = pcum - psz;
            const uint32_t T = readlane(pcum, 63);
            wave_lds_sync();
            s_roff[lane] = poff;
            if (lane == 0) s_roff[64] = T;
            wave_lds_sync();
            for (uint32_t t0 = 0; t0 < T; t0 += kWave) {
                const uint32_t t = t0 + lane;
                const uint32_t r = upper_bound_u32(s_roff, 65, t < T ? t : T - 1) - 1;
                const uint32_t g = __shfl(gsrc, r);
                if (t < T) s_val[t] = dst[g + (t - s_roff[r])];
            }
            for (uint32_t w = lane; w < (T + 1) / 2; w += kWave) s_cnt[w] = 0;
            wave_lds_sync();

            // items of this pass: candidate window in v's row
            const bool inpass = in && ((pass >> h) & 1ull);
            const uint32_t h_poff = __shfl(poff, h);
            const uint32_t h_q = __shfl(q, h);
            const uint32_t h_rsize = __shfl(rsize, h);
            const bool staged = h_rsize <= cap;
            const bool has = inpass && live && q + 1 < ulen && vlen > 0;
            uint32_t eslot, slo, shi;
            if (staged) {
                eslot = h_poff + (q - h_q);
                slo = eslot + 1;
                shi = h_poff + (ulen - h_q);
            } else {
                eslot = e;
                slo = ub + q + 1;
                shi = ub + ulen;
            }
            uint32_t lo = 0, hi = 0;
            if (has) {
                const uint32_t* rows = staged ? s_val : dst;
                const uint32_t umin = rows[slo] & KT_VMASK;
                const uint32_t umax = rows[shi - 1] & KT_VMASK;
                if (vlen <= 16) {
                    hi = vlen;
                } else {
                    lo = lower_bound_masked(dst + vb, 0, vlen, umin);
                    hi = lower_bound_masked(dst + vb, lo, vlen, umax + 1);
                }
            }
            if (inpass && live) {
                acc_rec += 1;
                acc_merge += (unsigned long long)ulen + vlen;
            }
            const uint32_t work = hi - lo;
            const uint32_t wincl = wave_incl_scan(work, lane);
            const uint32_t pre = wincl - work;
            const uint32_t W = readlane(wincl, 63);
            s_pre[lane] = pre;
            if (lane == 0) s_pre[64] = W;
            FlatItem it;
            it.vb = vb;
            it.lo = lo;
            it.pre = pre;
            it.eslot = eslot;
            it.slo = slo;
            it.shi = shi;
            it.hbm = staged ? 0u : 1u;
            it.pad = 0;
            s_item[lane] = it;
            wave_lds_sync();

            if (MERGE) {
                // lane-contiguous share [t, tend) of the concatenated windows
                const uint32_t per = (W + kWave - 1) / kWave;
                uint32_t t = lane * per;
                const uint32_t tend = t + per < W ? t + per : W;
                if (t < tend) {
                    uint32_t i = upper_bound_u32(s_pre, 65, t) - 1;
                    FlatItem x = s_item[i];
                    uint32_t nextb = s_pre[i + 1];
                    uint32_t p = 0, ecount = 0;
                    bool fresh = true;
                    for (; t < tend; ++t) {
                        while (t >= nextb) {  // next item (empty windows skipped)
                            if (ecount) {
                                if (x.hbm) atomicAdd(sup + x.eslot, ecount);
                                else lds_credit16_n(s_cnt, x.eslot, ecount);
                            }
                            ecount = 0;
                            ++i;
                            x = s_item[i];
                            nextb = s_pre[i + 1];
                            fresh = true;
                        }
                        const uint32_t j = x.lo + (t - x.pre);
                        const uint32_t w = dst[x.vb + j];
                        if (w & KT_DEAD) continue;
                        const uint32_t* cand = x.hbm ? dst : s_val;
                        if (fresh) {
                            p = lower_bound_masked(cand, x.slo, x.shi, w);
                            fresh = false;
                        } else {
                            // advance the cursor: a few linear steps, then search
                            uint32_t steps = 0;
                            while (p < x.shi && (cand[p] & KT_VMASK) < w && steps < 4) {
                                ++p;
                                ++steps;
                            }
                            if (p < x.shi && (cand[p] & KT_VMASK) < w)
                                p = lower_bound_masked(cand, p, x.shi, w);
                        }
                        if (p < x.shi && cand[p] == w) {
                            if (x.hbm) atomicAdd(sup + p, 1u);
                            else lds_credit16(s_cnt, p);
                            ++ecount;
                            atomicAdd(sup + x.vb + j, 1u);
                        }
                    }
                    if (ecount) {
                        if (x.hbm) atomicAdd(sup + x.eslot, ecount);
                        else lds_credit16_n(s_cnt, x.eslot, ecount);
                    }
                }
            }
            for (uint32_t t0 = 0; !MERGE && t0 < W; t0 += kWave) {
                const uint32_t t = t0 + lane;
                if (t < W) {
                    const uint32_t i = upper_bound_u32(s_pre, 65, t) - 1;
                    const FlatItem x = s_item[i];
                    const uint32_t j = x.lo + (t - x.pre);
                    const uint32_t w = dst[x.vb + j];
                    if (!(w & KT_DEAD)) {
                        if (!x.hbm) {
                            const uint32_t p = lower_bound_masked(s_val, x.slo, x.shi, w);
                            if (p < x.shi && s_val[p] == w) {
                                lds_credit16(s_cnt, p);
                                lds_credit16(s_cnt, x.eslot);
                                atomicAdd(sup + x.vb + j, 1u);
                            }
                        } else {
                            const uint32_t p = lower_bound_masked(dst, x.slo, x.shi, w);
                            if (p < x.shi && dst[p] == w) {
                                atomicAdd(sup + p, 1u);
                                atomicAdd(sup + x.eslot, 1u);
                                atomicAdd(sup + x.vb + j, 1u);
                            }
                        }
                    }
                }
            }
            wave_lds_sync();
            if (W) {
                for (uint32_t t0 = 0; t0 < T; t0 += kWave) {
                    const uint32_t t = t0 + lane;
                    const uint32_t r = upper_bound_u32(s_roff, 65, t < T ? t : T - 1) - 1;
                    const uint32_t g = __shfl(gsrc, r);
                    if (t < T) {
                        const uint32_t c = (s_cnt[t >> 1] >> ((t & 1) << 4)) & 0xFFFFu;
                        if (c) atomicAdd(sup + g + (t - s_roff[r]), c);
                    }
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc_rec += __shfl_xor(acc_rec, off);
        acc_merge += __shfl_xor(acc_merge, off);
    }
    if (lane == 0 && acc_rec) stripe_stats(a, acc_merge, acc_rec);
}

// ---------------------------------------------------------------------------
// Directed support count on the COO+CSR layout, row-staged and pipelined.
//
// Lanes load 64 items at once and, lane-parallel, narrow each item's v-row to
// the window of values its u-candidates span: [next entry of u's row, last
// entry of u's row], found by two binary searches (the first bound is simply
// the entry after e in HBM).  The wave then walks (item, 64-entry chunk)
// pairs in order; the load of the next chunk -- of the same item or of the
// next one -- is issued before the current chunk is searched, so one HBM round
// trip overlaps the LDS work of the previous chunk.  u's row is staged in LDS
// once per run of items of that row; credits to (u,v) and (u,w) go to 16-bit
// LDS counters, the credit to (v,w) is one global atomic.
//
// BATCH: the kernel is persistent over a task list of many graphs stored as
// one disjoint union; each wave walks tasks wave_id, wave_id + W, ... and a
// task is `chunk` consecutive entries of one graph.
constexpr uint32_t kCountAhead = 2;  // window chunks queued per wave (loads in flight)
template <bool BATCH>
__global__ __launch_bounds__(1024) void k_count_dir(CountArgs a, TaskList tl) {
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t nw = blockDim.x / kWave;
    const uint32_t cap = a.cap_lds;
    uint32_t* s_val = lds + wid * (cap + cap / 2);
    uint32_t* s_cnt = s_val + cap;
    const uint32_t task0 = xcd_block(blockIdx.x, gridDim.x) * nw + wid;
    const uint32_t* ptr = a.ptr;
    const uint32_t* dst = a.dst;
    uint32_t* sup = a.sup;
    // staged row
    uint32_t su = 0xFFFFFFFFu, sub = 0, sulen = 0;
    bool s_in_lds = false;
    unsigned long long acc_merge = 0, acc_rec = 0;

    auto flush = [&]() {
        if (su == 0xFFFFFFFFu || !s_in_lds) return;
        wave_lds_sync();
        for (uint32_t j = lane; j < sulen; j += kWave) {
            const uint32_t c = (s_cnt[j >> 1] >> ((j & 1) << 4)) & 0xFFFFu;
            if (c) atomicAdd(sup + sub + j, c);
        }
    };
    auto stage = [&](uint32_t u) {
        su = u;
        sub = ptr[u];
        sulen = ptr[u + 1] - sub;
        s_in_lds = sulen <= cap;
        if (s_in_lds) {
            wave_lds_sync();
            for (uint32_t j = lane; j < sulen; j += kWave) s_val[j] = dst[sub + j];
            for (uint32_t j = lane; j < (sulen + 1) / 2; j += kWave) s_cnt[j] = 0;
            wave_lds_sync();
        }
    };

    auto run = [&](uint32_t i0, uint32_t i1) {
    for (uint32_t base = i0; base < i1; base += kWave) {
        const uint32_t my = base + lane;
        uint32_t e = 0, u = 0, v = KT_DEAD, ub = 0, ulen = 0, vb = 0, vlen = 0;
        if (my < i1) {
            e = a.wl ? a.wl[my] : my;
            v = dst[e];
            if (!(v & KT_DEAD)) {
                u = a.src[e];
                ub = ptr[u];
                ulen = ptr[u + 1] - ub;
                vb = ptr[v];
                vlen = ptr[v + 1] - vb;
            }
        }
        const bool live = my < i1 && !(v & KT_DEAD);
        if (live) {
            acc_rec += 1;
            acc_merge += (unsigned long long)ulen + vlen;
        }
        uint32_t lo = 0, hi = 0;
        if (live && e + 1 < ub + ulen && vlen > 0) {
            const uint32_t umin = dst[e + 1] & KT_VMASK;
            const uint32_t umax = dst[ub + ulen - 1] & KT_VMASK;
            if (vlen <= 16) {
                hi = vlen;
            } else {
                lo = lower_bound_masked(dst + vb, 0, vlen, umin);
                hi = lower_bound_masked(dst + vb, lo, vlen, umax + 1);
            }
        }
        uint64_t todo = ballot(hi > lo);
        if (!todo) continue;
        // The batch's (item, 64-entry chunk of v's window) pairs in order,
        // kCountAhead of them queued with their window loads in flight, so a
        // chunk's load arrives while the chunk before it is searched.  The
        // slots rotate in place: the next chunk's entries held in the
        // register the load went to (copied to the current chunk's, the
        // copy waited for every load before it, the prefetch just issued
        // included, so no search overlapped a load).  Two slots: R-MAT 22's
        // first count 104.7 -> 99.6 ms; three or four, 139 ms.  Slot q:
        // item lane ql, first window slot qj, its entries qw.
        uint32_t pl = (uint32_t)__ffsll((long long)todo) - 1;  // the item being cut into chunks
        todo &= todo - 1;
        uint32_t pj = readlane(lo, pl), phi = readlane(hi, pl), pvb = readlane(vb, pl);
        bool pdone = false;
        auto produce = [&](uint32_t& ql, uint32_t& qj, uint32_t& qw, bool& qv) {
            qv = !pdone;
            ql = pl;
            qj = pj;
            // unconditional, clamped to the window (a load under a condition,
            // or a select on its result, waits for it here); lanes past the
            // window are masked when the chunk is searched
            qw = dst[pvb + min(pj + lane, phi - 1)];
            if (pdone) return;
            pj += kWave;
            if (pj >= phi) {
                if (todo) {
                    pl = (uint32_t)__ffsll((long long)todo) - 1;
                    todo &= todo - 1;
                    pj = readlane(lo, pl);
                    phi = readlane(hi, pl);
                    pvb = readlane(vb, pl);
                } else {
                    pdone = true;
                }
            }
        };
        static_assert(kCountAhead == 2, "two queue slots below");
        uint32_t l0, l1, j0, j1, w0, w1;
        bool v0, v1;
        produce(l0, j0, w0, v0);
        produce(l1, j1, w1, v1);
        uint32_t c = 0;
        // the oldest slot's chunk is searched after its slot is refilled; the
        // slots rotate in place (a register copy of an entry still in flight
        // would wait for every load before it)
        auto consume = [&](uint32_t& ql, uint32_t& qj, uint32_t& qw, bool& qv) {
            const uint32_t cl = ql, cj = qj;
            const uint32_t cu = readlane(u, cl), ce = readlane(e, cl), cvb = readlane(vb, cl);
            const uint32_t chi = readlane(hi, cl);
            const uint32_t w = cj + lane < chi ? qw : KT_DEAD;
            produce(ql, qj, qw, qv);  // issued before this chunk's search
            if (cu != su) {
                flush();
                stage(cu);
                // the rare path's loads and atomics all done, so the compiler
                // keeps counting the queued chunk loads precisely
                __builtin_amdgcn_s_waitcnt(0);
            }
            const uint32_t q = ce - sub;
            bool hit = false;
            // candidates are row positions [q+1, sulen): the same range for the whole chunk
            if (s_in_lds) {
                const uint32_t p = lower_bound_uniform(s_val, q + 1, sulen - q - 1, w & KT_VMASK);
                hit = !(w & KT_DEAD) && p < sulen && s_val[p] == w;
                if (hit) lds_credit16(s_cnt, p);
            } else {
                const uint32_t p = lower_bound_uniform(dst, sub + q + 1, sulen - q - 1, w & KT_VMASK);
                hit = !(w & KT_DEAD) && p < sub + sulen && dst[p] == w;
                if (hit) atomicAdd(sup + p, 1u);
                __builtin_amdgcn_s_waitcnt(0);
            }
            if (hit) atomicAdd(sup + cvb + cj + lane, 1u);
            c += (uint32_t)__popcll(ballot(hit));
            if (cj + kWave >= chi) {  // last chunk of this item: credit (u, v)
                if (c && lane == 0) {
                    if (s_in_lds) lds_credit16_n(s_cnt, q, c);
                    else atomicAdd(sup + ce, c);
                }
                c = 0;
            }
        };
        while (v0) {
            consume(l0, j0, w0, v0);
            if (!v1) break;
            consume(l1, j1, w1, v1);
        }
    }
    };
    if (BATCH) {
        const uint32_t ntasks = *tl.ntasks;
        const uint32_t W = gridDim.x * nw;
        for (uint32_t t = task0; t < ntasks; t += W) {
            const uint4 d = tl.desc[t];
            run(d.x, d.y);
        }
    } else {  // persistent too: tasks of `chunk` items, wave_id + W, ...
        const uint32_t nitems = a.nitems;
        const uint32_t ntasks = (uint32_t)(((uint64_t)nitems + a.chunk - 1) / a.chunk);
        const uint32_t W = gridDim.x * nw;
        for (uint32_t t = task0; t < ntasks; t += W) {
            const uint32_t i0 = t * a.chunk;
            run(i0, nitems - i0 < a.chunk ? nitems : i0 + a.chunk);
        }
    }
    flush();
    for (int off = 32; off > 0; off >>= 1) {
        acc_rec += __shfl_xor(acc_rec, off);
        acc_merge += __shfl_xor(acc_merge, off);
    }
    if (lane == 0 && acc_rec) stripe_stats(a, acc_merge, acc_rec);
}

// ---------------------------------------------------------------------------
// Directed support count on the COO+CSR layout, window-staged.
//
// The triangles that close at entry e = (u, v) are the w of u's row after e
// that v's row holds.  u's row is the shorter side: over R-MAT graphs its
// entries after e number about a third of v's window [first entry of u's
// row after e, last entry of u's row] (R-MAT 20: 1.23G against 3.89G over
// all entries), so v's window is staged in the wave's LDS slice and u's
// entries are searched in it, instead of v's window entries being searched
// in u's row.  Lanes load 64 items at once and narrow each v-row to its
// window by two binary searches; the wave then takes the items in order:
// the window's loads and the first 64 candidates' load are issued together,
// the window is stored to LDS and each lane binary-searches its candidate.
// Credits: (u, w) and (u, v) to 16-bit LDS counters over u's row (flushed
// once per row), (v, w) one global atomic.  A window larger than the slice
// is counted the other way round, its entries searched in u's row in HBM;
// a row longer than the counters credits in HBM.
//
// BATCH: persistent over a task list of many graphs (as k_count_dir).
constexpr uint32_t kWinBatch = 8;  // window loads per lane in flight (512 entries per round trip)
// The window-staged count wins while the layout and supports (8 bytes per
// entry) stay in the 256 MB last-level cache, the row-staged one beyond it:
// full counts of R-MAT 20 / 21 / 22 (15.7M / 31.7M / 64.2M entries) took
// 12.0 / 42.8 / 127.4 ms window-staged against 17.7 / 42.4 / 106.6 ms.
constexpr uint64_t kWinCacheBytes = 192ull << 20;
inline bool count_windowed(uint64_t entries) { return entries * 8 <= kWinCacheBytes; }
// LDS words per wave: u's row counters (16-bit) and the window
__host__ __device__ constexpr uint32_t win_words(uint32_t cap) { return cap / 2 + cap; }

template <bool BATCH>
__global__ __launch_bounds__(1024) void k_count_win(CountArgs a, TaskList tl) {
    extern __shared__ uint32_t lds[];
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const uint32_t nw = blockDim.x / kWave;
    const uint32_t cap = a.cap_lds;
    uint32_t* s_cnt = lds + wid * win_words(cap);
    uint32_t* s_win = s_cnt + cap / 2;
    const uint32_t task0 = xcd_block(blockIdx.x, gridDim.x) * nw + wid;
    const uint32_t* ptr = a.ptr;
    const uint32_t* dst = a.dst;
    uint32_t* sup = a.sup;
    uint32_t su = 0xFFFFFFFFu, sub = 0, sulen = 0;  // the row whose counters are in LDS
    bool s_in_lds = false;
    unsigned long long acc_merge = 0, acc_rec = 0;

    auto flush = [&]() {
        if (su == 0xFFFFFFFFu || !s_in_lds) return;
        wave_lds_sync();
        for (uint32_t j = lane; j < sulen; j += kWave) {
            const uint32_t c = (s_cnt[j >> 1] >> ((j & 1) << 4)) & 0xFFFFu;
            if (c) atomicAdd(sup + sub + j, c);
        }
    };
    auto stage = [&](uint32_t u) {
        su = u;
        sub = ptr[u];
        sulen = ptr[u + 1] - sub;
        s_in_lds = sulen <= cap;
        if (s_in_lds) {
            wave_lds_sync();
            for (uint32_t j = lane; j < (sulen + 1) / 2; j += kWave) s_cnt[j] = 0;
            wave_lds_sync();
        }
    };
    // credit of u's row slot j (its counter, or HBM)
    auto credit_u = [&](uint32_t j, uint32_t c) {
        if (s_in_lds) lds_credit16_n(s_cnt, j, c);
        else atomicAdd(sup + sub + j, c);
    };

    auto run = [&](uint32_t i0, uint32_t i1) {
        for (uint32_t base = i0; base < i1; base += kWave) {
            const uint32_t my = base + lane;
            uint32_t e = 0, u = 0, v = KT_DEAD, ub = 0, ulen = 0, vb = 0, vlen = 0;
            if (my < i1) {
                e = a.wl ? a.wl[my] : my;
                v = dst[e];
                if (!(v & KT_DEAD)) {
                    u = a.src[e];
                    ub = ptr[u];
                    ulen = ptr[u + 1] - ub;
                    vb = ptr[v];
                    vlen = ptr[v + 1] - vb;
                }
            }
            const bool live = my < i1 && !(v & KT_DEAD);
            if (live) {
                acc_rec += 1;
                acc_merge += (unsigned long long)ulen + vlen;
            }
            uint32_t lo = 0, hi = 0;
            if (live && e + 1 < ub + ulen && vlen > 0) {
                const uint32_t umin = dst[e + 1] & KT_VMASK;
                const uint32_t umax = dst[ub + ulen - 1] & KT_VMASK;
                if (vlen <= 16) {
                    hi = vlen;
                } else {
                    lo = lower_bound_masked(dst + vb, 0, vlen, umin);
                    hi = lower_bound_masked(dst + vb, lo, vlen, umax + 1);
                }
            }
            uint64_t todo = ballot(hi > lo);
            if (!todo) continue;
            // item l's window (its first 512 entries) and its first 64
            // candidates -- u's entries after e, dst[e + 1 ...] -- loaded
            // into registers; the next item's are loaded while this one's
            // candidates are searched
            uint32_t wv[kWinBatch], wc;
            auto load_item = [&](uint32_t l) {
                const uint32_t le = readlane(e, l);
                const uint32_t wbeg = readlane(vb, l) + readlane(lo, l), last = readlane(hi, l) - readlane(lo, l) - 1;
                const uint32_t nr = last / kWave + 1;  // loads that hold window entries
#pragma unroll
                for (uint32_t r = 0; r < kWinBatch; ++r) {
                    const uint32_t j = lane + r * kWave;
                    if (r < nr) wv[r] = dst[wbeg + (j < last ? j : last)];
                }
                const uint32_t na = readlane(ub, l) + readlane(ulen, l) - le - 1;
                wc = lane < na ? dst[le + 1 + lane] : KT_DEAD;
            };
            uint32_t l = (uint32_t)__ffsll((long long)todo) - 1;
            todo &= todo - 1;
            load_item(l);
            while (true) {
                const uint32_t cu = readlane(u, l), ce = readlane(e, l);
                const uint32_t wbeg = readlane(vb, l) + readlane(lo, l);  // v's window in HBM
                const uint32_t b = readlane(hi, l) - readlane(lo, l);
                if (cu != su) {
                    flush();
                    stage(cu);
                }
                const uint32_t q = ce - sub;        // e's slot in u's row
                const uint32_t na = sulen - q - 1;  // u's entries after e (>= 1)
                const bool fits = b <= cap;
                if (fits) {
                    wave_lds_sync();
#pragma unroll
                    for (uint32_t r = 0; r < kWinBatch; ++r) {
                        const uint32_t j = lane + r * kWave;
                        if (j < b) s_win[j] = wv[r];
                    }
                    for (uint32_t j0 = kWinBatch * kWave; j0 < b; j0 += kWinBatch * kWave) {
                        const uint32_t last = b - 1;
#pragma unroll
                        for (uint32_t r = 0; r < kWinBatch; ++r) {
                            const uint32_t j = j0 + lane + r * kWave;
                            wv[r] = dst[wbeg + (j < last ? j : last)];
                        }
#pragma unroll
                        for (uint32_t r = 0; r < kWinBatch; ++r) {
                            const uint32_t j = j0 + lane + r * kWave;
                            if (j < b) s_win[j] = wv[r];
                        }
                    }
                    wave_lds_sync();
                }
                uint32_t w = wc;
                const bool more = todo != 0;
                const uint32_t ln = more ? (uint32_t)__ffsll((long long)todo) - 1 : l;
                if (more) {
                    todo &= todo - 1;
                    load_item(ln);
                }
                uint32_t c = 0;
                if (fits) {
                    for (uint32_t t0 = 0; t0 < na; t0 += kWave) {
                        const uint32_t t = t0 + lane;
                        // the next candidates' load before this search
                        const uint32_t wn = t0 + kWave < na && t + kWave < na ? dst[ce + 1 + t + kWave] : KT_DEAD;
                        const uint32_t p = lower_bound_uniform(s_win, 0, b, w & KT_VMASK);
                        const bool hit = !(w & KT_DEAD) && p < b && s_win[p] == w;
                        if (hit) {
                            credit_u(q + 1 + t, 1u);
                            atomicAdd(sup + wbeg + p, 1u);
                        }
                        c += (uint32_t)__popcll(ballot(hit));
                        w = wn;
                    }
                } else {
                    // a window beyond the slice: its entries searched in u's row
                    for (uint32_t j0 = 0; j0 < b; j0 += kWave) {
                        const uint32_t j = j0 + lane;
                        const uint32_t x = j < b ? dst[wbeg + j] : KT_DEAD;
                        const uint32_t p = lower_bound_uniform(dst, sub + q + 1, na, x & KT_VMASK);
                        const bool hit = !(x & KT_DEAD) && p < sub + sulen && dst[p] == x;
                        if (hit) {
                            credit_u(p - sub, 1u);
                            atomicAdd(sup + wbeg + j, 1u);
                        }
                        c += (uint32_t)__popcll(ballot(hit));
                    }
                }
                if (c && lane == 0) credit_u(q, c);  // (u, v)
                if (!more) break;
                l = ln;
            }
        }
    };
    if (BATCH) {
        const uint32_t ntasks = *tl.ntasks;
        const uint32_t W = gridDim.x * nw;
        for (uint32_t t = task0; t < ntasks; t += W) {
            const uint4 d = tl.desc[t];
            run(d.x, d.y);
        }
    } else {
        const uint32_t nitems = a.nitems;
        const uint32_t ntasks = (uint32_t)(((uint64_t)nitems + a.chunk - 1) / a.chunk);
        const uint32_t W = gridDim.x * nw;
        for (uint32_t t = task0; t < ntasks; t += W) {
            const uint32_t i0 = t * a.chunk;
            run(i0, nitems - i0 < a.chunk ? nitems : i0 + a.chunk);
        }
    }
    flush();
    for (int off = 32; off > 0; off >>= 1) {
        acc_rec += __shfl_xor(acc_rec, off);
        acc_merge += __shfl_xor(acc_merge, off);
    }
    if (lane == 0 && acc_rec) stripe_stats(a, acc_merge, acc_rec);
}

// ---------------------------------------------------------------------------
// Streaming kernels: grid-stride, 256 threads.

constexpr uint32_t kBlock = 256;

__device__ __forceinline__ uint32_t gsize() { return gridDim.x * blockDim.x; }
__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

uint32_t grid_for(uint64_t count) {
    uint64_t b = (count + kBlock - 1) / kBlock;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return (uint32_t)b;
}

// Adds the stripe lines to the counters and clears them (one workgroup).
__global__ __launch_bounds__(256) void k_stripes_fold(unsigned long long* __restrict__ st,
                                                      DevCounters* ctr) {
    __shared__ unsigned long long s_m[256], s_r[256];
    unsigned long long m = 0, r = 0;
    for (uint32_t l = threadIdx.x; l < kStripes; l += 256) {
        unsigned long long* q = st + (size_t)l * kStripeWords;
        m += q[0];
        r += q[1];
        q[0] = 0;
        q[1] = 0;
    }
    s_m[threadIdx.x] = m;
    s_r[threadIdx.x] = r;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s_m[threadIdx.x] += s_m[threadIdx.x + off];
            s_r[threadIdx.x] += s_r[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctr->merge_elems += s_m[0];
        ctr->recount_edges += s_r[0];
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t value, uint64_t count) {
    for (uint64_t i = gtid(); i < count; i += gsize()) p[i] = value;
}

__global__ void k_fill_u8(uint8_t* p, uint8_t value, uint64_t count) {
    for (uint64_t i = gtid(); i < count; i += gsize()) p[i] = value;
}

__global__ void k_iota_u32(uint32_t* p, uint64_t count) {
    for (uint64_t i = gtid(); i < count; i += gsize()) p[i] = (uint32_t)i;
}

// clearIterInfoOnDevice's numTriangles memset (info.cu:75), live entries only.
__global__ void k_zero_live(const uint32_t* __restrict__ dst, uint32_t* __restrict__ sup,
                            uint32_t E) {
    for (uint32_t i = gtid(); i < E; i += gsize())
        if (!(dst[i] & KT_DEAD)) sup[i] = 0;
}

__global__ void k_zero_listed(const uint32_t* __restrict__ wl, uint32_t n, uint32_t* __restrict__ sup) {
    for (uint32_t i = gtid(); i < n; i += gsize()) sup[wl[i]] = 0;
}

// Grids of the list-building passes below: a grid-stride loop over at most
// kListGrid workgroups, each of which reserves room in the global list (or
// adds to the global count) once per flush of its LDS list, not once per
// wave: atomics of every wave on one address are served one at a time.
constexpr uint32_t kListGrid = 1024;
constexpr uint32_t kListCap = 2048;  // LDS list entries per workgroup

uint32_t list_grid(uint32_t items) {
    const uint32_t b = (items + kBlock - 1) / kBlock;
    return b < kListGrid ? (b ? b : 1u) : kListGrid;
}

// Copies the workgroup's LDS list to the global list at a reserved offset
// and empties it.  Every thread calls it after a barrier, with no addition to
// the count pending.
__device__ __forceinline__ void list_flush(const uint32_t* s_out, uint32_t* s_cnt, uint32_t* s_base,
                                           uint32_t* out, unsigned int* total) {
    const uint32_t c = *s_cnt;
    if (!c) return;
    if (threadIdx.x == 0) *s_base = atomicAdd(total, c);
    __syncthreads();
    const uint32_t base = *s_base;
    for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) out[base + i] = s_out[i];
    __syncthreads();
    if (threadIdx.x == 0) *s_cnt = 0;
    __syncthreads();
}

// mark_deleted_edges_kernel (kernel_ddc.cu:56-77 and its three twins): an
// entry whose support is below k-2 dies; its truss number is k-1 (when tau is
// given).  The deletions are counted per workgroup in LDS and added once.
__global__ __launch_bounds__(kBlock) void k_mark_deleted(uint32_t* __restrict__ dst,
                                                         const uint32_t* __restrict__ src,
                                                         const uint32_t* __restrict__ eid,
                                                         const uint32_t* __restrict__ sup,
                                                         const uint8_t* __restrict__ state,
                                                         uint8_t* __restrict__ nflag,
                                                         uint32_t* __restrict__ tau,
                                                         const uint32_t* __restrict__ wl,
                                                         uint32_t nitems, uint32_t k, int affected,
                                                         DevCounters* ctr) {
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & (kWave - 1);
    uint32_t mine = 0;
    for (uint32_t i = gtid(); i < nitems; i += gsize()) {
        const uint32_t e = wl ? wl[i] : i;
        const uint32_t v = dst[e];
        if (!(v & KT_DEAD) && (!affected || state[e] == KT_DIRECTLY_AFFECTED) && sup[e] < k - 2) {
            ++mine;
            dst[e] = v | KT_DEAD;
            if (tau) tau[eid[e]] = k - 1;  // no truss numbers kept: tau and eid are null
            if (affected) {
                nflag[src[e]] = KT_DIRECTLY_AFFECTED;
                nflag[v] = KT_DIRECTLY_AFFECTED;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
    if (lane == 0 && mine) atomicAdd(&s_cnt, mine);
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) atomicAdd(&ctr->ndel, s_cnt);
}

// mark_directly_affected_edges_kernel (kernel_ddc.cu:79-98, kernel_duc.cu:79-93).
__global__ void k_mark_directly(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                                uint8_t* __restrict__ state, uint8_t* nflag, uint32_t E,
                                int directed) {
    for (uint32_t e = gtid(); e < E; e += gsize()) {
        uint8_t st = KT_NOT_AFFECTED;
        const uint32_t v = dst[e];
        if (!(v & KT_DEAD)) {
            const uint32_t u = src[e];
            const bool du = nflag[u] == KT_DIRECTLY_AFFECTED;
            const bool dv = nflag[v] == KT_DIRECTLY_AFFECTED;
            if (du || dv) {
                st = KT_DIRECTLY_AFFECTED;
                // the endpoint that is not DIRECTLY becomes INDIRECTLY
                if (directed) {
                    if (!du) nflag[u] = KT_INDIRECTLY_AFFECTED;
                    else if (!dv) nflag[v] = KT_INDIRECTLY_AFFECTED;
                }
            }
        }
        state[e] = st;
    }
}

// mark_indirectly_affected_edges_kernel (kernel_ddc.cu:100-114) fused with the
// worklist build: every DIRECTLY / INDIRECTLY live entry is appended and its
// support zeroed before the next count.  The workgroup lists its entries in
// LDS and reserves room in the worklist once per flush.
__global__ __launch_bounds__(kBlock) void k_build_worklist(const uint32_t* __restrict__ src,
                                                           const uint32_t* __restrict__ dst,
                                                           uint8_t* __restrict__ state,
                                                           const uint8_t* __restrict__ nflag,
                                                           uint32_t* __restrict__ sup,
                                                           uint32_t* __restrict__ wl, uint32_t E,
                                                           int directed, DevCounters* ctr) {
    __shared__ uint32_t s_out[kListCap];
    __shared__ uint32_t s_cnt, s_base;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint64_t below = (1ull << lane) - 1ull;
    // uniform trip count: every thread reaches the barriers
    for (uint32_t base = blockIdx.x * blockDim.x; base < E; base += gsize()) {
        const uint32_t e = base + threadIdx.x;
        bool take = false;
        if (e < E) {
            const uint32_t v = dst[e];
            if (!(v & KT_DEAD)) {
                uint8_t st = state[e];
                if (directed && st != KT_DIRECTLY_AFFECTED) {
                    const uint32_t u = src[e];
                    if (nflag[u] == KT_INDIRECTLY_AFFECTED || nflag[v] == KT_INDIRECTLY_AFFECTED) {
                        st = KT_INDIRECTLY_AFFECTED;
                        state[e] = st;
                    }
                }
                take = st != KT_NOT_AFFECTED;
            }
        }
        const uint64_t m = ballot(take);
        if (m) {
            uint32_t b = 0;
            if (lane == 0) b = atomicAdd(&s_cnt, (uint32_t)__popcll(m));  // LDS
            b = __shfl(b, 0);
            if (take) {
                s_out[b + (uint32_t)__popcll(m & below)] = e;
                sup[e] = 0;
            }
        }
        // the count is read between two barriers, so no wave's next
        // additions can reach it before every wave has read it
        __syncthreads();
        const uint32_t c = s_cnt;
        __syncthreads();
        if (c > kListCap - kBlock) list_flush(s_out, &s_cnt, &s_base, wl, &ctr->wl_size);
    }
    __syncthreads();
    list_flush(s_out, &s_cnt, &s_base, wl, &ctr->wl_size);
}

// Minimum live support.  Grid-stride over at most 256 workgroups, wave
// butterflies, then one atomic per workgroup: the counter sees a few
// hundred atomics instead of one per wave of the whole array.
__global__ void k_min_live(const uint32_t* __restrict__ dst, const uint32_t* __restrict__ sup,
                           uint32_t E, DevCounters* ctr) {
    __shared__ uint32_t s_min[kBlock / kWave];
    uint32_t m = 0xFFFFFFFFu;
    for (uint32_t e = gtid(); e < E; e += gsize())
        if (!(dst[e] & KT_DEAD) && sup[e] < m) m = sup[e];
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
}

// ---------------------------------------------------------------------------
// Stable compaction (removeCOOCSRDeletedEdgesOnDevice, graph.cu:520-546):
// tile counts -> single-workgroup scan of tile counts -> scatter with in-tile
// ballot prefix sums -> row pointers read off the per-position prefix.
constexpr uint32_t kTile = 2048;  // entries per compaction tile (256 threads x 8)

__global__ void k_tile_count(const uint32_t* __restrict__ dst, uint32_t E,
                             uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t s_w[kBlock / kWave];
    const uint32_t t = blockIdx.x;
    const uint32_t base = t * kTile;
    uint32_t c = 0;
    for (uint32_t i = base + threadIdx.x; i < base + kTile && i < E; i += kBlock)
        c += !(dst[i] & KT_DEAD);
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    if ((threadIdx.x & (kWave - 1)) == 0) s_w[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (uint32_t w = 0; w < kBlock / kWave; ++w) s += s_w[w];
        tile_sums[t] = s;
    }
}

// Exclusive scan of nt tile counts in place by one workgroup of 1024 threads;
// the total lands in tile_sums[nt] and ctr->live.
__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t* tile_sums, uint32_t nt,
                                                     DevCounters* ctr) {
    __shared__ uint32_t s_part[1024];
    const uint32_t per = (nt + 1023) / 1024;
    const uint32_t b = threadIdx.x * per;
    uint32_t sum = 0;
    for (uint32_t i = b; i < b + per && i < nt; ++i) sum += tile_sums[i];
    s_part[threadIdx.x] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partials
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint32_t x = threadIdx.x >= off ? s_part[threadIdx.x - off] : 0;
        __syncthreads();
        s_part[threadIdx.x] += x;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? s_part[threadIdx.x - 1] : 0;
    for (uint32_t i = b; i < b + per && i < nt; ++i) {
        const uint32_t c = tile_sums[i];
        tile_sums[i] = run;
        run += c;
    }
    if (threadIdx.x == 1023) {
        tile_sums[nt] = s_part[1023];
        ctr->live = s_part[1023];
    }
}

__global__ void k_scatter(uint32_t E, const uint32_t* __restrict__ src_in,
                          const uint32_t* __restrict__ dst_in, const uint32_t* __restrict__ eid_in,
                          const uint32_t* __restrict__ sup_in, uint32_t* __restrict__ src_out,
                          uint32_t* __restrict__ dst_out, uint32_t* __restrict__ eid_out,
                          uint32_t* __restrict__ sup_out, const uint32_t* __restrict__ tile_sums,
                          uint32_t* __restrict__ xs) {
    __shared__ uint32_t s_w[kBlock / kWave];
    const uint32_t t = blockIdx.x;
    const uint32_t lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    uint32_t run = tile_sums[t];
    for (uint32_t r = 0; r < kTile / kBlock; ++r) {
        const uint32_t i = t * kTile + r * kBlock + threadIdx.x;
        const bool inr = i < E;
        const uint32_t d = inr ? dst_in[i] : KT_DEAD;
        const bool live = inr && !(d & KT_DEAD);
        const uint64_t m = ballot(live);
        if (lane == 0) s_w[wid] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = run;
        for (uint32_t w = 0; w < wid; ++w) before += s_w[w];
        uint32_t round = 0;
        for (uint32_t w = 0; w < kBlock / kWave; ++w) round += s_w[w];
        const uint32_t pos = before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (inr) xs[i] = pos;
        if (live) {
            src_out[pos] = src_in[i];
            dst_out[pos] = d;
            eid_out[pos] = eid_in[i];
            sup_out[pos] = sup_in[i];
        }
        run += round;
        __syncthreads();
    }
}

__global__ void k_rebuild_ptr(const uint32_t* __restrict__ ptr_in, const uint32_t* __restrict__ xs,
                              uint32_t* __restrict__ ptr_out, uint32_t n) {
    for (uint32_t r = gtid(); r <= n; r += gsize()) ptr_out[r] = xs[ptr_in[r]];
}

__global__ void k_set_tail(uint32_t* xs, uint32_t E, const uint32_t* tile_sums, uint32_t nt) {
    if (gtid() == 0) xs[E] = tile_sums[nt];
}

}  // namespace

namespace dev {

hipError_t stripes_fold(unsigned long long* stripes, DevCounters* ctr, hipStream_t s) {
    hipLaunchKernelGGL(k_stripes_fold, dim3(1), dim3(256), 0, s, stripes, ctr);
    return hipGetLastError();
}

hipError_t fill_u32(uint32_t* p, uint32_t value, uint64_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(count)), dim3(kBlock), 0, s, p, value, count);
    return hipGetLastError();
}

hipError_t fill_u8(uint8_t* p, uint8_t value, uint64_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(count)), dim3(kBlock), 0, s, p, value, count);
    return hipGetLastError();
}

hipError_t iota_u32(uint32_t* p, uint64_t count, hipStream_t s) {
    if (!count) return hipSuccess;
    hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(count)), dim3(kBlock), 0, s, p, count);
    return hipGetLastError();
}

// Workgroups of `kernel` (block threads, lds dynamic LDS bytes) resident on
// the whole device at once; cached per (kernel, block, lds).
uint32_t resident_blocks(const void* kernel, uint32_t block, size_t lds) {
    struct Key {
        const void* k;
        uint32_t b;
        size_t l;
        uint32_t v;
    };
    static std::vector<Key> cache;
    static std::mutex mu;
    std::lock_guard<std::mutex> g(mu);
    for (const Key& c : cache)
        if (c.k == kernel && c.b == block && c.l == lds) return c.v;
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, (int)block, lds) != hipSuccess || per < 1)
        per = 1;
    const uint32_t v = (uint32_t)per * (uint32_t)(ncu > 0 ? ncu : 256);
    cache.push_back({kernel, block, lds, v});
    return v;
}

hipError_t count_supports(const CountArgs& a, bool directed, bool tiled, uint32_t algo,
                          uint32_t block, hipStream_t s) {
    if (!a.nitems) return hipSuccess;
    const uint32_t nw = block / kWave;
    const uint64_t tasks = (a.nitems + (uint64_t)a.chunk - 1) / a.chunk;
    const uint32_t blocks = (uint32_t)((tasks + nw - 1) / nw);
    const dim3 grid(blocks, tiled ? a.ytiles : 1);
    if (directed && !tiled && algo != KT_COUNT_ROW) {
        const size_t lds = (size_t)nw * flat_words(a.cap_lds) * sizeof(uint32_t);
        if (algo == KT_COUNT_SEARCH) hipLaunchKernelGGL((k_count_flat<false>), grid, dim3(block), lds, s, a);
        else hipLaunchKernelGGL((k_count_flat<true>), grid, dim3(block), lds, s, a);
        return hipGetLastError();
    }
    if (directed && !tiled) {
        // persistent: as many workgroups as stay resident at once (a
        // workgroup keeps its LDS until its slowest wave ends, so a grid of
        // one wave per task leaves CUs waiting behind hub rows)
        const size_t lds = (size_t)nw * win_words(a.cap_lds) * sizeof(uint32_t);
        if (count_windowed(a.nitems)) {
            const uint32_t res = resident_blocks((const void*)k_count_win<false>, block, lds);
            const uint32_t pb = blocks < res ? blocks : res;
            hipLaunchKernelGGL(k_count_win<false>, dim3(pb ? pb : 1), dim3(block), lds, s, a, TaskList{});
        } else {
            const uint32_t res = resident_blocks((const void*)k_count_dir<false>, block, lds);
            const uint32_t pb = blocks < res ? blocks : res;
            hipLaunchKernelGGL(k_count_dir<false>, dim3(pb ? pb : 1), dim3(block), lds, s, a, TaskList{});
        }
        return hipGetLastError();
    }
    // tiled: grid (task groups, x-tile classes)
    const size_t lds = (size_t)nw * (directed ? 2 : 1) * a.cap_lds * sizeof(uint32_t);
    if (directed) {
        hipLaunchKernelGGL((k_count<true, true>), grid, dim3(block), lds, s, a);
    } else {
        if (tiled) hipLaunchKernelGGL((k_count<false, true>), grid, dim3(block), lds, s, a);
        else hipLaunchKernelGGL((k_count<false, false>), grid, dim3(block), lds, s, a);
    }
    return hipGetLastError();
}

hipError_t tiled_row_totals(const uint32_t* ptr, uint32_t n, uint32_t S, uint32_t Y, uint32_t* out, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_tiled_row_totals, dim3(grid_for(n)), dim3(kBlock), 0, s, ptr, n, S, Y, out);
    return hipGetLastError();
}

hipError_t count_supports_batch(const CountArgs& a, const TaskList& tl, uint32_t block,
                                uint32_t blocks, hipStream_t s) {
    const size_t lds = (size_t)(block / kWave) * win_words(a.cap_lds) * sizeof(uint32_t);
    if (count_windowed(a.nitems)) hipLaunchKernelGGL(k_count_win<true>, dim3(blocks), dim3(block), lds, s, a, tl);
    else hipLaunchKernelGGL(k_count_dir<true>, dim3(blocks), dim3(block), lds, s, a, tl);
    return hipGetLastError();
}

hipError_t zero_live_supports(const uint32_t* dst, uint32_t* sup, uint32_t E, hipStream_t s) {
    if (!E) return hipSuccess;
    hipLaunchKernelGGL(k_zero_live, dim3(grid_for(E)), dim3(kBlock), 0, s, dst, sup, E);
    return hipGetLastError();
}

hipError_t zero_listed_supports(const uint32_t* wl, uint32_t n, uint32_t* sup, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_zero_listed, dim3(grid_for(n)), dim3(kBlock), 0, s, wl, n, sup);
    return hipGetLastError();
}

hipError_t mark_deleted(uint32_t* dst, const uint32_t* src, const uint32_t* eid,
                        const uint32_t* sup, const uint8_t* state, uint8_t* nflag,
                        uint32_t* tau, const uint32_t* wl, uint32_t nitems, uint32_t k,
                        bool affected, DevCounters* ctr, hipStream_t s) {
    if (!nitems) return hipSuccess;
    hipLaunchKernelGGL(k_mark_deleted, dim3(list_grid(nitems)), dim3(kBlock), 0, s, dst, src, eid,
                       sup, state, nflag, tau, wl, nitems, k, affected ? 1 : 0, ctr);
    return hipGetLastError();
}

hipError_t mark_directly(const uint32_t* src, const uint32_t* dst, uint8_t* state, uint8_t* nflag,
                         uint32_t E, bool directed, hipStream_t s) {
    if (!E) return hipSuccess;
    hipLaunchKernelGGL(k_mark_directly, dim3(grid_for(E)), dim3(kBlock), 0, s, src, dst, state,
                       nflag, E, directed ? 1 : 0);
    return hipGetLastError();
}

hipError_t build_worklist(const uint32_t* src, const uint32_t* dst, uint8_t* state,
                          const uint8_t* nflag, uint32_t* sup, uint32_t* wl, uint32_t E,
                          bool directed, DevCounters* ctr, hipStream_t s) {
    if (!E) return hipSuccess;
    hipLaunchKernelGGL(k_build_worklist, dim3(list_grid(E)), dim3(kBlock), 0, s, src, dst, state,
                       nflag, sup, wl, E, directed ? 1 : 0, ctr);
    return hipGetLastError();
}

hipError_t min_live_support(const uint32_t* dst, const uint32_t* sup, uint32_t E,
                            DevCounters* ctr, hipStream_t s) {
    if (!E) return hipSuccess;
    const uint32_t blocks = grid_for(E) < 256 ? grid_for(E) : 256;  // one atomic per workgroup
    hipLaunchKernelGGL(k_min_live, dim3(blocks), dim3(kBlock), 0, s, dst, sup, E, ctr);
    return hipGetLastError();
}

uint32_t compact_tiles(uint32_t E) { return (E + kTile - 1) / kTile; }

uint32_t count_lds_bytes(uint32_t cap_lds, uint32_t block, bool directed, bool tiled,
                         uint32_t algo) {
    const uint32_t nw = block / kWave;
    if (directed && !tiled && algo != KT_COUNT_ROW) return nw * flat_words(cap_lds) * 4u;
    if (directed && !tiled) return nw * win_words(cap_lds) * 4u;
    return nw * (directed ? 2u : 1u) * cap_lds * 4u;
}

hipError_t compact(uint32_t n, uint32_t E, const uint32_t* ptr_in, const uint32_t* src_in,
                   const uint32_t* dst_in, const uint32_t* eid_in, const uint32_t* sup_in,
                   uint32_t* ptr_out, uint32_t* src_out, uint32_t* dst_out, uint32_t* eid_out,
                   uint32_t* sup_out, uint32_t* tile_sums, uint32_t* xs, DevCounters* ctr,
                   hipStream_t s) {
    const uint32_t nt = compact_tiles(E);
    if (nt) {
        hipLaunchKernelGGL(k_tile_count, dim3(nt), dim3(kBlock), 0, s, dst_in, E, tile_sums);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, tile_sums, nt, ctr);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    if (nt) {
        hipLaunchKernelGGL(k_scatter, dim3(nt), dim3(kBlock), 0, s, E, src_in, dst_in, eid_in,
                           sup_in, src_out, dst_out, eid_out, sup_out, tile_sums, xs);
        err = hipGetLastError();
        if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL(k_set_tail, dim3(1), dim3(64), 0, s, xs, E, tile_sums, nt);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_rebuild_ptr, dim3(grid_for((uint64_t)n + 1)), dim3(kBlock), 0, s, ptr_in,
                       xs, ptr_out, n);
    return hipGetLastError();
}

}  // namespace dev
}  // namespace kt
