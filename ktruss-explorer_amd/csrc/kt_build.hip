// kt_build.hip -- gfx950 construction of the stored layouts (the graph's
// preprocessing: degree relabel, orientation, CSR rows, full adjacency).
//
// Stored layouts (sortByDegree + undirected2directedCOO + coo2coocsr,
// graph.cu:152-237 and :372): for the degree orientation, vertices are
// relabelled by ascending degree, ties by id (a stable radix sort of the
// degrees over the ids; the reference's quicksort tie order is not
// result-bearing); the other orientations keep ids.  Every edge becomes
// (min, max) of its ends, and an undirected layout also stores (max, min);
// the entries are sorted by that pair, which gives the rows by source with
// destinations ascending (the tiled formats use the same rows).  The full
// adjacency of an oriented layout (the DELTA lists): every entry (s, d)
// appears in s's list as d and in d's list as s; sorting the 2E
// (vertex, neighbour) pairs gives each vertex's list ascending (in-neighbours,
// which are smaller, then out-neighbours).  Both produce exactly the arrays
// the host builders (kt_host.cpp build_layout / build_adjacency) produce.
// Row starts are lower bounds of each vertex in the sorted keys.
#include <hip/hip_runtime.h>

#include <chrono>

#include <rocprim/device/device_radix_sort.hpp>

#include "kt_internal.h"

namespace kt {
namespace {

constexpr uint32_t kBlock = 256;

uint32_t grid(uint64_t n) {
    const uint64_t b = (n + kBlock - 1) / kBlock;
    return (uint32_t)(b < 65535 ? (b ? b : 1) : 65535);
}

uint32_t bits_for(uint64_t v) {  // bits of v (0 for 0)
    uint32_t b = 0;
    while (b < 64 && (v >> b)) ++b;
    return b;
}

__global__ void k_iota(uint32_t* __restrict__ out, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = i;
}

__global__ void k_invert(const uint32_t* __restrict__ new2old, uint32_t* __restrict__ old2new, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        old2new[new2old[i]] = i;
}

// key of edge e: (min, max) of its relabelled ends; value: e
// edge e's stored entries: (min, max) of its relabelled ends (old2new NULL:
// ids kept), and for an undirected layout also (max, min) at m + e
__global__ void k_orient(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                         const uint32_t* __restrict__ old2new, uint32_t m, int both, uint64_t* __restrict__ keys,
                         uint32_t* __restrict__ vals) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
        const uint32_t x = old2new ? old2new[a[e]] : a[e], y = old2new ? old2new[b[e]] : b[e];
        const uint32_t s = x < y ? x : y, d = x < y ? y : x;
        keys[e] = ((uint64_t)s << 32) | d;
        vals[e] = e;
        if (both) {
            keys[(size_t)m + e] = ((uint64_t)d << 32) | s;
            vals[(size_t)m + e] = e;
        }
    }
}

// Tiled formats: the sort key's high word becomes the entry's block row,
// ((tile(s) * Y + class(tile(d))) * S + s mod S) with tiles of S vertices and
// class(t) = t mod Y, so entries sort by (source tile, destination class,
// source, destination).
__global__ void k_tile_keys(uint64_t* __restrict__ keys, uint32_t E, uint32_t S, uint32_t Y) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < E; p += gridDim.x * blockDim.x) {
        const uint64_t k = keys[p];
        const uint32_t s = (uint32_t)(k >> 32), d = (uint32_t)k;
        const uint32_t row = ((s / S) * Y + (d / S) % Y) * S + s % S;
        keys[p] = ((uint64_t)row << 32) | d;
    }
}

// an edge's position: its one entry in a directed layout, the (max, min)
// copy (the one in the higher row) in an undirected layout
// (tiled: the source is recovered from the block row, S and Y as k_tile_keys)
__global__ void k_unpack_layout(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                uint32_t E, int both, uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                uint32_t* __restrict__ pos, uint32_t S, uint32_t Y) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < E; p += gridDim.x * blockDim.x) {
        const uint64_t k = keys[p];
        const uint32_t hi = (uint32_t)(k >> 32), d = (uint32_t)k;
        const uint32_t s = Y ? (hi / S / Y) * S + hi % S : hi;
        src[p] = s;
        dst[p] = d;
        if (!both || s > d) pos[vals[p]] = p;
    }
}

// starts[v] = first index of sorted keys whose high word is >= v, v <= n
__global__ void k_row_starts(const uint64_t* __restrict__ keys, uint32_t len, uint32_t n,
                             uint32_t* __restrict__ starts) {
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += gridDim.x * blockDim.x) {
        const uint64_t x = (uint64_t)v << 32;
        uint32_t lo = 0, hi = len;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (keys[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        starts[v] = lo;
    }
}

__global__ void k_zero_word(uint32_t* w) {
    if (threadIdx.x == 0) *w = 0;
}

__global__ void k_max_row(const uint32_t* __restrict__ ptr, uint32_t n, uint32_t* max_row) {
    uint32_t m = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
        const uint32_t l = ptr[v + 1] - ptr[v];
        m = l > m ? l : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && m) atomicMax(max_row, m);
}

// two list items per entry p = (s, d): d in s's list, s in d's list
__global__ void k_adj_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint32_t E,
                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < E; p += gridDim.x * blockDim.x) {
        const uint32_t s = src[p], d = dst[p];
        keys[2 * (size_t)p] = ((uint64_t)s << 32) | d;
        keys[2 * (size_t)p + 1] = ((uint64_t)d << 32) | s;
        vals[2 * (size_t)p] = p;
        vals[2 * (size_t)p + 1] = p;
    }
}

// slot j of the sorted items: its neighbour and entry; the entry's slot in its
// source's list (neighbour above the list's vertex) is sa, the other sb
__global__ void k_unpack_adj(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                             uint32_t S, uint32_t* __restrict__ snbr, uint32_t* __restrict__ seid,
                             uint32_t* __restrict__ sa, uint32_t* __restrict__ sb) {
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < S; j += gridDim.x * blockDim.x) {
        const uint64_t k = keys[j];
        const uint32_t x = (uint32_t)(k >> 32), nb = (uint32_t)k, p = vals[j];
        snbr[j] = nb;
        seid[j] = p;
        if (nb > x) sa[p] = j;
        else sb[p] = j;
    }
}

// device buffer released when the scope ends (construction is not timed)
struct Tmp {
    void* p = nullptr;
    ~Tmp() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 1); }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

#define KT_BUILD_TRY(x)                   \
    do {                                  \
        const hipError_t e_ = (x);        \
        if (e_ != hipSuccess) return e_;  \
    } while (0)

// sorts (keys, vals) of len items on their low `bits` bits into (ko, vo)
hipError_t sort_pairs(const uint64_t* ki, uint64_t* ko, const uint32_t* vi, uint32_t* vo, uint32_t len,
                      uint32_t bits, hipStream_t s) {
    size_t bytes = 0;
    KT_BUILD_TRY(rocprim::radix_sort_pairs(nullptr, bytes, ki, ko, vi, vo, len, 0u, bits, s));
    Tmp t;
    KT_BUILD_TRY(t.alloc(bytes));
    KT_BUILD_TRY(rocprim::radix_sort_pairs(t.p, bytes, ki, ko, vi, vo, len, 0u, bits, s));
    return hipStreamSynchronize(s);  // t is released on return
}

}  // namespace

namespace dev {

hipError_t build_layout_device(uint32_t n, uint32_t m, const uint32_t* a, const uint32_t* b, int directedness,
                               const uint32_t* degree, uint32_t max_degree, uint32_t* ptr, uint32_t* src,
                               uint32_t* dst, uint32_t* new2old, uint32_t* pos, uint32_t* max_row,
                               hipStream_t s, double* stage_ms, uint32_t tile_size, uint32_t classes) {
    using clk = std::chrono::steady_clock;
    clk::time_point t0 = clk::now();
    // stage boundary: with stage_ms, wait for the stream and book the wall time
    auto stage = [&](int i) -> hipError_t {
        if (!stage_ms) return hipSuccess;
        const hipError_t e = hipStreamSynchronize(s);
        const clk::time_point t1 = clk::now();
        stage_ms[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        t0 = t1;
        return e;
    };
    const bool by_degree = directedness == KT_DIRECTED_BY_DEGREE, both = directedness == KT_UNDIRECTED;
    const uint64_t E64 = both ? 2ull * m : (uint64_t)m;
    if (E64 >= 0xFFFFFFFFull) return hipErrorInvalidValue;
    const uint32_t E = (uint32_t)E64;
    // rows of the stored layout: vertices, or block rows of a tiled format
    const uint32_t S = classes ? tile_size : 0, Y = classes;
    const uint64_t rows64 = Y ? tiled_rows(n, S, Y) : (uint64_t)n;
    if (rows64 >= 0xFFFFFFFFull || (Y && S == 0)) return hipErrorInvalidValue;
    const uint32_t rows = (uint32_t)rows64;
    // relabel: vertices sorted by degree, stable in id (other orientations
    // keep ids)
    Tmp ids, old2new;
    KT_BUILD_TRY(ids.alloc((size_t)n * 4));
    KT_BUILD_TRY(old2new.alloc((size_t)n * 4));
    hipLaunchKernelGGL(k_iota, dim3(grid(n)), dim3(kBlock), 0, s, by_degree ? ids.as<uint32_t>() : new2old, n);
    KT_BUILD_TRY(hipGetLastError());
    if (by_degree && n) {
        size_t bytes = 0;
        const uint32_t bits = bits_for(max_degree) ? bits_for(max_degree) : 1u;
        Tmp dsorted;
        KT_BUILD_TRY(dsorted.alloc((size_t)n * 4));
        KT_BUILD_TRY(rocprim::radix_sort_pairs(nullptr, bytes, degree, dsorted.as<uint32_t>(), ids.as<uint32_t>(),
                                               new2old, n, 0u, bits, s));
        Tmp t;
        KT_BUILD_TRY(t.alloc(bytes));
        KT_BUILD_TRY(rocprim::radix_sort_pairs(t.p, bytes, degree, dsorted.as<uint32_t>(), ids.as<uint32_t>(),
                                               new2old, n, 0u, bits, s));
        KT_BUILD_TRY(hipStreamSynchronize(s));
    }
    if (by_degree) {
        hipLaunchKernelGGL(k_invert, dim3(grid(n)), dim3(kBlock), 0, s, new2old, old2new.as<uint32_t>(), n);
        KT_BUILD_TRY(hipGetLastError());
    }
    KT_BUILD_TRY(stage(0));
    // entries sorted by (source, destination)
    Tmp k0, k1, v0, v1;
    KT_BUILD_TRY(k0.alloc((size_t)E * 8));
    KT_BUILD_TRY(k1.alloc((size_t)E * 8));
    KT_BUILD_TRY(v0.alloc((size_t)E * 4));
    KT_BUILD_TRY(v1.alloc((size_t)E * 4));
    if (m) {
        hipLaunchKernelGGL(k_orient, dim3(grid(m)), dim3(kBlock), 0, s, a, b,
                           by_degree ? old2new.as<uint32_t>() : (const uint32_t*)nullptr, m, both ? 1 : 0,
                           k0.as<uint64_t>(), v0.as<uint32_t>());
        KT_BUILD_TRY(hipGetLastError());
        if (Y) {
            hipLaunchKernelGGL(k_tile_keys, dim3(grid(E)), dim3(kBlock), 0, s, k0.as<uint64_t>(), E, S, Y);
            KT_BUILD_TRY(hipGetLastError());
        }
    }
    KT_BUILD_TRY(stage(1));
    if (m) {
        KT_BUILD_TRY(sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint32_t>(), v1.as<uint32_t>(), E,
                                32 + (bits_for(rows) ? bits_for(rows) : 1u), s));
        hipLaunchKernelGGL(k_unpack_layout, dim3(grid(E)), dim3(kBlock), 0, s, k1.as<uint64_t>(),
                           v1.as<uint32_t>(), E, both ? 1 : 0, src, dst, pos, S, Y);
        KT_BUILD_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_row_starts, dim3(grid((uint64_t)rows + 1)), dim3(kBlock), 0, s, k1.as<uint64_t>(), E,
                       rows, ptr);
    KT_BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(64), 0, s, max_row);
    KT_BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_max_row, dim3(grid(rows)), dim3(kBlock), 0, s, ptr, rows, max_row);
    KT_BUILD_TRY(hipGetLastError());
    KT_BUILD_TRY(stage(2));
    return hipStreamSynchronize(s);
}

hipError_t build_adjacency_device(uint32_t n, uint32_t E, const uint32_t* src, const uint32_t* dst, uint32_t* sptr,
                                  uint32_t* snbr, uint32_t* seid, uint32_t* sa, uint32_t* sb, hipStream_t s) {
    const uint32_t S = 2 * E;
    Tmp k0, k1, v0, v1;
    KT_BUILD_TRY(k0.alloc((size_t)S * 8));
    KT_BUILD_TRY(k1.alloc((size_t)S * 8));
    KT_BUILD_TRY(v0.alloc((size_t)S * 4));
    KT_BUILD_TRY(v1.alloc((size_t)S * 4));
    if (E) {
        hipLaunchKernelGGL(k_adj_keys, dim3(grid(E)), dim3(kBlock), 0, s, src, dst, E, k0.as<uint64_t>(),
                           v0.as<uint32_t>());
        KT_BUILD_TRY(hipGetLastError());
        KT_BUILD_TRY(sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint32_t>(), v1.as<uint32_t>(), S,
                                32 + (bits_for(n) ? bits_for(n) : 1u), s));
        hipLaunchKernelGGL(k_unpack_adj, dim3(grid(S)), dim3(kBlock), 0, s, k1.as<uint64_t>(), v1.as<uint32_t>(),
                           S, snbr, seid, sa, sb);
        KT_BUILD_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_row_starts, dim3(grid((uint64_t)n + 1)), dim3(kBlock), 0, s, k1.as<uint64_t>(), S, n,
                       sptr);
    KT_BUILD_TRY(hipGetLastError());
    return hipStreamSynchronize(s);
}

}  // namespace dev
}  // namespace kt

// ---------------------------------------------------------------------------
// Input validation (kt_graph_create) on the device: range and self-loop
// checks (graph.cu:54), then the entries sorted by (min, max, direction): a
// symmetric input is valid when every (min, max) pair holds exactly one
// entry of each direction (graph.cu:52-53 asserts symmetry), a pair input
// when no pair repeats.  A valid input's unique edges are then the sorted
// pairs in order, so edge ids are sorted positions (halved for a symmetric
// input).  Errors report the lowest offending input entry.
namespace kt {
namespace {

constexpr uint32_t kNoError = 0xFFFFFFFFu;

__global__ void k_check_entries(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint32_t m,
                                uint32_t n, uint32_t* bad_range, uint32_t* bad_loop) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const uint32_t s = src[i], d = dst[i];
        if (s >= n || d >= n) atomicMin(bad_range, i);
        else if (s == d) atomicMin(bad_loop, i);
    }
}

// key: min << 33 | max << 1 | (src > dst); value: the entry
__global__ void k_pair_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint32_t m,
                            uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) {
        const uint32_t s = src[i], d = dst[i];
        const uint32_t a = s < d ? s : d, b = s < d ? d : s;
        keys[i] = ((uint64_t)a << 33) | ((uint64_t)b << 1) | (s > d ? 1u : 0u);
        vals[i] = i;
    }
}

// sym: position p starts a pair (even p) that must be (p, p + 1) with
// directions 0 then 1; a single entry is asymmetric, anything else a
// duplicate.  pairs: p repeating p - 1's pair is a duplicate.
__global__ void k_check_pairs(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t m,
                              int sym, uint32_t* bad_asym, uint32_t* bad_dup) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < m; p += gridDim.x * blockDim.x) {
        const uint64_t pk = keys[p] >> 1;
        const bool start = p == 0 || (keys[p - 1] >> 1) != pk;
        if (sym) {
            if (!start) continue;
            const bool one = p + 1 >= m || (keys[p + 1] >> 1) != pk;
            if (one) {
                atomicMin(bad_asym, vals[p]);
                continue;
            }
            const bool two = p + 2 >= m || (keys[p + 2] >> 1) != pk;
            if (!two || (keys[p] & 1) != 0 || (keys[p + 1] & 1) != 1) atomicMin(bad_dup, vals[p]);
        } else if (!start) {
            atomicMin(bad_dup, vals[p - 1] < vals[p] ? vals[p - 1] : vals[p]);
        }
    }
}

// unique edge q = p >> shift: its ends, and every input entry's edge id
__global__ void k_unique_edges(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint32_t m,
                               uint32_t shift, uint32_t* __restrict__ ua, uint32_t* __restrict__ ub,
                               uint32_t* __restrict__ uid, uint32_t* __restrict__ degree) {
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < m; p += gridDim.x * blockDim.x) {
        const uint32_t q = p >> shift;
        uid[vals[p]] = q;
        if ((p & ((1u << shift) - 1)) == 0) {
            const uint64_t k = keys[p];
            const uint32_t a = (uint32_t)(k >> 33), b = (uint32_t)((k >> 1) & 0xFFFFFFFFu);
            ua[q] = a;
            ub[q] = b;
            atomicAdd(degree + a, 1u);
            atomicAdd(degree + b, 1u);
        }
    }
}

__global__ void k_max_u32(const uint32_t* __restrict__ x, uint32_t n, uint32_t* out) {
    uint32_t m = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        m = x[v] > m ? x[v] : m;
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

}  // namespace

namespace dev {

hipError_t validate_input(uint32_t n, uint32_t m, const uint32_t* src, const uint32_t* dst, bool sym,
                          InputCheck& res, std::vector<uint32_t>& ua, std::vector<uint32_t>& ub,
                          std::vector<uint32_t>& uid, std::vector<uint32_t>& degree, uint32_t& max_degree) {
    hipStream_t s = nullptr;
    KT_BUILD_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    Tmp ds, dd, flags;
    KT_BUILD_TRY(ds.alloc((size_t)m * 4));
    KT_BUILD_TRY(dd.alloc((size_t)m * 4));
    KT_BUILD_TRY(flags.alloc(5 * 4));
    uint32_t* f = flags.as<uint32_t>();  // range, loop, asymmetric, duplicate, max degree
    KT_BUILD_TRY(hipMemcpyAsync(ds.p, src, (size_t)m * 4, hipMemcpyHostToDevice, s));
    KT_BUILD_TRY(hipMemcpyAsync(dd.p, dst, (size_t)m * 4, hipMemcpyHostToDevice, s));
    KT_BUILD_TRY(hipMemsetAsync(f, 0xFF, 4 * 4, s));
    KT_BUILD_TRY(hipMemsetAsync(f + 4, 0, 4, s));
    hipLaunchKernelGGL(k_check_entries, dim3(grid(m)), dim3(kBlock), 0, s, ds.as<uint32_t>(), dd.as<uint32_t>(), m,
                       n, f, f + 1);
    KT_BUILD_TRY(hipGetLastError());
    uint32_t hf[5];
    KT_BUILD_TRY(hipMemcpyAsync(hf, f, 2 * 4, hipMemcpyDeviceToHost, s));
    KT_BUILD_TRY(hipStreamSynchronize(s));
    res = InputCheck{};
    if (hf[0] != kNoError || hf[1] != kNoError) {  // the lowest entry that is wrong
        const bool range = hf[0] <= hf[1];
        res.status = range ? KT_ERR_RANGE : KT_ERR_SELF_LOOP;
        res.at = range ? hf[0] : hf[1];
        return hipSuccess;
    }
    Tmp k0, k1, v0, v1;
    KT_BUILD_TRY(k0.alloc((size_t)m * 8));
    KT_BUILD_TRY(k1.alloc((size_t)m * 8));
    KT_BUILD_TRY(v0.alloc((size_t)m * 4));
    KT_BUILD_TRY(v1.alloc((size_t)m * 4));
    hipLaunchKernelGGL(k_pair_keys, dim3(grid(m)), dim3(kBlock), 0, s, ds.as<uint32_t>(), dd.as<uint32_t>(), m,
                       k0.as<uint64_t>(), v0.as<uint32_t>());
    KT_BUILD_TRY(hipGetLastError());
    KT_BUILD_TRY(sort_pairs(k0.as<uint64_t>(), k1.as<uint64_t>(), v0.as<uint32_t>(), v1.as<uint32_t>(), m,
                            33 + (bits_for(n) ? bits_for(n) : 1u), s));
    hipLaunchKernelGGL(k_check_pairs, dim3(grid(m)), dim3(kBlock), 0, s, k1.as<uint64_t>(), v1.as<uint32_t>(), m,
                       sym ? 1 : 0, f + 2, f + 3);
    KT_BUILD_TRY(hipGetLastError());
    KT_BUILD_TRY(hipMemcpyAsync(hf + 2, f + 2, 2 * 4, hipMemcpyDeviceToHost, s));
    KT_BUILD_TRY(hipStreamSynchronize(s));
    if (hf[2] != kNoError || hf[3] != kNoError) {
        const bool asym = hf[2] <= hf[3];
        res.status = asym ? KT_ERR_ASYMMETRIC : KT_ERR_DUPLICATE;
        res.at = asym ? hf[2] : hf[3];
        return hipSuccess;
    }
    const uint32_t shift = sym ? 1u : 0u;
    const uint32_t mu = m >> shift;
    Tmp dua, dub, duid, ddeg;
    KT_BUILD_TRY(dua.alloc((size_t)mu * 4));
    KT_BUILD_TRY(dub.alloc((size_t)mu * 4));
    KT_BUILD_TRY(duid.alloc((size_t)m * 4));
    KT_BUILD_TRY(ddeg.alloc((size_t)n * 4));
    KT_BUILD_TRY(hipMemsetAsync(ddeg.p, 0, (size_t)n * 4, s));
    hipLaunchKernelGGL(k_unique_edges, dim3(grid(m)), dim3(kBlock), 0, s, k1.as<uint64_t>(), v1.as<uint32_t>(), m,
                       shift, dua.as<uint32_t>(), dub.as<uint32_t>(), duid.as<uint32_t>(), ddeg.as<uint32_t>());
    KT_BUILD_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_max_u32, dim3(grid(n)), dim3(kBlock), 0, s, ddeg.as<uint32_t>(), n, f + 4);
    KT_BUILD_TRY(hipGetLastError());
    ua.resize(mu);
    ub.resize(mu);
    uid.resize(m);
    degree.resize(n);
    KT_BUILD_TRY(hipMemcpyAsync(ua.data(), dua.p, (size_t)mu * 4, hipMemcpyDeviceToHost, s));
    KT_BUILD_TRY(hipMemcpyAsync(ub.data(), dub.p, (size_t)mu * 4, hipMemcpyDeviceToHost, s));
    KT_BUILD_TRY(hipMemcpyAsync(uid.data(), duid.p, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    KT_BUILD_TRY(hipMemcpyAsync(degree.data(), ddeg.p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    KT_BUILD_TRY(hipMemcpyAsync(&max_degree, f + 4, 4, hipMemcpyDeviceToHost, s));
    return hipStreamSynchronize(s);
}

}  // namespace dev
}  // namespace kt
