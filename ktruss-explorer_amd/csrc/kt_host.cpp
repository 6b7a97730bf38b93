// kt_host.cpp -- host-side graph handling of the k-truss engine: input
// validation and edge numbering, TSV I/O, the edge-set comparator and the
// seeded synthetic generators (R-MAT, Chung-Lu with planted cliques,
// Erdos-Renyi).  Parallel loops use std::thread; every result is independent
// of the thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_set>
#include <vector>

#include "kt_host.h"
#include "kt_internal.h"

namespace kt {

// ---------------------------------------------------------------------------
// Errors

namespace {
thread_local std::string g_last_error;
}

void set_error(int status, const std::string& msg) {
    g_last_error = std::string(kt_status_string(status)) + ": " + msg;
}

int fail(int status, const std::string& msg) {
    set_error(status, msg);
    return status;
}

const char* last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Parallel helpers

unsigned num_threads() {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    if (t > 16) t = 16;
    if (const char* env = std::getenv("KT_HOST_THREADS")) {
        const int v = std::atoi(env);
        if (v > 0) t = (unsigned)v;
    }
    return t;
}

namespace {
thread_local bool tl_worker = false;  // nested parallel_for calls run inline
}

void parallel_for(size_t n, size_t grain, const std::function<void(size_t, size_t)>& body) {
    if (n == 0) return;
    size_t T = tl_worker ? 1 : num_threads();
    if (grain == 0) grain = 1;
    T = std::min(T, (n + grain - 1) / grain);
    if (T <= 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T);
    for (size_t t = 0; t < T; ++t) {
        const size_t b = n * t / T, e = n * (t + 1) / T;
        th.emplace_back([&body, b, e] {
            tl_worker = true;
            body(b, e);
        });
    }
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// Bucketing by the smaller endpoint.  Entries are grouped by min(u, v) with a
// counting sort and each bucket sorted by (max(u, v), direction); this is the
// COO -> CSR conversion of coo2coocsr (graph.cu:372-412) with the per-row
// sort done on (key, payload) pairs.

namespace {

struct Bucketed {
    std::vector<uint64_t> off;     // n + 1
    std::vector<uint64_t> key;     // (max << 1 | dir), bucket-sorted
    std::vector<uint64_t> idx;     // originating entry
};

void bucket_by_min(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                   Bucketed& B) {
    B.off.assign((size_t)n + 1, 0);
    for (uint64_t i = 0; i < m; ++i) {
        const uint32_t a = std::min(src[i], dst[i]);
        B.off[a + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) B.off[v + 1] += B.off[v];
    B.key.resize(m);
    B.idx.resize(m);
    std::vector<uint64_t> cur(B.off.begin(), B.off.end() - 1);
    for (uint64_t i = 0; i < m; ++i) {
        const uint32_t s = src[i], d = dst[i];
        const uint32_t a = std::min(s, d), b = std::max(s, d);
        const uint64_t p = cur[a]++;
        B.key[p] = ((uint64_t)b << 1) | (s > d ? 1u : 0u);
        B.idx[p] = i;
    }
    parallel_for(n, 4096, [&](size_t v0, size_t v1) {
        std::vector<std::pair<uint64_t, uint64_t>> tmp;
        for (size_t v = v0; v < v1; ++v) {
            const uint64_t b = B.off[v], e = B.off[v + 1];
            if (e - b < 2) continue;
            tmp.resize(e - b);
            for (uint64_t p = b; p < e; ++p) tmp[p - b] = {B.key[p], B.idx[p]};
            std::sort(tmp.begin(), tmp.end());
            for (uint64_t p = b; p < e; ++p) {
                B.key[p] = tmp[p - b].first;
                B.idx[p] = tmp[p - b].second;
            }
        }
    });
}

}  // namespace

// inputs of at least this many entries are validated on the device
constexpr uint64_t kDeviceValidateMin = 1ull << 16;

int graph_create(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, uint32_t flags,
                 kt_graph** out) {
    if (!out) return fail(KT_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (m && (!src || !dst)) return fail(KT_ERR_ARG, "src/dst NULL");
    if (n >= (1u << 31)) return fail(KT_ERR_RANGE, "n must be below 2^31");
    if (flags != KT_INPUT_SYMMETRIC && flags != KT_INPUT_PAIRS)
        return fail(KT_ERR_ARG, "unknown input flags");
    if (m >= (1ull << 32)) return fail(KT_ERR_ARG, "too many entries");
    // large inputs are validated on the device once it is initialised
    // (kt_build.hip: the same checks and the same edges, ids and degrees);
    // ensure_device makes it current on this thread
    if (m >= kDeviceValidateMin && device_initialized() && ensure_device() == KT_OK) {
        auto* g = new (std::nothrow) kt_graph();
        if (!g) return fail(KT_ERR_NOMEM, "graph");
        dev::InputCheck chk;
        const hipError_t e = dev::validate_input(n, (uint32_t)m, src, dst, flags == KT_INPUT_SYMMETRIC, chk, g->ua,
                                                 g->ub, g->uid, g->degree, g->max_degree);
        if (e != hipSuccess) {
            delete g;
            return fail(KT_ERR_HIP, std::string("input validation: ") + hipGetErrorString(e));
        }
        if (chk.status != KT_OK) {
            const uint64_t i = chk.at;
            char buf[160];
            if (chk.status == KT_ERR_RANGE || chk.status == KT_ERR_SELF_LOOP)
                std::snprintf(buf, sizeof buf, "entry %llu = (%u, %u) with n = %u", (unsigned long long)i, src[i],
                              dst[i], n);
            else
                std::snprintf(buf, sizeof buf, "entry %llu = (%u, %u)%s", (unsigned long long)i, src[i], dst[i],
                              chk.status == KT_ERR_ASYMMETRIC ? " has no reverse entry" : " repeats");
            delete g;
            return fail(chk.status, buf);
        }
        g->n = n;
        g->m_input = m;
        g->m = g->ua.size();
        *out = g;
        return KT_OK;
    }
    // range and self-loop checks (graph.cu:54); like the device checks, the
    // lowest offending entry is reported (a range error before a self-loop at
    // the same entry)
    std::atomic<uint64_t> bad_range{~0ull}, bad_loop{~0ull};
    auto atomic_min = [](std::atomic<uint64_t>& a, uint64_t v) {
        uint64_t cur = a.load();
        while (v < cur && !a.compare_exchange_weak(cur, v)) {
        }
    };
    parallel_for(m, 1 << 16, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            if (src[i] >= n || dst[i] >= n) {
                atomic_min(bad_range, i);
                return;
            }
            if (src[i] == dst[i]) {
                atomic_min(bad_loop, i);
                return;
            }
        }
    });
    if (bad_range.load() != ~0ull || bad_loop.load() != ~0ull) {
        const bool range = bad_range.load() <= bad_loop.load();
        const uint64_t i = range ? bad_range.load() : bad_loop.load();
        char buf[160];
        std::snprintf(buf, sizeof buf, "entry %llu = (%u, %u) with n = %u",
                      (unsigned long long)i, src[i], dst[i], n);
        return fail(range ? KT_ERR_RANGE : KT_ERR_SELF_LOOP, buf);
    }
    Bucketed B;
    bucket_by_min(n, m, src, dst, B);
    auto* g = new (std::nothrow) kt_graph();
    if (!g) return fail(KT_ERR_NOMEM, "graph");
    g->n = n;
    g->m_input = m;
    g->uid.resize(m);
    // per bucket: validate and count unique edges
    const bool sym = flags == KT_INPUT_SYMMETRIC;
    std::vector<uint64_t> nuniq((size_t)n + 1, 0);
    // every pair group (same (min, max)) is checked; the reported entry is
    // the lowest over the offending groups, each group naming its first entry
    // in (direction, entry) order (a repeated pair of a pair input: its lowest
    // entry), as the device checks do
    std::atomic<uint64_t> bad_asym{~0ull}, bad_dup{~0ull};
    parallel_for(n, 4096, [&](size_t v0, size_t v1) {
        for (size_t v = v0; v < v1; ++v) {
            const uint64_t b = B.off[v], e = B.off[v + 1];
            uint64_t c = 0;
            for (uint64_t p = b; p < e;) {
                const uint64_t hi = B.key[p] >> 1;
                uint64_t q = p;
                while (q < e && (B.key[q] >> 1) == hi) ++q;
                const uint64_t cnt = q - p;
                if (sym) {
                    if (cnt == 2 && (B.key[p] & 1) == 0 && (B.key[p + 1] & 1) == 1) ++c;
                    else atomic_min(cnt == 1 ? bad_asym : bad_dup, B.idx[p]);
                } else if (cnt == 1) {
                    ++c;
                } else {
                    uint64_t lo = B.idx[p];
                    for (uint64_t r = p + 1; r < q; ++r) lo = std::min<uint64_t>(lo, B.idx[r]);
                    atomic_min(bad_dup, lo);
                }
                p = q;
            }
            nuniq[v + 1] = c;
        }
    });
    if (bad_asym.load() != ~0ull || bad_dup.load() != ~0ull) {
        const bool asym = bad_asym.load() <= bad_dup.load();
        const uint64_t i = asym ? bad_asym.load() : bad_dup.load();
        char buf[160];
        std::snprintf(buf, sizeof buf, "entry %llu = (%u, %u)%s", (unsigned long long)i, src[i],
                      dst[i], asym ? " has no reverse entry" : " repeats");
        delete g;
        return fail(asym ? KT_ERR_ASYMMETRIC : KT_ERR_DUPLICATE, buf);
    }
    for (uint32_t v = 0; v < n; ++v) nuniq[v + 1] += nuniq[v];
    g->m = nuniq[n];
    g->ua.resize(g->m);
    g->ub.resize(g->m);
    parallel_for(n, 4096, [&](size_t v0, size_t v1) {
        for (size_t v = v0; v < v1; ++v) {
            uint64_t id = nuniq[v];
            const uint64_t b = B.off[v], e = B.off[v + 1];
            const uint64_t step = sym ? 2 : 1;
            for (uint64_t p = b; p < e; p += step) {
                g->ua[id] = (uint32_t)v;
                g->ub[id] = (uint32_t)(B.key[p] >> 1);
                g->uid[B.idx[p]] = (uint32_t)id;
                if (sym) g->uid[B.idx[p + 1]] = (uint32_t)id;
                ++id;
            }
        }
    });
    g->degree.assign(n, 0);
    for (uint64_t e = 0; e < g->m; ++e) {
        g->degree[g->ua[e]]++;
        g->degree[g->ub[e]]++;
    }
    g->max_degree = 0;
    for (uint32_t v = 0; v < n; ++v) g->max_degree = std::max(g->max_degree, g->degree[v]);
    *out = g;
    return KT_OK;
}

// ---------------------------------------------------------------------------
// Stored layout (host).  directed-by-degree relabels by ascending degree with
// ties broken by id (sortByDegree, graph.cu:152-185, uses an unstable
// quicksort: its tie order is not result-bearing); directed variants keep
// (min, max) of the relabelled pair (undirected2directedCOO, graph.cu:194-206);
// undirected keeps both directions.  Rows sorted by destination.

int build_layout(const kt_graph* g, int directedness, HostLayout& L) {
    const uint32_t n = g->n;
    L.n = n;
    L.directed = directedness != KT_UNDIRECTED;
    L.old2new.resize(n);
    L.new2old.resize(n);
    if (directedness == KT_DIRECTED_BY_DEGREE) {
        // counting sort by degree, stable in id
        std::vector<uint64_t> cnt((size_t)g->max_degree + 2, 0);
        for (uint32_t v = 0; v < n; ++v) cnt[g->degree[v] + 1]++;
        for (size_t d = 0; d + 1 < cnt.size(); ++d) cnt[d + 1] += cnt[d];
        for (uint32_t v = 0; v < n; ++v) {
            const uint32_t nv = (uint32_t)cnt[g->degree[v]]++;
            L.old2new[v] = nv;
            L.new2old[nv] = v;
        }
    } else {
        for (uint32_t v = 0; v < n; ++v) L.old2new[v] = L.new2old[v] = v;
    }
    const uint64_t m = g->m;
    const uint64_t E = L.directed ? m : 2 * m;
    if (E >= (1ull << 32) - 1) return fail(KT_ERR_ARG, "too many stored entries");
    L.E = (uint32_t)E;
    L.ptr.assign((size_t)n + 1, 0);
    auto endpoint = [&](uint64_t e, uint32_t& s, uint32_t& d) {
        uint32_t a = L.old2new[g->ua[e]], b = L.old2new[g->ub[e]];
        if (a > b) std::swap(a, b);
        s = a;
        d = b;
    };
    for (uint64_t e = 0; e < m; ++e) {
        uint32_t s, d;
        endpoint(e, s, d);
        L.ptr[s + 1]++;
        if (!L.directed) L.ptr[d + 1]++;
    }
    for (uint32_t v = 0; v < n; ++v) L.ptr[v + 1] += L.ptr[v];
    L.src.resize(E);
    L.dst.resize(E);
    std::vector<uint32_t> id(E);
    {
        std::vector<uint32_t> cur(L.ptr.begin(), L.ptr.end() - 1);
        for (uint64_t e = 0; e < m; ++e) {
            uint32_t s, d;
            endpoint(e, s, d);
            uint32_t p = cur[s]++;
            L.src[p] = s;
            L.dst[p] = d;
            id[p] = (uint32_t)e;
            if (!L.directed) {
                p = cur[d]++;
                L.src[p] = d;
                L.dst[p] = s;
                id[p] = (uint32_t)e;
            }
        }
    }
    parallel_for(n, 2048, [&](size_t v0, size_t v1) {
        std::vector<std::pair<uint32_t, uint32_t>> tmp;
        for (size_t v = v0; v < v1; ++v) {
            const uint32_t b = L.ptr[v], e = L.ptr[v + 1];
            if (e - b < 2) continue;
            tmp.resize(e - b);
            for (uint32_t p = b; p < e; ++p) tmp[p - b] = {L.dst[p], id[p]};
            std::sort(tmp.begin(), tmp.end());
            for (uint32_t p = b; p < e; ++p) {
                L.dst[p] = tmp[p - b].first;
                id[p] = tmp[p - b].second;
            }
        }
    });
    L.pos_of_edge.resize(m);
    for (uint32_t p = 0; p < (uint32_t)E; ++p) L.pos_of_edge[id[p]] = p;
    L.max_row = 0;
    for (uint32_t v = 0; v < n; ++v) L.max_row = std::max(L.max_row, L.ptr[v + 1] - L.ptr[v]);
    return KT_OK;
}

// Full adjacency of an oriented layout (every stored entry e = (u, v) has
// u < v): vertex x's list is its in-neighbours (ascending, since entries are
// stored in source order) followed by its out-neighbours (ascending), which is
// one sorted list.  seid maps a list slot to its entry, sa / sb map an entry
// to its slot in u's / v's list.
int build_adjacency(uint32_t n, uint32_t E, const uint32_t* ptr, const uint32_t* src,
                    const uint32_t* dst, std::vector<uint32_t>& sptr, std::vector<uint32_t>& snbr,
                    std::vector<uint32_t>& seid, std::vector<uint32_t>& sa,
                    std::vector<uint32_t>& sb) {
    std::vector<uint32_t> indeg(n, 0);
    for (uint32_t e = 0; e < E; ++e) {
        if (src[e] >= dst[e]) return fail(KT_ERR_STATE, "adjacency lists need an oriented layout");
        indeg[dst[e]]++;
    }
    sptr.assign((size_t)n + 1, 0);
    for (uint32_t x = 0; x < n; ++x) sptr[x + 1] = sptr[x] + indeg[x] + (ptr[x + 1] - ptr[x]);
    snbr.resize(2 * (size_t)E);
    seid.resize(2 * (size_t)E);
    sa.resize(E);
    sb.resize(E);
    std::vector<uint32_t> fill(sptr.begin(), sptr.end() - 1);
    for (uint32_t e = 0; e < E; ++e) {  // in-neighbours, in source order
        const uint32_t x = dst[e];
        const uint32_t p = fill[x]++;
        snbr[p] = src[e];
        seid[p] = e;
        sb[e] = p;
    }
    parallel_for(n, 4096, [&](size_t x0, size_t x1) {
        for (size_t x = x0; x < x1; ++x) {
            const uint32_t base = sptr[x] + indeg[x];
            for (uint32_t e = ptr[x]; e < ptr[x + 1]; ++e) {
                const uint32_t p = base + (e - ptr[x]);
                snbr[p] = dst[e];
                seid[p] = e;
                sa[e] = p;
            }
        }
    });
    return KT_OK;
}

int verify_edges(uint64_t ma, const uint32_t* sa, const uint32_t* da, uint64_t mb,
                 const uint32_t* sb, const uint32_t* db) {
    if (ma != mb) return 1;
    std::vector<uint64_t> a(ma), b(mb);
    for (uint64_t i = 0; i < ma; ++i) a[i] = ((uint64_t)sa[i] << 32) | da[i];
    for (uint64_t i = 0; i < mb; ++i) b[i] = ((uint64_t)sb[i] << 32) | db[i];
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    return a == b ? 0 : 1;
}

// ---------------------------------------------------------------------------
// Random numbers: splitmix64 (seeding) and xoshiro256** (streams), both from
// Blackman & Vigna's published definitions.

namespace {

inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t x = seed;
        for (auto& w : s) w = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    inline double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    inline uint64_t below(uint64_t bound) {  // Lemire's multiply-shift, bias < 2^-64 * bound
        return (uint64_t)(((__uint128_t)next() * bound) >> 64);
    }
};

// Stream seed for block `blk` of a generator run.
inline uint64_t block_seed(uint64_t seed, uint64_t salt, uint64_t blk) {
    uint64_t x = seed ^ (salt * 0xD1B54A32D192ED03ull);
    splitmix64(x);
    x ^= blk * 0x9E3779B97F4A7C15ull;
    return splitmix64(x);
}

constexpr uint64_t kGenBlock = 1 << 15;  // edges per RNG stream

// Sorted unique pairs (a < b) from raw endpoint lists; self-loops dropped.
void clean_pairs(uint32_t n, std::vector<uint32_t>& a, std::vector<uint32_t>& b,
                 std::vector<uint32_t>& ua, std::vector<uint32_t>& ub) {
    const uint64_t m = a.size();
    std::vector<uint64_t> off((size_t)n + 1, 0);
    for (uint64_t i = 0; i < m; ++i)
        if (a[i] != b[i]) off[std::min(a[i], b[i]) + 1]++;
    for (uint32_t v = 0; v < n; ++v) off[v + 1] += off[v];
    std::vector<uint32_t> hi(off[n]);
    {
        std::vector<uint64_t> cur(off.begin(), off.end() - 1);
        for (uint64_t i = 0; i < m; ++i)
            if (a[i] != b[i]) hi[cur[std::min(a[i], b[i])]++] = std::max(a[i], b[i]);
    }
    std::vector<uint32_t>().swap(a);
    std::vector<uint32_t>().swap(b);
    std::vector<uint64_t> keep((size_t)n + 1, 0);
    parallel_for(n, 4096, [&](size_t v0, size_t v1) {
        for (size_t v = v0; v < v1; ++v) {
            auto first = hi.begin() + (ptrdiff_t)off[v], last = hi.begin() + (ptrdiff_t)off[v + 1];
            std::sort(first, last);
            keep[v + 1] = (uint64_t)(std::unique(first, last) - first);
        }
    });
    for (uint32_t v = 0; v < n; ++v) keep[v + 1] += keep[v];
    ua.resize(keep[n]);
    ub.resize(keep[n]);
    parallel_for(n, 4096, [&](size_t v0, size_t v1) {
        for (size_t v = v0; v < v1; ++v) {
            const uint64_t c = keep[v + 1] - keep[v];
            for (uint64_t i = 0; i < c; ++i) {
                ua[keep[v] + i] = (uint32_t)v;
                ub[keep[v] + i] = hi[off[v] + i];
            }
        }
    });
}

int export_pairs(const std::vector<uint32_t>& ua, const std::vector<uint32_t>& ub, uint64_t* m,
                 uint32_t** src, uint32_t** dst) {
    const uint64_t mm = ua.size();
    uint32_t* s = (uint32_t*)std::malloc((mm ? mm : 1) * sizeof(uint32_t));
    uint32_t* d = (uint32_t*)std::malloc((mm ? mm : 1) * sizeof(uint32_t));
    if (!s || !d) {
        std::free(s);
        std::free(d);
        return fail(KT_ERR_NOMEM, "generator output");
    }
    if (mm) {
        std::memcpy(s, ua.data(), mm * sizeof(uint32_t));
        std::memcpy(d, ub.data(), mm * sizeof(uint32_t));
    }
    *m = mm;
    *src = s;
    *dst = d;
    return KT_OK;
}

}  // namespace

int generate_rmat(uint32_t scale, uint32_t ef, double A, double B, double C, uint64_t seed,
                  uint32_t* n_out, uint64_t* m, uint32_t** src, uint32_t** dst) {
    if (!n_out || !m || !src || !dst) return fail(KT_ERR_ARG, "NULL argument");
    if (scale == 0 || scale > 30) return fail(KT_ERR_ARG, "scale must be in [1, 30]");
    if (A < 0 || B < 0 || C < 0 || A + B + C > 1.0) return fail(KT_ERR_ARG, "bad R-MAT probabilities");
    const uint32_t n = 1u << scale;
    const uint64_t M = (uint64_t)ef << scale;
    std::vector<uint32_t> a(M), b(M);
    const double ab = A + B, abc = A + B + C;
    const uint64_t nblk = (M + kGenBlock - 1) / kGenBlock;
    parallel_for(nblk, 1, [&](size_t b0, size_t b1) {
        for (size_t blk = b0; blk < b1; ++blk) {
            Xoshiro rng(block_seed(seed, 1, blk));
            const uint64_t e1 = std::min<uint64_t>(M, (blk + 1) * kGenBlock);
            for (uint64_t e = blk * kGenBlock; e < e1; ++e) {
                uint32_t u = 0, v = 0;
                for (uint32_t l = 0; l < scale; ++l) {
                    const double r = rng.uniform();
                    const uint32_t bit = 1u << (scale - 1 - l);
                    if (r < A) {
                    } else if (r < ab) {
                        v |= bit;
                    } else if (r < abc) {
                        u |= bit;
                    } else {
                        u |= bit;
                        v |= bit;
                    }
                }
                a[e] = u;
                b[e] = v;
            }
        }
    });
    std::vector<uint32_t> ua, ub;
    clean_pairs(n, a, b, ua, ub);
    *n_out = n;
    return export_pairs(ua, ub, m, src, dst);
}

int generate_erdos_renyi(uint32_t n, double avg_degree, uint64_t seed, uint64_t* m,
                         uint32_t** src, uint32_t** dst) {
    if (!m || !src || !dst) return fail(KT_ERR_ARG, "NULL argument");
    if (n < 2 || n >= (1u << 31) || !(avg_degree > 0)) return fail(KT_ERR_ARG, "bad n or degree");
    const uint64_t M = (uint64_t)std::llround((double)n * avg_degree / 2.0);
    std::vector<uint32_t> a(M), b(M);
    const uint64_t nblk = (M + kGenBlock - 1) / kGenBlock;
    parallel_for(nblk, 1, [&](size_t b0, size_t b1) {
        for (size_t blk = b0; blk < b1; ++blk) {
            Xoshiro rng(block_seed(seed, 2, blk));
            const uint64_t e1 = std::min<uint64_t>(M, (blk + 1) * kGenBlock);
            for (uint64_t e = blk * kGenBlock; e < e1; ++e) {
                a[e] = (uint32_t)rng.below(n);
                b[e] = (uint32_t)rng.below(n);
            }
        }
    });
    std::vector<uint32_t> ua, ub;
    clean_pairs(n, a, b, ua, ub);
    return export_pairs(ua, ub, m, src, dst);
}

int generate_chung_lu(uint32_t n, double gamma, double avg_degree, double max_weight,
                      uint32_t num_cliques, uint32_t cmin, uint32_t cmax, uint64_t seed,
                      uint64_t* m, uint32_t** src, uint32_t** dst) {
    if (!m || !src || !dst) return fail(KT_ERR_ARG, "NULL argument");
    if (n < 2 || n >= (1u << 31) || !(gamma > 1.0) || !(avg_degree > 0))
        return fail(KT_ERR_ARG, "bad Chung-Lu parameters");
    if (num_cliques && (cmin < 2 || cmax < cmin || cmax > n))
        return fail(KT_ERR_ARG, "bad clique sizes");
    // Weights w_i = c (i + i0)^(-beta), beta = 1/(gamma-1); c fixes the mean
    // degree, i0 the largest weight (bisection on log i0).
    const double beta = 1.0 / (gamma - 1.0);
    const double target_sum = (double)n * avg_degree;
    auto sum_for = [&](double i0) {
        double s = 0;
        for (uint32_t i = 0; i < n; ++i) s += std::pow((double)i + i0, -beta);
        return s;
    };
    double i0 = 1.0;
    if (max_weight > 0) {
        double lo = std::log(1e-3), hi = std::log((double)n);
        for (int it = 0; it < 60; ++it) {
            const double mid = 0.5 * (lo + hi);
            const double x = std::exp(mid);
            const double c = target_sum / sum_for(x);
            const double wmax = c * std::pow(x, -beta);
            if (wmax > max_weight) lo = mid; else hi = mid;
        }
        i0 = std::exp(0.5 * (lo + hi));
    }
    const double c = target_sum / sum_for(i0);
    std::vector<double> cdf((size_t)n + 1, 0.0);
    for (uint32_t i = 0; i < n; ++i) cdf[i + 1] = cdf[i] + c * std::pow((double)i + i0, -beta);
    const double total = cdf[n];
    const uint64_t M = (uint64_t)std::llround(target_sum / 2.0);
    std::vector<uint32_t> a(M), b(M);
    const uint64_t nblk = (M + kGenBlock - 1) / kGenBlock;
    auto draw = [&](Xoshiro& rng) {
        const double x = rng.uniform() * total;
        size_t i = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
        if (i == 0) i = 1;
        if (i > n) i = n;
        return (uint32_t)(i - 1);
    };
    parallel_for(nblk, 1, [&](size_t b0, size_t b1) {
        for (size_t blk = b0; blk < b1; ++blk) {
            Xoshiro rng(block_seed(seed, 3, blk));
            const uint64_t e1 = std::min<uint64_t>(M, (blk + 1) * kGenBlock);
            for (uint64_t e = blk * kGenBlock; e < e1; ++e) {
                a[e] = draw(rng);
                b[e] = draw(rng);
            }
        }
    });
    // planted cliques on seeded vertex sets
    Xoshiro crng(block_seed(seed, 4, 0));
    for (uint32_t q = 0; q < num_cliques; ++q) {
        const uint32_t size = cmin + (uint32_t)crng.below((uint64_t)(cmax - cmin) + 1);
        std::vector<uint32_t> members;
        std::unordered_set<uint32_t> seen;
        while (members.size() < size) {
            const uint32_t v = (uint32_t)crng.below(n);
            if (seen.insert(v).second) members.push_back(v);
        }
        for (uint32_t i = 0; i < size; ++i)
            for (uint32_t j = i + 1; j < size; ++j) {
                a.push_back(members[i]);
                b.push_back(members[j]);
            }
    }
    std::vector<uint32_t> ua, ub;
    clean_pairs(n, a, b, ua, ub);
    return export_pairs(ua, ub, m, src, dst);
}

}  // namespace kt
