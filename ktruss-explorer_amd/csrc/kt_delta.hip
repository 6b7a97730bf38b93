// kt_delta.hip -- gfx950 kernels of DELTA support maintenance.
//
// Directed COO+CSR only (every undirected edge is one oriented entry e, and
// its support sup[e] is the exact number of triangles through it).  After a
// mark pass the deleted entries carry KT_DYING and sit in the list dl; the
// supports are brought to those of the surviving graph by
//   k_delta     one wave per deleted entry d = (u, v): the full adjacency
//               list of the endpoint with the shorter list is read 64
//               entries at a time and each neighbour w is binary-searched
//               (uniform trip count) in the other endpoint's list.  A hit is a
//               triangle {u, v, w} that was intact before the pass.  It is
//               handled by the deleted edge of lowest entry index among its
//               three edges, which decrements the supports of the triangle's
//               edges that survive -- so every destroyed triangle is removed
//               exactly once from every surviving edge;
//   k_finalize  the deleted entries become KT_DEAD in the oriented layout and
//               in both of their adjacency-list slots.
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

// Wave-aggregated append to a device counter; returns this lane's slot.
__device__ __forceinline__ uint32_t wave_append(unsigned int* ctr, bool pred, uint32_t lane) {
    const uint64_t m = ballot(pred);
    if (!m) return 0;
    const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(ctr, (unsigned int)__popcll(m));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

// mark_deleted_edges_kernel (kernel_ddc.cu:56-77), deferred: a deleted entry
// is flagged KT_DYING and listed; it becomes KT_DEAD in k_finalize.
__global__ void k_mark_dying(uint32_t* __restrict__ dst, const uint32_t* __restrict__ eid,
                             const uint32_t* __restrict__ sup, uint32_t* __restrict__ tau,
                             uint32_t* __restrict__ dl, uint32_t E, uint32_t k, DevCounters* ctr) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    for (uint32_t base = blockIdx.x * blockDim.x; base < E; base += gridDim.x * blockDim.x) {
        const uint32_t e = base + threadIdx.x;
        bool del = false;
        uint32_t v = 0;
        if (e < E) {
            v = dst[e];
            del = !(v & (KT_DEAD | KT_DYING)) && sup[e] < k - 2;
        }
        const uint32_t pos = wave_append(&ctr->ndel, del, lane);
        if (del) {
            dst[e] = v | KT_DYING;
            tau[eid[e]] = k - 1;
            dl[pos] = e;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_delta(const uint32_t* __restrict__ dl, uint32_t nd,
                                                  const uint32_t* __restrict__ src,
                                                  const uint32_t* __restrict__ dst,
                                                  const uint32_t* __restrict__ sptr,
                                                  const uint32_t* __restrict__ snbr,
                                                  const uint32_t* __restrict__ seid,
                                                  uint32_t* sup, DevCounters* ctr) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    unsigned long long acc_elems = 0, acc_tri = 0;
    for (uint32_t t = gwave(); t < nd; t += nwaves()) {
        const uint32_t d = dl[t];
        const uint32_t u = src[d];
        const uint32_t v = dst[d] & KT_IDMASK;
        const uint32_t ub = sptr[u], ue = sptr[u + 1];
        const uint32_t vb = sptr[v], ve = sptr[v + 1];
        const bool ushort = ue - ub <= ve - vb;
        const uint32_t qb = ushort ? ub : vb, qe = ushort ? ue : ve;
        const uint32_t lb = ushort ? vb : ub, ll = ushort ? ve - vb : ue - ub;
        const uint32_t skip = ushort ? v : u;  // the edge d itself
        acc_elems += qe - qb;
        if (ll == 0) continue;
        for (uint32_t j0 = qb; j0 < qe; j0 += kWave) {
            const uint32_t j = j0 + lane;
            const uint32_t x = j < qe ? snbr[j] : KT_DEAD;
            const bool cand = !(x & KT_DEAD) && x != skip;
            // lower bound of x in the other list, the same trip count on every lane
            uint32_t base = lb, len = ll;
            while (len > 1) {
                const uint32_t half = len >> 1;
                if ((snbr[base + half] & KT_VMASK) < x) base += half;
                len -= half;
            }
            const uint32_t p = base + ((snbr[base] & KT_VMASK) < x ? 1u : 0u);
            const bool hit = cand && p < lb + ll && snbr[p] == x;
            uint32_t tri = 0;
            if (hit) {
                const uint32_t e1 = seid[j], e2 = seid[p];
                const bool y1 = (dst[e1] & KT_DYING) != 0;
                const bool y2 = (dst[e2] & KT_DYING) != 0;
                // the deleted edge of lowest index owns the triangle
                if (!((y1 && e1 < d) || (y2 && e2 < d))) {
                    if (!y1) atomicSub(sup + e1, 1u);
                    if (!y2) atomicSub(sup + e2, 1u);
                    tri = 1;
                }
            }
            acc_tri += (uint32_t)__popcll(ballot(tri != 0));
        }
    }
    if (lane == 0 && (acc_elems || acc_tri)) {
        atomicAdd(&ctr->delta_elems, acc_elems);
        atomicAdd(&ctr->delta_tris, acc_tri);
    }
}

__global__ void k_finalize(const uint32_t* __restrict__ dl, uint32_t nd, uint32_t* __restrict__ dst,
                           uint32_t* __restrict__ snbr, const uint32_t* __restrict__ sa,
                           const uint32_t* __restrict__ sb) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nd; t += gridDim.x * blockDim.x) {
        const uint32_t d = dl[t];
        dst[d] = (dst[d] & ~KT_DYING) | KT_DEAD;
        snbr[sa[d]] |= KT_DEAD;
        snbr[sb[d]] |= KT_DEAD;
    }
}

uint32_t blocks_for(uint64_t items, uint32_t per_block) {
    uint64_t b = (items + per_block - 1) / per_block;
    if (b > 8192) b = 8192;
    return b ? (uint32_t)b : 1u;
}

}  // namespace

namespace dev {

hipError_t mark_dying(uint32_t* dst, const uint32_t* eid, const uint32_t* sup, uint32_t* tau,
                      uint32_t* dl, uint32_t E, uint32_t k, DevCounters* ctr, hipStream_t s) {
    if (!E) return hipSuccess;
    hipLaunchKernelGGL(k_mark_dying, dim3(blocks_for(E, kBlock)), dim3(kBlock), 0, s, dst, eid, sup,
                       tau, dl, E, k, ctr);
    return hipGetLastError();
}

hipError_t delta_update(const uint32_t* dl, uint32_t nd, const uint32_t* src, const uint32_t* dst,
                        const uint32_t* sptr, const uint32_t* snbr, const uint32_t* seid,
                        uint32_t* sup, DevCounters* ctr, hipStream_t s) {
    if (!nd) return hipSuccess;
    hipLaunchKernelGGL(k_delta, dim3(blocks_for(nd, kBlock / kWave)), dim3(kBlock), 0, s, dl, nd,
                       src, dst, sptr, snbr, seid, sup, ctr);
    return hipGetLastError();
}

hipError_t finalize_deleted(const uint32_t* dl, uint32_t nd, uint32_t* dst, uint32_t* snbr,
                            const uint32_t* sa, const uint32_t* sb, hipStream_t s) {
    if (!nd) return hipSuccess;
    hipLaunchKernelGGL(k_finalize, dim3(blocks_for(nd, kBlock)), dim3(kBlock), 0, s, dl, nd, dst,
                       snbr, sa, sb);
    return hipGetLastError();
}

}  // namespace dev
}  // namespace kt
