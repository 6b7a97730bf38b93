/*
 * ktoracle.h -- CPU oracle for the k-truss peel path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libktruss_hip.so, the ktruss CLI) never links or calls it.
 *
 * It restates, single-threaded and in plain C, the reference's Host path
 * (KTrussExplorer, "-p cpu"):
 *   kernel.cu:9-39            relabel -> peel -> unsort
 *   kernel_h.cu:11-68         orientation, variant dispatch, re-symmetrize
 *   kernel_hdc.cu:11-201      directed   / COO+CSR   count, mark, affected, peel loop
 *   kernel_hdt.cu:11-214      directed   / tiled     count, mark, affected, peel loop
 *   kernel_huc.cu:11-179      undirected / COO+CSR   count, mark, affected, peel loop
 *   kernel_hut.cu:11-190      undirected / tiled     count, mark, affected, peel loop
 *   info.cu:7-91              per-iteration state (changed, numTriangles, affected flags)
 *   graph.cu:37-104,152-300,372-412,480-509,656-707,782-818   I/O, relabel,
 *                             orientation, CSR and tiled builds, compaction
 *   verify.cu:12-40           edge-set comparator
 * and adds the k++ outer loop that turns per-k peeling into a truss number per
 * edge (the reference computes one k per run).
 *
 * Parity pinning (what checks each function, so a plausible mistake fails):
 *   - per-k edge sets: networkx k_truss (an independent implementation of the
 *     same "every edge in >= k-2 triangles" definition) on seeded graphs for
 *     all 12 variant combinations, and the known-answer graph whose outputs
 *     were recorded from the reference's CPU path (tests/golden);
 *   - truss numbers: closed forms (K_n -> n, cycles and trees -> 2, K_n minus
 *     an edge, wheels, planted cliques), brute-force enumeration on tiny graphs,
 *     nesting T_{k+1} within T_k, every edge of T_k having >= k-2 triangles
 *     inside T_k;
 *   - per-iteration deletion traces: identical across every variant and
 *     schedule (the peel is a Jacobi iteration whose deletion sets do not
 *     depend on layout or orientation);
 *   - B_alg counters: closed form on K_n (every pass, every row length known).
 */
#ifndef KTORACLE_H
#define KTORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KTO_DELETED 0xFFFFFFFFu /* graph.h:8 */

enum { KTO_UNDIRECTED = 0, KTO_DIRECTED_BY_INDEX = 1, KTO_DIRECTED_BY_DEGREE = 2 }; /* config.h:75-79 */
enum { KTO_COOCSR = 0, KTO_TILED_COOCSR = 1 };                                      /* config.h:111-114 */
enum { KTO_RECOUNT_ALL = 0, KTO_RECOUNT_AFFECTED = 1 };                             /* config.h:137-140 */

/* Decomposition schedule bits (the k++ loop is not in the reference).
 * 0                  = plain: k = 3, 4, 5, ... each phase a fresh per-k peel
 * KTO_SCHED_JUMP     = when a phase's first count is exact, jump k to
 *                      (min live support) + 3: the skipped phases delete nothing
 * KTO_SCHED_REUSE    = directed recount-all only: a phase that ended on a count
 *                      with no deletion hands its exact supports to the next
 *                      phase, which then skips its first count */
enum { KTO_SCHED_JUMP = 1, KTO_SCHED_REUSE = 2 };

/* Error codes. */
enum {
    KTO_OK = 0,
    KTO_ERR_ARG = -1,       /* bad argument (NULL, k < 2, bad enum)          */
    KTO_ERR_SELF_LOOP = -2, /* graph.cu:54 asserts src != dst                 */
    KTO_ERR_ASYMMETRIC = -3,/* input must hold (u,v) and (v,u)                */
    KTO_ERR_DUPLICATE = -4, /* repeated entry                                 */
    KTO_ERR_RANGE = -5,     /* vertex id >= n                                 */
    KTO_ERR_NOMEM = -6,
    KTO_ERR_CAPACITY = -7,  /* output or trace buffer too small               */
    KTO_ERR_IO = -8
};

typedef struct {
    int directedness;               /* -d */
    int format;                     /* -f */
    int recount;                    /* -r */
    uint32_t num_edge_remove_iter;  /* -e */
    uint32_t tiles_per_dim;         /* -t */
} kto_variant;

/* Counters behind the algorithmic-byte model
 *   B_alg = sum_count_passes [ sum_{e in R_p} (4*(L_u+L_v) + 24) + 12*|E_p| ]
 *         + sum_compactions 16*|E_c|
 * L = stored length (tombstones included) of every row segment the variant's
 * merge would scan without early exit; |E_p| = live stored entries at the pass;
 * |E_c| = stored entries scanned by a compaction. */
typedef struct {
    uint64_t count_passes;
    uint64_t recount_edges;   /* sum |R_p|                 */
    uint64_t merge_elems;     /* sum_{e in R_p} (L_u+L_v)  */
    uint64_t live_in_passes;  /* sum |E_p|                 */
    uint64_t compactions;
    uint64_t compact_elems;   /* sum |E_c|                 */
    uint64_t iterations;      /* peel iterations, all phases */
    uint64_t phases;          /* phases actually run         */
    uint64_t triangles;       /* triangles found by the first exact full count (directed) */
    uint32_t k_max;           /* largest truss number (decompose)  */
    uint32_t pad;
} kto_stats;

double kto_balg_bytes(const kto_stats* s);

/* Per-iteration record: k, iteration index within the phase, undirected
 * edges deleted in that iteration, edges recounted in that iteration. */
typedef struct {
    uint32_t k;
    uint32_t iter;
    uint64_t deleted;
    uint64_t recounted;
} kto_trace_rec;

typedef struct {
    kto_trace_rec* recs;
    uint64_t cap;
    uint64_t len;
} kto_trace;

/* Single-k run: the reference's ./ktruss -p cpu -k k on a symmetric COO graph
 * (both (u,v) and (v,u) present, 0-based).  Writes the k-truss as a symmetric
 * COO into out_src/out_dst (capacity cap entries, in the reference's order:
 * oriented survivors then their reverses for directed variants).  trace and
 * stats may be NULL. */
int kto_ktruss(uint32_t n, uint64_t m_sym, const uint32_t* src, const uint32_t* dst,
               uint32_t k, const kto_variant* v,
               uint32_t* out_src, uint32_t* out_dst, uint64_t cap, uint64_t* out_m,
               kto_stats* stats, kto_trace* trace);

/* Full decomposition via the k++ loop: tau[i] = truss number of input entry i
 * (max k with the edge in the k-truss; 2 for an edge in no triangle). */
int kto_decompose(uint32_t n, uint64_t m_sym, const uint32_t* src, const uint32_t* dst,
                  const kto_variant* v, uint32_t schedule,
                  uint32_t* tau, kto_stats* stats, kto_trace* trace);

/* Brute force for tiny graphs (n <= 4096): truss numbers by repeatedly
 * deleting one edge whose triangle count, enumerated over all vertex
 * triples' adjacency-matrix entries, is below k-2.  Shares no code with the
 * peel. */
int kto_bruteforce_tau(uint32_t n, uint64_t m_sym, const uint32_t* src, const uint32_t* dst,
                       uint32_t* tau);

/* graph.cu:37-79: read "dst src w" lines (1-based) into a COO.  The src and dst arrays are
 * malloc'ed; free with kto_free. */
int kto_read_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst);
/* graph.cu:96-104 */
int kto_write_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst);
/* verify.cu:12-40: 0 when both COO graphs hold the same entries. */
int kto_verify(uint32_t n, uint64_t m_a, const uint32_t* src_a, const uint32_t* dst_a,
               uint64_t m_b, const uint32_t* src_b, const uint32_t* dst_b);
void kto_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
