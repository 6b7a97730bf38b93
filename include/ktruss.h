/*
 * ktruss.h -- C ABI of the MI355X k-truss engine (libktruss_hip.so).
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point
 * returns KT_OK (0) or a negative kt_status; kt_last_error() gives the text.
 * Host buffers are caller-owned unless an entry point says it allocates
 * (free those with kt_free).  Thread safety: one plan per host thread.
 *
 * The reference (KTrussExplorer) has no FFI; its boundary is the C++ call
 *     void ktruss(COOGraph* graph, COOGraph* truss, Config config)
 * (kernel.h:12, defined kernel.cu:9-39), specialised per variant into
 * Device::{Directed,Undirected}::{COOCSRInput,TiledCOOCSRInput}::ktruss
 * (kernel.h:34-52).  Each entry point below names the reference code it
 * replaces.
 */
#ifndef KTRUSS_H
#define KTRUSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KT_ABI_VERSION 1

typedef enum {
    KT_OK = 0,
    KT_ERR_ARG = -1,        /* NULL pointer, k < 2, unknown enum value, bad size      */
    KT_ERR_SELF_LOOP = -2,  /* graph.cu:54 asserts no self-loop; here an error        */
    KT_ERR_ASYMMETRIC = -3, /* symmetric input lacks the reverse of an entry          */
    KT_ERR_DUPLICATE = -4,  /* an entry or pair appears twice                         */
    KT_ERR_RANGE = -5,      /* vertex id >= n, or n >= 2^31                           */
    KT_ERR_NOMEM = -6,      /* host or device allocation failed                      */
    KT_ERR_CAPACITY = -7,   /* caller buffer too small; the required size is returned */
    KT_ERR_IO = -8,         /* file missing or malformed                             */
    KT_ERR_HIP = -9,        /* a HIP runtime call failed                              */
    KT_ERR_NO_DEVICE = -10, /* no MI355X (gfx950) device visible                      */
    KT_ERR_STATE = -11      /* call order violated (e.g. results before a run)        */
} kt_status;

/* config.h:75-79, 111-114, 137-140 */
enum { KT_UNDIRECTED = 0, KT_DIRECTED_BY_INDEX = 1, KT_DIRECTED_BY_DEGREE = 2 };
enum { KT_COOCSR = 0, KT_TILED_COOCSR = 1 };
enum { KT_RECOUNT_ALL = 0, KT_RECOUNT_AFFECTED = 1 };

/* Decomposition schedule (the reference has one k per run; these govern the
 * k++ loop added on top).  JUMP: after an exact full count, go straight to
 * k = min live support + 3 (phases in between delete nothing).  REUSE
 * (directed, recount-all): a phase that ends on a deletion-free count hands
 * its exact supports to the next phase. */
enum { KT_SCHED_JUMP = 1u, KT_SCHED_REUSE = 2u };

/* Support-count strategies (directed COO+CSR; the other layouts use ROW).
 * ROW   : a wave walks its entries in order, staging each row of u in LDS and
 *         streaming v's row against it.
 * SEARCH: 64 entries at a time, all touched row tails staged, the candidate
 *         windows of v's rows concatenated and binary-searched lane by lane.
 * MERGE : as SEARCH, but each lane merges a contiguous share of the windows
 *         against the staged rows.
 * DELTA : support maintenance instead of a recount (directed COO+CSR,
 *         recount all): after a mark pass, every triangle destroyed by the
 *         deleted set is found by intersecting the full (in + out) adjacency
 *         lists of each deleted edge's endpoints, once per triangle, and the
 *         supports of its surviving edges are decremented.  The supports are
 *         then those a full recount would give, so the next mark deletes the
 *         same edges.  A pass with nothing to update (the first count) uses
 *         ROW.  Plans that use DELTA compact their entries (and the
 *         adjacency lists with them) once a third of the stored entries are dead. */
enum { KT_COUNT_AUTO = 0u, KT_COUNT_ROW = 1u, KT_COUNT_SEARCH = 2u, KT_COUNT_MERGE = 3u,
       KT_COUNT_DELTA = 4u };

/* Input flags for kt_graph_create. */
enum {
    KT_INPUT_SYMMETRIC = 0u, /* both (u,v) and (v,u) present, as createCOOFromFile
                                returns them (graph.cu:37-79)                       */
    KT_INPUT_PAIRS = 1u      /* each undirected edge exactly once, either direction */
};

/* One point of the design space.  The first seven fields mirror Config
 * (config.h:163-181); the rest are this engine's own knobs. */
typedef struct kt_variant {
    int32_t directedness;          /* -d                                                  */
    int32_t format;                /* -f                                                  */
    int32_t recount;               /* -r                                                  */
    uint32_t num_edge_remove_iter; /* -e: compact tombstones after the first e iterations */
    uint32_t block_size;           /* -b: threads per count workgroup (64..1024, x64)    */
    uint32_t tiles_per_dim;        /* -t: tiles per dimension (tiled format)             */
    uint32_t num_parallel_tiles;   /* -y: x-tile groups counted in parallel (tiled)      */
    uint32_t tile_size;            /* vertices per tile; 0 = ceil(n / tiles_per_dim)      */
    uint32_t chunk;                /* entries per count task; 0 = 128 (DELTA: 64)        */
    uint32_t schedule;             /* KT_SCHED_* bits for kt_plan_decompose              */
    uint32_t count_kernel;         /* KT_COUNT_*: support-count strategy                 */
    uint32_t cap_lds;              /* LDS slice per wave in entries; 0 = automatic       */
} kt_variant;

/* The reference's defaults (config.h:185-199): directed-by-degree, coocsr,
 * recount all, -e 0, -b 512, -t 8, -y 8; schedule JUMP|REUSE. */
void kt_variant_default(kt_variant* v);

/* Counters and times of the last ktruss / decompose run.  B_alg =
 *   sum over count passes [ sum_{e in R_p} (4 (L_u+L_v) + 24) + 12 |E_p| ]
 *   + sum over compactions 16 |E_c|
 * with L the stored lengths of the row segments a merge would scan. */
typedef struct kt_stats {
    uint64_t count_passes;
    uint64_t recount_edges;   /* sum |R_p|                        */
    uint64_t merge_elems;     /* sum_{e in R_p} (L_u + L_v)       */
    uint64_t live_in_passes;  /* sum |E_p|                        */
    uint64_t compactions;
    uint64_t compact_elems;   /* sum |E_c|                        */
    uint64_t iterations;      /* peel iterations, all phases      */
    uint64_t phases;          /* phases run                       */
    uint32_t k_max;           /* largest truss number             */
    uint32_t k_last;          /* k of the last phase              */
    double ms_total;          /* device time of the whole run (hipEvents)        */
    double ms_count;          /* summed device time of count-kernel launches (recounts) */
    uint64_t count_launches;  /* count-kernel launches timed into ms_count        */
    double balg_count_bytes;  /* B_alg of the count passes only                   */
    uint64_t delta_passes;    /* support updates done by DELTA (counted in count_passes) */
    uint64_t delta_edges;     /* deleted edges processed by DELTA                  */
    uint64_t delta_elems;     /* sum over them of the shorter adjacency list       */
    uint64_t delta_triangles; /* destroyed triangles found                         */
    double ms_delta;          /* summed device time of the DELTA update kernel     */
    double balg_delta_bytes;  /* DELTA algorithmic bytes: 4 (candidates + window
                                 entries read) + 32 per destroyed triangle          */
    uint64_t delta_live;      /* DELTA candidates whose edge was still alive       */
    double ms_setup;          /* wall time of the one-time set-up this run did before
                                 ms_total started (the DELTA adjacency lists, built on
                                 a plan's first DELTA run); 0 when there was none  */
} kt_stats;

typedef struct kt_graph kt_graph; /* validated host-side edge set            */
typedef struct kt_plan kt_plan;   /* device-resident layout of one variant    */

/* ---- runtime ------------------------------------------------------------ */
int kt_abi_version(void);
const char* kt_last_error(void);
const char* kt_status_string(int status);
/* Select the device (hipSetDevice) and check it is gfx950.  Optional: the
 * first plan calls it with device 0. */
int kt_init(int device);
/* hipDeviceSynchronize on the current device (used to bracket timed regions). */
int kt_device_synchronize(void);
void kt_free(void* p);

/* ---- graphs (host) -------------------------------------------------------
 * Replaces createCOOFromFile / the COOGraph handed to ktruss (kernel.cu:9):
 * validates what the reference asserts or assumes (no self-loop, graph.cu:54;
 * symmetric input, kernel_h.cu:24), numbers the undirected edges and keeps the
 * map from every input entry to its edge.  Inputs of 2^16 entries or more
 * are checked and numbered on the device once kt_init (or a plan) has
 * initialised it, others on the host; either way an error names the lowest
 * offending entry. */
int kt_graph_create(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                    uint32_t flags, kt_graph** out);
void kt_graph_destroy(kt_graph* g);
/* n, undirected edge count, input entry count, maximum degree. */
int kt_graph_info(const kt_graph* g, uint32_t* n, uint64_t* m_undirected, uint64_t* m_input,
                  uint32_t* max_degree);

/* createCOOFromFile (graph.cu:37-79): "dst src w" per line, 1-based.  Returns
 * malloc'ed arrays (kt_free) of 0-based entries. */
int kt_read_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst);
/* writeCOOGraphToFile (graph.cu:96-104). */
int kt_write_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst);
/* Truss numbers as text: the graph TSV's layout with the weight column holding
 * each entry's truss number, "dst+1 \t src+1 \t tau" per line (what
 * `ktruss -D -o` writes).  No counterpart in the reference, which outputs a
 * k-truss only; kt_read_tsv reads such a file as a graph. */
int kt_write_truss_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst,
                       const uint32_t* tau);
/* kt_read_tsv that also returns the third column (malloc'ed, kt_free). */
int kt_read_truss_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst,
                      uint32_t** tau);
/* Binary edge cache: a 32-byte little-endian header ("KTEDGES1", n, flags,
 * m, checksum) then src[m] and dst[m] as u32, so large inputs skip text
 * parsing.  No counterpart in the reference.  The reader checks the size
 * against m and the checksum, and returns malloc'ed arrays (kt_free) and the
 * KT_INPUT_* flags the entries were written with (flags may be NULL). */
int kt_write_edges_bin(const char* path, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                       uint32_t flags);
int kt_read_edges_bin(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst,
                      uint32_t* flags);
/* 1 when the file starts with the binary edge cache's magic, else 0. */
int kt_is_edges_bin(const char* path);
/* verify (verify.cu:12-40) without the abort: 0 when both COO graphs hold the
 * same entries, 1 when they differ. */
int kt_verify_edges(uint64_t m_a, const uint32_t* src_a, const uint32_t* dst_a,
                    uint64_t m_b, const uint32_t* src_b, const uint32_t* dst_b);

/* ---- synthetic graphs (undirected pairs u < v, sorted, deduplicated,
 * self-loops removed; arrays malloc'ed, kt_free) ---------------------------- */
int kt_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                     uint64_t seed, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst);
/* Chung-Lu with power-law weights of exponent gamma, mean degree avg_degree,
 * weights capped at max_weight (0 = no cap), plus num_cliques planted cliques
 * of sizes uniform in [clique_min, clique_max] on seeded vertex sets. */
int kt_generate_chung_lu(uint32_t n, double gamma, double avg_degree, double max_weight,
                         uint32_t num_cliques, uint32_t clique_min, uint32_t clique_max,
                         uint64_t seed, uint64_t* m, uint32_t** src, uint32_t** dst);
/* Erdos-Renyi G(n, m = n*avg_degree/2) sampled pairs. */
int kt_generate_erdos_renyi(uint32_t n, double avg_degree, uint64_t seed, uint64_t* m,
                            uint32_t** src, uint32_t** dst);

/* ---- plans (device) ------------------------------------------------------
 * Replaces the set-up half of Device::ktruss (kernel_d.cu:11-33) and of the
 * variant wrappers (kernel_ddc.cu:216-224, kernel_ddt.cu:226-234,
 * kernel_duc.cu:193-201, kernel_dut.cu:205-213): relabel (kernel.cu:13-20),
 * orientation (graph.cu:194-237), CSR / tiled build (graph.cu:372-478,
 * 656-780), device allocation of the peel state (info.cu:22-35).  The layout
 * stays resident; every run starts from it. */
int kt_plan_create(const kt_graph* g, const kt_variant* v, kt_plan** out);
void kt_plan_destroy(kt_plan* p);
/* Stored entries of the layout (directed: m; undirected: 2m), max row length. */
int kt_plan_info(const kt_plan* p, uint64_t* entries, uint32_t* max_row);

/* One k, the reference's ./ktruss -k k: the peel loop of
 * Device::...::ktruss(COOCSRGraph* / TiledCOOCSRGraph*, Info, Config)
 * (kernel_ddc.cu:116-214, kernel_ddt.cu:125-224, kernel_duc.cu:95-191,
 * kernel_dut.cu:106-203) and the conversion back to a symmetric COO in the
 * caller's labels (kernel_ddc.cu:246-250, kernel_d.cu:57-79, kernel.cu:31-37).
 * out_src/out_dst receive 2*|T_k| entries; *out_m is set even when cap is too
 * small (then KT_ERR_CAPACITY). */
int kt_plan_ktruss(kt_plan* p, uint32_t k, uint32_t* out_src, uint32_t* out_dst, uint64_t cap,
                   uint64_t* out_m, kt_stats* stats);

/* One k of the directed COO+CSR peel, recount all, on a layout the caller
 * holds in device memory, allocating nothing (no hipMalloc, no hipFree, no
 * allocating library call).  No counterpart in the reference: it is for
 * callers that cannot allocate device memory -- serving loops and harnesses
 * that forbid hipMalloc.  It is kt_plan_ktruss without the plan: the layout
 * is the caller's COOCSRGraph (srcIdx, dstIdx, srcPtr of graph.h:40-47) and
 * the peel state (supports, counters) lives in a fixed device array of the
 * library, about 32 MiB per device, held until the process ends.
 *   ptr[n+1], src[E], dst[E]: device arrays of the current device, oriented by
 *     id (src[e] < dst[e] < n, each undirected edge once), rows by source,
 *     destinations strictly ascending per row (a degree-ordered relabelling
 *     makes it the reference's directed-by-degree layout);
 *   dst is also the output: entries outside the k-truss get bit 31 set (the
 *     vertex id stays below it), the others are unchanged;
 *   stream: a hipStream_t of the current device (NULL: the null stream); the
 *     call returns after the stream has run its kernels (it reads a
 *     deletion count back every iteration, so it cannot be captured);
 *   out_live (optional): entries left in the k-truss.
 * E above KT_NOALLOC_MAX_ENTRIES is KT_ERR_CAPACITY; a layout that breaks the
 * rules above is KT_ERR_ARG naming the first offending entry or vertex,
 * found on the device before any entry changes.  Every call on one device
 * shares that device's array (each call clears what must start at zero), so
 * calls that overlap on one device -- from two host threads or on two
 * streams -- race and must not be made. */
#define KT_NOALLOC_MAX_ENTRIES (1u << 23)
int kt_ktruss_noalloc(uint32_t n, uint32_t E, const uint32_t* ptr, const uint32_t* src, uint32_t* dst,
                      uint32_t k, void* stream, uint64_t* out_live);

/* Full decomposition (k++ loop over the same peel) on the device-resident
 * layout; truss numbers stay on the device.  stats->ms_total is the hipEvent
 * time from the resident oriented layout to the resident truss numbers. */
int kt_plan_decompose(kt_plan* p, kt_stats* stats);
/* Timing (on by default): the support passes are bracketed by hipEvents,
 * which fill stats->ms_count / ms_delta and the time column of the trace.  Off:
 * DELTA iterations queued ahead of the host record no events (each costs the
 * device a few microseconds between dependent kernels); their ms figures and
 * trace times stay 0.  Results are the same either way. */
int kt_plan_set_timing(kt_plan* p, int on);
/* Copy the last decomposition's truss numbers: one per input entry of the
 * graph the plan was made from (count = m_input). */
int kt_plan_truss_numbers(kt_plan* p, uint32_t* tau, uint64_t count);
/* Per-iteration trace of the last run: (k, iteration, undirected edges
 * deleted, entries recounted, count-kernel nanoseconds) as 5 uint64 per
 * record; *len is the number of records (all of them even when cap is
 * smaller). */
int kt_plan_trace(const kt_plan* p, uint64_t* recs, uint64_t cap, uint64_t* len);

/* Per-iteration log of the last run, KT_ITER_REC uint64 per iteration -- the
 * quantities the reference prints per iteration at -v 2 (kernel_ddc.cu:126-198
 * and its three twins):
 *   [0] k, [1] iteration within the phase, [2] undirected edges deleted,
 *   [3] entries recounted, [4] support-pass ns ("Triangle counting time"),
 *   [5] mark ns ("Mark deleted edges"), [6] removal ns ("Remove deleted
 *   edges", the first -e iterations), [7] stored entries after that removal
 *   ("# edges remaining"), [8] affected-marking ns ("Mark affected edges"),
 *   [9] removal ns after the phase's loop (kernel_ddc.cu:207-212; last
 *   iteration of a phase only), [10] KT_ITER_* flags saying which of them
 *   happened, [11] 0.
 * Times are device times (hipEvents) when plan timing is on (the default),
 * removal times host wall times around the removal (it reads the kept count
 * back); with timing off, or for steps fused into another kernel (DELTA
 * marks), the KT_ITER_*_TIMED flag is clear and the time is 0. */
#define KT_ITER_REC 12
enum {
    KT_ITER_COUNTED = 1u,        /* a support pass ran (not an iteration 0 handed exact supports) */
    KT_ITER_MARK_TIMED = 2u,     /* [5] holds the mark's time                           */
    KT_ITER_REMOVED = 4u,        /* deleted entries were removed in the iteration ([6], [7]) */
    KT_ITER_AFFECTED = 8u,       /* affected marking ran ([8])                          */
    KT_ITER_FINAL_REMOVAL = 16u, /* deleted entries were removed after the loop ([9])   */
    KT_ITER_COUNT_TIMED = 32u,   /* [4] holds the support pass's time                   */
    KT_ITER_DELTA = 64u          /* the support pass was a DELTA update                 */
};
int kt_plan_iteration_log(const kt_plan* p, uint64_t* recs, uint64_t cap, uint64_t* len);

/* Wall times (ms, host clock around device work that is synchronised) of the
 * stages of kt_plan_create and of the last kt_plan_ktruss, named after the
 * reference's -v 1 lines for the same work (kernel.cu:19, 36;
 * kernel_d.cu:20-84; kernel_ddc.cu:223, 249):
 *   [0] relabel by degree (directed-by-degree; 0 otherwise)
 *   [1] copy of the edge list to the GPU
 *   [2] allocation of the peel state
 *   [3] orientation (directed variants; 0 for undirected)
 *   [4] build of the stored COO+CSR rows
 *   [5] free of the edge list's device copy
 *   [6] gathering the surviving entries (COO+CSR -> COO)
 *   [7] adding the reverse entries (directed -> undirected COO; directed only)
 *   [8] copy of the stored entries from the GPU
 *   [9] restoring the caller's vertex labels */
#define KT_PLAN_STAGES 10
int kt_plan_stage_times(const kt_plan* p, double* ms, uint32_t n);
/* kt_plan_destroy that times what it frees, in the reference's order
 * (kernel_ddc.cu:254, kernel_d.cu:66, 72): ms[0] the working layout
 * ("Deallocate COOCSR"), ms[1] the pristine copy of the layout ("Deallocate
 * directed COO"), ms[2] the peel state and the rest ("Deallocate intermediate
 * data structures").  ms may be NULL. */
void kt_plan_destroy_timed(kt_plan* p, double* ms);

/* ---- autotuner of one plan ------------------------------------------------
 * No counterpart in the reference (its shapes come from -b, -t, -y).
 * Coordinate descent, one pass over each axis in turn:
 *   tiled plans:   tile size (the plan's own, the reference's n/8, 65536,
 *                  16384, 4096 vertices), destination classes (-y: 1, 2, 4,
 *                  8, 16) and workgroup size (256, 512, 1024);
 *   COO+CSR plans: count strategy (directed: ROW, SEARCH, MERGE; a DELTA plan
 *                  keeps DELTA and tunes the shape of its opening count),
 *                  workgroup size, LDS slice (automatic, 512, 1024, 2048) and
 *                  task size (64, 128, 256).
 * A COO+CSR plan first tries each of its per-phase candidates
 * (kt_plan_set_candidates; a DELTA candidate as the count that opens its
 * runs), so the log holds every candidate's cost on the whole graph: a
 * training sample of the cost model (learn.py appends them).
 * A trial is one full support count of the plan's pristine layout from zeroed
 * supports at k, timed with hipEvents, best of `repeats`; a tiled plan's layout
 * is rebuilt for each tile shape tried.  The fastest configuration is applied
 * to the plan (and returned in *best when best is not NULL); every
 * configuration counts the same supports, so results are unchanged.  Call it
 * before a run: it clears the last run's results. */
typedef struct kt_tune_trial {
    uint32_t tile_size;          /* vertices per tile (tiled) or 0            */
    uint32_t num_parallel_tiles; /* destination classes (tiled) or 1          */
    uint32_t count_kernel;       /* KT_COUNT_*                                */
    uint32_t block_size;
    uint32_t cap_lds;            /* 0 = automatic                             */
    uint32_t chunk;
    float ms;                    /* best of the repeats                       */
    uint32_t pad;
} kt_tune_trial;
int kt_plan_autotune(kt_plan* p, uint32_t k, uint32_t repeats, kt_variant* best);
/* Trials of the last kt_plan_autotune; *len is their number (all of them
 * even when cap is smaller). */
int kt_plan_tune_log(const kt_plan* p, kt_tune_trial* recs, uint64_t cap, uint64_t* len);

/* ---- learned count-kernel picker -----------------------------------------
 * No counterpart in the reference (its variant is fixed per run).  A count
 * configuration is the part of a variant that can change between phases
 * without touching the layout.  Candidates are tried per phase:
 *   FIXED  every phase uses the plan's own variant;
 *   SWEEP  every count pass runs every candidate (each from zeroed supports)
 *          and times it: the exhaustive baseline and the training labels;
 *   MODEL  at each phase start a device kernel computes KT_MLP_IN features
 *          of the surviving graph and the MLP (inference on the device)
 *          picks the candidate with the lowest predicted cost. */
#define KT_MLP_IN 16
#define KT_MLP_HIDDEN 64
#define KT_MLP_MAX_OUT 16
#define KT_MAX_CANDIDATES 16
#define KT_PHASE_REC 36

typedef struct kt_count_config {
    uint32_t count_kernel; /* KT_COUNT_*                              */
    uint32_t block_size;   /* threads per workgroup, multiple of 64    */
    uint32_t cap_lds;      /* LDS slice per wave; 0 = automatic        */
    uint32_t chunk;        /* items per wave task; 0 = default (128;   */
                           /* 64 for DELTA, whose full count is the   */
                           /* pass that opens a run)                  */
} kt_count_config;

enum { KT_PICK_FIXED = 0u, KT_PICK_SWEEP = 1u, KT_PICK_MODEL = 2u };

typedef struct kt_mlp kt_mlp;

int kt_plan_set_candidates(kt_plan* p, const kt_count_config* cands, uint32_t n);
/* model is read when mode is KT_PICK_MODEL (its parameters are copied). */
int kt_plan_set_pick(kt_plan* p, uint32_t mode, const kt_mlp* model);
/* Sampled SWEEP (training data from graphs too large to sweep whole): in
 * KT_PICK_SWEEP mode every candidate is timed only in the phases whose index
 * is offset modulo every; the other phases run candidate `fill` alone and log
 * no per-candidate times.  every = 1 (the default) sweeps every phase; runs
 * with offsets 0 .. every-1 together sweep every phase once (the peel, and so
 * the phases, are the same in every run). */
int kt_plan_set_sweep_sampling(kt_plan* p, uint32_t every, uint32_t offset, uint32_t fill);
/* Phase log of the last decomposition, KT_PHASE_REC floats per phase:
 * [0] k, [1] iterations, [2] chosen candidate, [3] count ms of the chosen,
 * [4..19] features, [20..35] count ms per candidate (SWEEP; -1 otherwise). */
int kt_plan_phase_log(const kt_plan* p, float* recs, uint64_t cap, uint64_t* len);

/* MLP KT_MLP_IN -> KT_MLP_HIDDEN -> KT_MLP_HIDDEN -> n_out, ReLU, FP32, on the
 * device.  Loss: sum_s w_s sum_c mask_sc (y_sc - t_sc)^2 / sum_s w_s. */
int kt_mlp_create(uint32_t n_out, uint64_t seed, kt_mlp** out);
void kt_mlp_destroy(kt_mlp* m);
int kt_mlp_num_params(const kt_mlp* m, uint32_t* n);
int kt_mlp_get_params(const kt_mlp* m, float* out, uint32_t n);
int kt_mlp_set_params(kt_mlp* m, const float* in, uint32_t n);
/* Inference inputs (kt_mlp_forward and the picks of kt_plan_set_pick) are
 * clamped to [lo, hi] per feature -- set it to the training features' range so
 * the network never extrapolates; NULL bounds mean unbounded.  Training does
 * not clamp. */
int kt_mlp_set_input_range(kt_mlp* m, const float* lo, const float* hi);
int kt_mlp_get_input_range(const kt_mlp* m, float* lo, float* hi);
int kt_mlp_forward(kt_mlp* m, const float* x, uint32_t batch, float* y);
/* Loss and its gradient (deterministic); mask and w may be NULL. */
int kt_mlp_loss_grad(kt_mlp* m, const float* x, const float* t, const float* mask, const float* w,
                     uint32_t batch, float* loss, float* grad);
/* Adam on mini-batches drawn in a seeded order; *loss_out is the loss over
 * all n samples after training. */
int kt_mlp_train(kt_mlp* m, const float* x, const float* t, const float* mask, const float* w,
                 uint32_t n, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                 float* loss_out);

/* ---- batched decomposition ----------------------------------------------
 * No counterpart in the reference (its ktruss(COOGraph*, COOGraph*, Config),
 * kernel.h:12, takes one graph per process).  Many independent graphs are
 * laid out as one disjoint union on the device and decomposed together: every
 * peel step advances every unfinished graph by one iteration of its own peel
 * (its own k), and the steps are captured once as a HIP graph and replayed.
 * Truss numbers are identical to kt_plan_decompose on each graph alone.
 * Supported variants: directed (by degree or by index), COO+CSR, recount all;
 * count_kernel KT_COUNT_DELTA maintains the supports with DELTA updates after
 * the first full count (every step: mark, update, control). */
typedef struct kt_batch kt_batch;

typedef struct kt_batch_tuning {
    uint32_t block_size;       /* count workgroup size (64..1024, x64); 0 = 512   */
    uint32_t chunk;            /* entries per count task (x64); 0 = 128            */
    uint32_t cap_lds;          /* LDS slice per wave; 0 = automatic                */
    uint32_t blocks_per_cu;    /* persistent count workgroups per CU; 0 = 2        */
    uint32_t mark_chunk;       /* entries per mark task (x64); 0 = 1024            */
    uint32_t steps_per_replay; /* peel steps per captured graph replay; 0 = 16; rounded up to even */
} kt_batch_tuning;

typedef struct kt_batch_stats {
    uint64_t graphs;
    uint64_t entries;          /* stored entries of all graphs                     */
    uint64_t steps;            /* peel steps (mark passes over the active graphs)  */
    uint64_t replays;          /* graph replays launched                           */
    uint64_t count_passes;     /* sum over graphs                                  */
    uint64_t iterations;       /* sum over graphs                                  */
    uint64_t phases;           /* sum over graphs                                  */
    uint64_t recount_edges;
    uint64_t merge_elems;
    uint64_t live_in_passes;
    uint32_t k_max;            /* largest truss number over all graphs             */
    uint32_t pad;
    double ms_total;           /* device time from resident layouts to resident truss numbers */
    double balg_count_bytes;   /* B_alg of the count passes                        */
    double ms_count;           /* summed count-kernel time (timing mode only)      */
    uint64_t count_launches;   /* count-kernel launches                            */
    uint64_t delta_elems;      /* DELTA: adjacency entries read                    */
    uint64_t delta_triangles;  /* DELTA: destroyed triangles removed               */
    double ms_delta;           /* summed DELTA update time (timing mode only)      */
    double balg_delta_bytes;   /* DELTA algorithmic bytes (as kt_stats)            */
} kt_batch_stats;

/* graphs[count] stay referenced by the batch until kt_batch_destroy. */
int kt_batch_create(const kt_graph* const* graphs, uint32_t count, const kt_variant* v,
                    kt_batch** out);
void kt_batch_destroy(kt_batch* b);
/* Graphs, stored entries, longest row over the batch. */
int kt_batch_info(const kt_batch* b, uint32_t* graphs, uint64_t* entries, uint32_t* max_row);
int kt_batch_set_tuning(kt_batch* b, const kt_batch_tuning* t);
/* The resolved tuning in use (zeros replaced by the values chosen). */
int kt_batch_get_tuning(const kt_batch* b, kt_batch_tuning* t);
/* Autotuner: coordinate descent over block size, LDS slice, count chunk,
 * count workgroups per CU and mark chunk.  Each trial runs the first
 * trial_steps peel steps of the batch from its initial state, best of
 * `repeats`, timed with hipEvents; the fastest tuning is applied.  The trial
 * log is read with kt_batch_tune_log. */
int kt_batch_autotune(kt_batch* b, uint32_t trial_steps, uint32_t repeats, kt_batch_tuning* best);
/* Trials of the last autotune: tunings and milliseconds; *len is the number
 * of trials (all of them even when cap is smaller). */
int kt_batch_tune_log(const kt_batch* b, kt_batch_tuning* cfg, float* ms, uint64_t cap,
                      uint64_t* len);
/* Timing mode (on != 0): decompose launches the peel steps one by one instead
 * of replaying the captured graph and brackets every count launch with
 * hipEvents, filling stats->ms_count.  Results are the same either way. */
int kt_batch_set_timing(kt_batch* b, int on);
int kt_batch_decompose(kt_batch* b, kt_batch_stats* stats);
/* Truss numbers of graph `graph` of the batch, one per input entry of that
 * graph (count = its m_input). */
int kt_batch_truss_numbers(kt_batch* b, uint32_t graph, uint32_t* tau, uint64_t count);
/* Per-graph counters of the last run: info[0] largest truss number,
 * [1] peel iterations, [2] phases, [3] count passes. */
int kt_batch_graph_info(kt_batch* b, uint32_t graph, uint32_t* info);

#ifdef __cplusplus
}
#endif
#endif
