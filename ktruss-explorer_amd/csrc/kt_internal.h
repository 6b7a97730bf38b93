// kt_internal.h -- shared declarations of the MI355X k-truss engine.
//
// Device layout of one plan (all arrays in HBM, uint32 unless noted):
//   ptr[n+1]            row starts of the stored (oriented or symmetric) CSR
//   src[E], dst[E]      COO+CSR entries (graph.h:40-47); a deleted entry keeps
//                       its vertex id and gets bit 31 set (KT_DEAD), so a row
//                       stays sorted under (x & KT_VMASK) and can be binary
//                       searched with tombstones in place.  The reference
//                       overwrites dstIdx with UINT_MAX instead (graph.h:8).
//   eid[E]              pristine position of each entry (truss numbers are
//                       written at tau[eid])
//   sup[E]              support (Info.numTriangles, info.h:16)
//   state[E] (u8)       NOT / DIRECTLY / INDIRECTLY affected (info.h:10-12)
//   nflag[n] (u8)       per-vertex affected state (Info.nodeAffected)
//   wl[E]               worklist of entries to recount (affected mode)
//   tau[E]              truss number per pristine position
#pragma once

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include "ktruss.h"

#define KT_DEAD 0x80000000u
#define KT_VMASK 0x7FFFFFFFu
// DELTA plans flag an entry deleted by the current mark pass with KT_DYING
// until its triangles are removed from the surviving supports; vertex ids
// then stay below 2^30.
#define KT_DYING 0x40000000u
#define KT_IDMASK 0x3FFFFFFFu
#define KT_NOT_AFFECTED 0
#define KT_DIRECTLY_AFFECTED 1
#define KT_INDIRECTLY_AFFECTED 2

namespace kt {

// Error plumbing: every failing path records a message for kt_last_error().
void set_error(int status, const std::string& msg);
int fail(int status, const std::string& msg);

#define KT_HIP_TRY(call)                                                              \
    do {                                                                              \
        hipError_t kt_e_ = (call);                                                    \
        if (kt_e_ != hipSuccess)                                                      \
            return ::kt::fail(KT_ERR_HIP, std::string(#call) + ": " +                 \
                                              hipGetErrorString(kt_e_));              \
    } while (0)

// Counters the device updates and the host reads back after each iteration.
// Support of an entry a DELTA mark deleted: above every threshold, so later
// marks find it without reading its flags.
constexpr uint32_t KT_SUP_GONE = 0xFFFFFFFFu;
// Support of an entry a DELTA mark deleted while it lay in no triangle (its
// support was 0): above every threshold as KT_SUP_GONE, and its update has
// nothing to find, so the task list gives it no tasks.
constexpr uint32_t KT_SUP_NONE = 0xFFFFFFFEu;

struct DevCounters {
    unsigned long long merge_elems;   // sum of L_u + L_v over recounted entries
    unsigned long long recount_edges; // recounted entries
    // the three deletion lists' counts, each followed by the DELTA task total
    // of its deletions: a mark that lists the tasks reserves (deletions,
    // tasks) with one 64-bit atomic on the pair, so the first tasks ascend
    // with the list positions
    unsigned int ndel;                // entries deleted by the last mark pass
    unsigned int ntask;
    unsigned int ndel2;               // second deletion count (DELTA iterations run ahead)
    unsigned int ntask2;
    unsigned int ndel3;               // third deletion count (DELTA iterations run ahead)
    unsigned int ntask3;
    unsigned int wl_size;             // worklist length
    unsigned int min_sup;             // minimum live support (JUMP)
    unsigned int live;                // entries kept by the last compaction
    unsigned int choice;              // candidate picked by the cost model
    unsigned int dtasks;              // DELTA tasks of the pending deletions (task-list kernel)
    unsigned int dticket;             // DELTA task list: workgroups done (reset by the last)
    unsigned int void_del;            // deletions of void DELTA iterations (stays 0)
    unsigned int mark_bad;            // entries the DELTA mark found deleted already (stays 0)
    unsigned int mticket;             // JUMP minimum: workgroups done (reset by the last)
    unsigned long long delta_elems;   // adjacency entries read by DELTA updates
    unsigned long long delta_tris;    // destroyed triangles removed by DELTA updates
    unsigned long long delta_win;     // longer-list window entries read by DELTA updates
    unsigned long long delta_live;    // candidates of DELTA updates whose edge was alive
};

static_assert(offsetof(DevCounters, ndel) % 8 == 0 && offsetof(DevCounters, ntask) == offsetof(DevCounters, ndel) + 4 &&
                  offsetof(DevCounters, ndel2) % 8 == 0 && offsetof(DevCounters, ntask2) == offsetof(DevCounters, ndel2) + 4 &&
                  offsetof(DevCounters, ndel3) % 8 == 0 && offsetof(DevCounters, ntask3) == offsetof(DevCounters, ndel3) + 4,
              "a deletion count and its task total form one aligned 64-bit pair");

// Arguments of the support-count kernels.
// Statistics every wave of a count launch adds to go to one of kStripes
// 128-byte lines (by wave index) instead of one address: tens of thousands of
// atomics on one address are served one at a time and hold back the launch.
// dev::stripes_fold adds the lines to the counters and clears them.
constexpr uint32_t kStripes = 512;
constexpr uint32_t kStripeWords = 16;  // 64-bit words per line: merge_elems, recount_edges
constexpr size_t kStripeBytes = (size_t)kStripes * kStripeWords * 8;

struct CountArgs {
    const uint32_t* ptr;
    const uint32_t* src;
    const uint32_t* dst;
    uint32_t* sup;
    const uint32_t* wl;   // nullptr: items are entry positions [0, nitems)
    uint32_t nitems;
    uint32_t chunk;       // items per wave task (multiple of 64)
    uint32_t cap_lds;     // LDS entries per wave for the staged row
    uint32_t cap;         // k - 2 (undirected early exit)
    uint32_t tile_size;   // vertices per tile (tiled format)
    uint32_t ytiles;      // x-tile classes counted in parallel (tiled); 1 otherwise
    uint32_t query_factor;// query the shorter list when it is this many times shorter
    uint32_t prefilter;   // undirected COO+CSR stored-length pre-filter (huc:20)
    DevCounters* ctr;
    unsigned long long* stripes;  // kStripes lines of the count statistics
    const uint32_t* rowlen;  // tiled undirected: stored row length per vertex (pre-filter); else null
};

// Device-resident task list of a batched decomposition: task t covers the
// entries [desc[t].x, desc[t].y) of graph desc[t].z of the batch.  Both
// fields are read on the device, so one captured launch serves whatever list
// the previous step built.
struct TaskList {
    const uint4* desc;
    const uint32_t* ntasks;
};

// Tiled COO+CSR format (TiledCOOCSRGraph, graph.h:67-76): tiles of S
// vertices, destination tiles grouped into Y classes (class(t) = t mod Y, the
// x-tile groups the reference's 2-D grid counts in parallel, kernel_ddt.cu:28,
// kernel_dut.cu:28).  Entries are stored block by block -- block (i, y) holds
// the entries from source tile i into tiles of class y -- and within a block
// by source, then destination: a CSR per block, whose rows are the block rows
// ((i * Y + y) * S + u mod S) of one pointer array of tiled_rows(n, S, Y) + 1
// entries.  With Y equal to the tiles per dimension this is the reference's
// tileSrcPtr ((tile(u) * T + tile(v)) * S + u mod S, graph.cu:674-676); with
// fewer classes it has T * Y * S + 1 pointers instead of T^2 * S + 1, so fine
// tiles (T = 1024 at 2^22 vertices) need no 2^32-entry array.
__host__ __device__ __forceinline__ uint32_t tiled_row(uint32_t u, uint32_t cls, uint32_t S, uint32_t Y) {
    return ((u / S) * Y + cls) * S + u % S;
}
inline uint64_t tiled_rows(uint32_t n, uint32_t S, uint32_t Y) {
    return (((uint64_t)n + S - 1) / S) * Y * S;
}

// Kernel launchers (kt_device.hip).  All asynchronous on `stream`.
namespace dev {
hipError_t fill_u32(uint32_t* p, uint32_t value, uint64_t count, hipStream_t s);
// Stored row length of every vertex of a tiled layout (the sum of its Y block
// rows): out[u] for u < n.
hipError_t tiled_row_totals(const uint32_t* ptr, uint32_t n, uint32_t S, uint32_t Y, uint32_t* out, hipStream_t s);
hipError_t stripes_fold(unsigned long long* stripes, DevCounters* ctr, hipStream_t s);
hipError_t fill_u8(uint8_t* p, uint8_t value, uint64_t count, hipStream_t s);
hipError_t iota_u32(uint32_t* p, uint64_t count, hipStream_t s);
hipError_t count_supports(const CountArgs& a, bool directed, bool tiled, uint32_t algo,
                          uint32_t block, hipStream_t s);
// Persistent directed COO+CSR count over a batch task list, `blocks` workgroups.
hipError_t count_supports_batch(const CountArgs& a, const TaskList& tl, uint32_t block,
                                uint32_t blocks, hipStream_t s);
hipError_t zero_live_supports(const uint32_t* dst, uint32_t* sup, uint32_t E, hipStream_t s);
hipError_t zero_listed_supports(const uint32_t* wl, uint32_t n, uint32_t* sup, hipStream_t s);
hipError_t mark_deleted(uint32_t* dst, const uint32_t* src, const uint32_t* eid,
                        const uint32_t* sup, const uint8_t* state, uint8_t* nflag,
                        uint32_t* tau, const uint32_t* wl, uint32_t nitems, uint32_t k,
                        bool affected, DevCounters* ctr, hipStream_t s);
hipError_t mark_directly(const uint32_t* src, const uint32_t* dst, uint8_t* state, uint8_t* nflag,
                         uint32_t E, bool directed, hipStream_t s);
// Marks INDIRECTLY entries (directed) and appends every DIRECTLY/INDIRECTLY
// live entry to wl (zeroing its support); undirected: append only.
hipError_t build_worklist(const uint32_t* src, const uint32_t* dst, uint8_t* state,
                          const uint8_t* nflag, uint32_t* sup, uint32_t* wl, uint32_t E,
                          bool directed, DevCounters* ctr, hipStream_t s);
hipError_t min_live_support(const uint32_t* dst, const uint32_t* sup, uint32_t E,
                            DevCounters* ctr, hipStream_t s);
// DELTA support maintenance (kt_delta.hip).  A deleted entry's support is
// KT_SUP_GONE from its mark on, so a support below a threshold is a live
// entry's.
// The deletions are appended to dl, counted in *ndel (zero on entry).  fin_dl
// / fin_nd (optional, at most fin_max entries): a previous deletion list
// whose update is done, finalized (KT_DEAD) by the same kernel.  ldeg
// (optional): live length of every oriented row, decremented at src[e] for
// every deleted entry e.  *bad counts entries below the threshold that were
// flagged already (none while the supports keep the rule above).
struct MarkTasks;
hipError_t mark_dying(uint32_t* dst, const uint32_t* eid, uint32_t* sup, uint32_t* tau,
                      uint32_t* dl, uint32_t n, uint32_t k, uint32_t* snbr, const uint32_t* sa,
                      const uint32_t* sb, unsigned int* ndel, const uint32_t* fin_dl,
                      const uint32_t* fin_nd, uint32_t fin_max, const uint32_t* src, uint32_t* ldeg,
                      unsigned int* bad, hipStream_t s, const MarkTasks* mt = nullptr);
// out[v] = ptr[v + 1] - ptr[v] for v < nv.
hipError_t row_lengths(const uint32_t* ptr, uint32_t nv, uint32_t* out, hipStream_t s);
// Minimum support of the live entries (those neither dead nor dying: the
// others' supports are KT_SUP_GONE).  host_min (optional, device address of
// pinned host memory): the last workgroup to finish stores the minimum there
// (kNoLiveSupport when nothing is live) with a system-scope store.
constexpr uint32_t kNoLiveSupport = 0xFFFFFFFEu;
// One thread at the start of a DELTA phase: ctr->void_del |= *from; *from = 0
// (from: a deletion count checked at the next counter read-back; optional),
// then ctr->ndel = 0 (zero_ndel) and ctr->min_sup = ~0 (reset_min).
hipError_t phase_start(unsigned int* from, DevCounters* ctr, bool zero_ndel, bool reset_min,
                       hipStream_t s);
hipError_t min_live_delta(uint32_t n, const uint32_t* sup, DevCounters* ctr, hipStream_t s,
                          uint32_t* host_min = nullptr);
// Scratch of a DELTA update for up to `cap` pending deletions.
struct DeltaScratch {
    unsigned long long* bstat;  // per k_delta workgroup: elems, win, live, tris
    uint4* meta;        // per deletion: short list [x, y), long list [z, z + w)
    uint32_t* ntask;    // per deletion: first task counted from its tile's first
                        // (delta_prepare), or its first task (a listing mark)
    uint32_t* tsum;     // per 2048-deletion tile, plus one
    uint4* mtmp;        // a listing mark's staged descriptors and task counts,
    uint32_t* ttmp;     // one buffer of list slots per mark workgroup
    uint32_t cap;
};
// A mark pass that lists the DELTA tasks of its deletions itself (mark_dying
// with MarkTasks): each deletion's descriptor and task count are made when it
// is flagged (both endpoints' list starts) and staged per workgroup; a
// workgroup's flush reserves its list slots and its task range with one
// 64-bit atomic on the list's (count, task total) pair, so delta_apply with
// DirectTasks needs no task-list kernel in between.
struct MarkTasks {
    unsigned long long* pair;  // the list's (deletion count, task total), zero on entry
    const uint32_t* sptr;      // adjacency list starts
    DeltaScratch ws;           // meta / ntask receive the list's descriptors and first tasks
};
struct DirectTasks {
    const uint32_t* ntask;           // the list's task total (the pair's high half)
    unsigned long long* zero_pair;   // optional: another list's pair, cleared by the update
    uint32_t* host_cnt;              // optional: device address of pinned memory, receives the count
};
size_t delta_scratch_bytes(uint32_t cap);
void delta_scratch_carve(void* base, uint32_t cap, DeltaScratch& s);
// nd_dev: the number of pending deletions, read on the device (dl[0, *nd_dev));
// nd_max bounds it and sizes the grids.  delta_prepare lists the tasks (one
// kernel), delta_apply is the update proper (k_delta).
// zero_cnt (optional): a counter set to 0 by the same kernel (the next list's);
// host_cnt (optional, device address of pinned host memory): receives *nd_dev;
// lm (optional): list work fused into the same kernel, see ListMark.
struct ListMark;
// sup (optional): the supports, read for each listed deletion -- one whose
// support is KT_SUP_NONE (or any of a list the same kernel flags at k = 3)
// lies in no triangle and gets no tasks.
hipError_t delta_prepare(const uint32_t* dl, const uint32_t* nd_dev, uint32_t nd_max,
                         const uint32_t* src, const uint32_t* dst, const uint32_t* sptr,
                         const DeltaScratch& ws, DevCounters* ctr, hipStream_t s,
                         unsigned int* zero_cnt = nullptr, uint32_t* host_cnt = nullptr,
                         const ListMark* lm = nullptr, const uint32_t* sup = nullptr);
// Work the task-list kernel can do besides the task list: flag the pending
// list when an update made it (its entries' supports fell below k - 2: KT_DYING
// on the entry and both adjacency slots, KT_SUP_GONE, truss number k - 1, live
// row lengths), and finalize the list before it (KT_DEAD), whose update is
// done.  A live entry's slots hold plain ids (sa[e], in src[e]'s list, holds
// dst[e]; sb[e] holds src[e]), so flags are stored, not read-modified-written.
struct ListMark {
    uint32_t* dst;             // the oriented layout (the same array as delta_prepare's dst)
    const uint32_t* eid;
    uint32_t* sup;
    uint32_t* tau;
    uint32_t* snbr;
    const uint32_t* sa;
    const uint32_t* sb;
    uint32_t* ldeg;            // optional: live row lengths
    uint32_t k;                // 0: the pending list is flagged already
    const uint32_t* fin_dl;    // optional: the list to finalize
    const unsigned int* fin_nd;
};
// The update's statistics are summed per workgroup slot; delta_stats_fold
// adds them to ctr->delta_* and clears the slots, delta_stats_reset clears
// them (neither may be captured in a graph that replays before stats are read).
hipError_t delta_stats_reset(const DeltaScratch& ws, hipStream_t s);
hipError_t delta_stats_fold(const DeltaScratch& ws, DevCounters* ctr, hipStream_t s);
// next (optional): the entries whose support the update takes from thr = k - 2
// to k - 3 are appended to next, counted in *next_cnt (zero on entry).  Every
// live entry is at k - 2 or above before the update (the ones below are the
// pending deletions) and supports only fall, so these are exactly the next
// mark pass's deletions, each listed once.  The decrements then return the old
// support: under contention on hub edges' supports that costs the update
// time, so it pays where the update is light.
hipError_t delta_apply(const uint32_t* dl, const uint32_t* nd_dev, uint32_t nd_max,
                       const uint32_t* snbr, const uint32_t* seid, uint32_t* sup,
                       const DeltaScratch& ws, DevCounters* ctr, hipStream_t s,
                       uint32_t thr = 0, uint32_t* next = nullptr, unsigned int* next_cnt = nullptr,
                       const uint32_t* gsmp = nullptr, const DirectTasks* dt = nullptr);
// gsmp[b] = id of adjacency slot 64 b (flags masked), b < ceil(S / 64): the
// block samples delta_apply (gsmp) stages contiguously for long lists; to be
// rebuilt whenever the lists are rebuilt.
hipError_t block_samples(const uint32_t* snbr, uint32_t S, uint32_t* gsmp, hipStream_t s);
hipError_t finalize_deleted(const uint32_t* dl, const uint32_t* nd_dev, uint32_t nd_max,
                            uint32_t* dst, uint32_t* snbr, const uint32_t* sa, const uint32_t* sb,
                            hipStream_t s);
// Stable compaction of the DELTA adjacency after a compaction of the oriented
// layout: drops the dead slots of the S-slot lists into (sptr2, snbr2,
// seid2), S2 kept slots expected, maps every kept slot's entry through remap
// (old position -> new position of the oriented layout) and rewrites sa / sb
// of the live entries from dst (the compacted layout).  tsum needs
// compact_tiles_for(S) + 1 words, xs S + 1.
uint32_t compact_tiles_for(uint32_t n);
hipError_t compact_adjacency(uint32_t nv, uint32_t S, const uint32_t* sptr, const uint32_t* snbr,
                             const uint32_t* seid, const uint32_t* remap, uint32_t* sptr2,
                             uint32_t* snbr2, uint32_t* seid2, uint32_t S2, const uint32_t* dst,
                             uint32_t* sa, uint32_t* sb, uint32_t* tsum, uint32_t* xs, hipStream_t s);
// Stable compaction of live entries into the *_out arrays and row-pointer
// rebuild: tile_sums needs compact_tiles(E) + 1 entries, xs needs E + 1.
uint32_t compact_tiles(uint32_t E);
// Dynamic LDS bytes of one count workgroup.
uint32_t count_lds_bytes(uint32_t cap_lds, uint32_t block, bool directed, bool tiled,
                         uint32_t algo);
hipError_t compact(uint32_t n, uint32_t E, const uint32_t* ptr_in, const uint32_t* src_in,
                   const uint32_t* dst_in, const uint32_t* eid_in, const uint32_t* sup_in,
                   uint32_t* ptr_out, uint32_t* src_out, uint32_t* dst_out, uint32_t* eid_out,
                   uint32_t* sup_out, uint32_t* tile_sums, uint32_t* xs, DevCounters* ctr,
                   hipStream_t s);
// Layout construction on the device (kt_build.hip), each the same arrays as
// its host twin in kt_host.cpp; temporaries are allocated and released inside
// and the stream is synchronized on return.
// Stored layout of m unique pairs (a[e], b[e]) over n vertices in the
// orientation `directedness` (KT_DIRECTED_BY_DEGREE relabels by the given
// degrees, at most max_degree; the others keep ids): ptr (n + 1), src / dst
// (m entries, or 2m for KT_UNDIRECTED; rows by source, destinations
// ascending), new2old (n), pos (m: pair -> its entry in the higher row) and
// *max_row (the longest row).  stage_ms (optional, 3 entries): wall times of
// the relabel, the orientation and the row build, each synchronised.
// classes > 0: the tiled format -- entries in block-row order (tiled_row),
// ptr with tiled_rows(n, tile_size, classes) + 1 entries, *max_row the
// longest block row.
hipError_t build_layout_device(uint32_t n, uint32_t m, const uint32_t* a, const uint32_t* b, int directedness,
                               const uint32_t* degree, uint32_t max_degree, uint32_t* ptr, uint32_t* src,
                               uint32_t* dst, uint32_t* new2old, uint32_t* pos, uint32_t* max_row,
                               hipStream_t s, double* stage_ms = nullptr, uint32_t tile_size = 0,
                               uint32_t classes = 0);
// Input validation of m entries over n vertices (kt_graph_create's checks):
// res.status is KT_OK or the error of the lowest offending entry res.at;
// for a valid input, the unique edges (ua < ub, sorted), every entry's edge
// (uid) and the degrees, in host vectors.
struct InputCheck {
    int status = KT_OK;
    uint64_t at = 0;
};
hipError_t validate_input(uint32_t n, uint32_t m, const uint32_t* src, const uint32_t* dst, bool sym,
                          InputCheck& res, std::vector<uint32_t>& ua, std::vector<uint32_t>& ub,
                          std::vector<uint32_t>& uid, std::vector<uint32_t>& degree, uint32_t& max_degree);
// Full adjacency of an oriented layout of E entries (src < dst each): sptr
// (n + 1), snbr / seid (2E), sa / sb (E: the entry's slot in src's / dst's list).
hipError_t build_adjacency_device(uint32_t n, uint32_t E, const uint32_t* src, const uint32_t* dst, uint32_t* sptr,
                                  uint32_t* snbr, uint32_t* seid, uint32_t* sa, uint32_t* sb, hipStream_t s);
}  // namespace dev

// kt_ktruss_noalloc (kt_noalloc.hip): one k of the directed COO+CSR peel on
// the caller's device layout, peel state in a fixed device array.
int ktruss_noalloc(uint32_t n, uint32_t E, const uint32_t* ptr, const uint32_t* src, uint32_t* dst,
                   uint32_t k, hipStream_t s, uint64_t* out_live);

// the gfx950 device has been initialised (kt_init or a plan)
bool device_initialized();
// makes the engine's device current on the calling thread (initialising it
// on first use)
int ensure_device();

}  // namespace kt

struct kt_graph {
    uint32_t n = 0;
    uint64_t m_input = 0;
    uint64_t m = 0;                 // undirected edges
    std::vector<uint32_t> ua, ub;   // edge endpoints, ua < ub, sorted by (ua, ub)
    std::vector<uint32_t> uid;      // per input entry: its edge
    std::vector<uint32_t> degree;   // undirected degree per vertex
    uint32_t max_degree = 0;
    std::atomic<int> refs{1};       // a plan holds a reference
};
