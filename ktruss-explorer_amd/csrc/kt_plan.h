// kt_plan.h -- device-resident plan of one variant.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "kt_internal.h"

namespace kt {

struct DBuf {
    void* p = nullptr;
    size_t size = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    int alloc(size_t bytes);
    void release();
    template <class T>
    T* as() const { return (T*)p; }
};

int init_device(int device);
// Makes the engine's device current (initialising it on first use).
int ensure_device();

// Count configuration resolved for a plan's layout.
struct CountCfg {
    uint32_t kernel = KT_COUNT_ROW, block = 512, cap = 64, chunk = 128;
};
const kt_graph* graph_retain(const kt_graph* g);
void graph_release(const kt_graph* g);

struct Plan {
    kt_variant v{};
    const kt_graph* graph = nullptr;
    bool directed = true, tiled = false, affected = false;
    uint32_t n = 0, E0 = 0, E = 0;
    uint32_t rows = 0;  // rows of the stored layout: n, or the tiled format's block rows
    uint64_t live = 0;
    uint32_t tile_size = 1, ytiles = 1;
    uint32_t cap_lds = 64, query_factor = 8;
    uint32_t max_row = 0;
    std::vector<uint32_t> new2old, pos_of_edge;

    DBuf ptr0, src0, dst0;                          // pristine layout
    DBuf ptr[2], src[2], dst[2], eid[2], sup[2];    // working layout, double-buffered
    int cur = 0;
    DBuf state, nflag, wl, tau, tiles, xs, ctr;
    DBuf rowlen;  // tiled undirected: stored row length per vertex of the working layout
    DevCounters* h_ctr = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr, ev_c0 = nullptr, ev_c1 = nullptr;
    // run-ahead DELTA iterations: kAhead queued beyond the one being read
    // back; per slot (update start, update end, count read-back) events and
    // the pinned deletion count
    static constexpr uint32_t kAhead = 1, kAheadSlots = kAhead + 2;
    static constexpr uint32_t kCountPending = 0xFFFFFFFFu;  // h_ra slot not written yet
    // a run-ahead iteration is light (its update lists the next deletions, no
    // mark pass) while the last two counts read back are at most this
    static constexpr uint32_t kLightDeletions = 64;
    hipEvent_t ev_ra[kAheadSlots][3] = {};
    // pinned words: kAheadSlots deletion counts, then the JUMP minimum and the
    // model's choice of a phase start (written by the kernels that make them)
    static constexpr uint32_t kSlotMin = kAheadSlots, kSlotChoice = kAheadSlots + 1;
    uint32_t* h_ra = nullptr;
    uint32_t* d_ra = nullptr;  // h_ra's device address
    // waits for pinned word h_ra[i] to leave kCountPending; a stream that went
    // idle without it is an error
    int poll_pinned(uint32_t i, uint32_t& v);

    std::vector<std::array<uint64_t, KT_ITER_REC>> trace;  // kt_plan_iteration_log records
    double stage_ms[KT_PLAN_STAGES] = {};                    // kt_plan_stage_times
    hipEvent_t ev_m0 = nullptr, ev_m1 = nullptr, ev_a0 = nullptr, ev_a1 = nullptr;  // mark, affected
    kt_stats stats{};
    bool have_tau = false;

    // count configurations: the variant's own, the candidates, the pick mode
    CountCfg fixed_cfg;
    std::vector<CountCfg> cands;
    uint32_t pick = KT_PICK_FIXED;
    // sampled SWEEP (kt_plan_set_sweep_sampling) and the current phase's mode
    uint32_t sweep_every = 1, sweep_offset = 0, sweep_fill = 0;
    uint32_t pick_now = KT_PICK_FIXED;
    CountCfg phase_fixed;
    void begin_phase();
    DBuf model;          // cost-model parameters (KT_PICK_MODEL)
    uint32_t model_C = 0, model_n = 0;
    DBuf feat_acc, ctr_scratch;
    DBuf feat_log;            // phase features, KT_MLP_IN floats per phase of the run
    uint32_t feat_cap = 0;    // phases feat_log holds
    std::vector<uint8_t> feat_on;  // per logged phase: features were computed
    bool phase_has_feat = false;
    bool timing = true;  // hipEvents around the support passes of run-ahead iterations
    int flush_phase_features();
    DBuf stripes, stripes_x;  // count statistics of ctr / of ctr_scratch (never read)
    bool stripes_dirty = false;  // a count added to `stripes` since the last fold
    hipEvent_t ev_sw[2 * KT_MAX_CANDIDATES] = {};
    std::vector<std::array<float, KT_PHASE_REC>> phase_log;

    // DELTA support maintenance: full adjacency lists (built on the device on
    // first use from the pristine oriented layout) and the lists of entries
    // deleted by mark passes whose triangles are still counted in the supports
    // adjacency lists: working pair (sptr/snbr/seid)[acur] of S_cur slots,
    // compacted into the other of the pair; pristine copies for every run
    DBuf sptr_b[2], snbr_b[2], seid_b[2], sa, sb;
    DBuf sptr0, snbr0, seid0, sa0, sb0;
    int acur = 0;
    uint32_t S_cur = 0;
    DBuf dl, dl_alt, dl3, sup_snap, dscratch, ctsum, cxs;  // dl_alt, dl3: run-ahead deletion lists
    DBuf ldeg;  // live length of every oriented row (phase features of DELTA plans)
    DBuf gsmp;  // block samples of the working adjacency lists (dev::block_samples)
    // the rows' live lengths are kept only when phases are picked (features)
    uint32_t* live_rows_ptr() const { return pick == KT_PICK_FIXED ? nullptr : ldeg.as<uint32_t>(); }
    dev::DeltaScratch dws{};
    bool delta_ready = false;
    bool last_pass_delta = false;  // the last support pass was a DELTA update
    uint32_t pending = 0;
    // count of the void iteration that ended the last run-ahead DELTA phase:
    // moved into ctr->void_del by take_void, which every read-back checks
    unsigned int* void_src = nullptr;
    bool delta_capable() const;

    int resolve(const kt_count_config& in, CountCfg& out) const;
    int set_candidates(const kt_count_config* c, uint32_t n);
    int set_pick(uint32_t mode, const float* params, uint32_t nparams, uint32_t C,
                 const float* range);

    ~Plan();
    // Frees the working layout, then the pristine layout, then the peel state
    // and the rest of the device arrays, timing each (ms[3], may be null); the
    // plan can only be destroyed afterwards.
    void release_timed(double* ms);
    int create(const kt_graph* g, const kt_variant* v);
    int set_tiles(uint32_t tile_size_req, uint32_t tiles_per_dim, uint32_t classes);
    int build_layout(const kt_graph* g);
    // kt_plan_autotune
    std::vector<kt_tune_trial> tune_log;
    int autotune(uint32_t k, uint32_t repeats);
    int tune_trial(kt_tune_trial& t, uint32_t k, uint32_t repeats);
    int reset();
    int decompose();
    int ktruss(uint32_t k, std::vector<uint32_t>& out_src, std::vector<uint32_t>& out_dst);
    int truss_numbers(uint32_t* out, uint64_t count);

   private:
    int run_phase(uint32_t& k, bool reuse, uint32_t schedule, bool& exact_out);
    int run_phase_delta(uint32_t& k, bool reuse, uint32_t schedule);
    int ensure_delta();
    int compact_delta_state();
    bool compaction_due() const;  // dead entries are more than a third of the stored ones
    int restore_delta();
    int support_pass(const CountCfg& cfg, uint32_t k, float& ms);
    int run_ahead_delta(uint32_t k, uint32_t& iter, double& phase_ms, bool& ended, bool head = false);
    // also clears ctr->ndel / resets ctr->min_sup in the same one-thread kernel
    int take_void(bool zero_ndel = false, bool reset_min = false);
    // flags: KT_ITER_* of the iteration
    int record_iteration(uint32_t k, uint32_t iter, uint32_t ndel, uint64_t recount, float count_ms,
                         uint32_t flags = 0);
    int sweep_pass(uint32_t k, std::vector<double>& cand_ms, float& best_ms);
    // defer: leave the model's choice on the device (read with the next
    // counter read-back) instead of synchronising for it now
    // live_rows: row features from the live row lengths (DELTA phases keep them)
    // host_choice: device address of pinned memory the choice is also stored to
    int phase_pick(uint32_t k, bool exact, CountCfg& cfg, uint32_t& chosen, bool defer = false,
                   bool live_rows = false, uint32_t* host_choice = nullptr);
    int compact_now();
    // Reads the device counters; fold: add the DELTA update's per-workgroup
    // statistics first (once a run's last update is done).
    int sync_counters(bool fold = false);
    int set_counter(size_t offset, uint32_t value);
    CountArgs count_args(uint32_t k, bool use_wl, uint32_t nitems, const CountCfg& cfg,
                         DevCounters* counters);
    void finish_stats();
};

}  // namespace kt

struct kt_plan {
    kt::Plan impl;
};
