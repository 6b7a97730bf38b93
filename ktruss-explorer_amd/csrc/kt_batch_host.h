// kt_batch_host.h -- host object of a batched decomposition.
#pragma once

#include <cstdint>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "kt_batch.h"
#include "kt_plan.h"

namespace kt {

struct Batch {
    kt_variant v{};
    std::vector<const kt_graph*> graphs;
    uint32_t G = 0, N = 0, E = 0, max_row = 0;
    std::vector<uint32_t> vbeg, gbeg;                 // G + 1 each
    std::vector<std::vector<uint32_t>> pos_of_edge;  // per graph: edge -> local entry

    DBuf ptr, src, dst0, dst, sup, tau;
    DBuf d_gbeg, gstate, lists, ddesc, adesc, bc;
    DBuf stripes;  // count statistics per stripe line (folded into bc->count at the end)
    bool delta = false;                               // DELTA support maintenance
    DBuf sptr, snbr0, snbr, seid, sa, sb, dl, dl2, dscratch;
    uint32_t parity = 0;  // DELTA steps alternate between (dl, ndel) and (dl2, ndel2)
    dev::DeltaScratch dws{};
    BatchCounters* h_bc = nullptr;
    BatchCounters* h_rb = nullptr;        // two read-back slots: replays run one ahead
    hipEvent_t ev_rb[2] = {nullptr, nullptr};
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr, ev_t0 = nullptr, ev_t1 = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t exec_steps = 0;
    uint32_t num_cu = 256;
    bool timed = false;                               // timing mode
    std::vector<hipEvent_t> tev;                      // count-launch events (pairs)

    kt_batch_tuning tune{};
    uint32_t cap = 64;          // resolved LDS slice
    BatchDev bd{};
    kt_batch_stats stats{};
    bool have_tau = false;
    std::vector<uint32_t> h_tau, h_ginfo;             // downloaded after a run (h_tau on first read)
    bool tau_on_host = false;
    std::vector<kt_batch_tuning> tune_log_cfg;
    std::vector<float> tune_log_ms;

    ~Batch();
    int create(const kt_graph* const* gs, uint32_t count, const kt_variant* var);
    int set_tuning(const kt_batch_tuning& t);
    int decompose();
    int autotune(uint32_t trial_steps, uint32_t repeats);
    int truss_numbers(uint32_t g, uint32_t* out, uint64_t count);
    int graph_info(uint32_t g, uint32_t* info);

   private:
    int resolve(const kt_batch_tuning& in, kt_batch_tuning& out, uint32_t& cap_out) const;
    int apply(const kt_batch_tuning& t);
    int prestep(hipEvent_t c0 = nullptr, hipEvent_t c1 = nullptr);
    int launch_step(hipEvent_t c0 = nullptr, hipEvent_t c1 = nullptr);
    int event_pair(size_t i, hipEvent_t& a, hipEvent_t& b);
    int capture(uint32_t steps);
    int reset();
    int trial(uint32_t steps, float& ms);
    uint32_t stream_blocks() const { return num_cu * 8; }
    uint32_t count_blocks() const;
};

}  // namespace kt

struct kt_batch {
    kt::Batch impl;
};
