// kt_plan.cpp -- device-resident plans and the peel / decomposition driver.
//
// The host drives one HIP stream per plan.  Per peel iteration it launches
// [zero supports] -> count -> mark, then reads a 32-byte counter block back
// (the reference reads its `changed` flag the same way, kernel_ddc.cu:166-167)
// to decide between compaction (-e), the k == 3 early exit, the affected-edge
// marking and the next iteration.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstring>
#include <deque>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "kt_host.h"
#include "kt_internal.h"
#include "kt_learn.h"
#include "kt_plan.h"

namespace kt {

int DBuf::alloc(size_t bytes) {
    release();
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        p = nullptr;
        (void)hipGetLastError();
        return fail(KT_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes: " +
                                      hipGetErrorString(e));
    }
    size = bytes;
    return KT_OK;
}

void DBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    size = 0;
}

namespace {

std::atomic<bool> g_device_checked{false};  // read by graph creation on any thread
int g_device = 0;

int next_pow2(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return (int)p;
}

}  // namespace

int init_device(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(KT_ERR_NO_DEVICE, "no HIP device visible");
    }
    if (device < 0 || device >= count) return fail(KT_ERR_ARG, "device index out of range");
    KT_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    KT_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(KT_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                          ", this build targets gfx950 (MI355X)");
    g_device = device;
    g_device_checked.store(true);
    return KT_OK;
}

bool device_initialized() { return g_device_checked.load(); }

int ensure_device() {
    if (g_device_checked) return hipSetDevice(g_device) == hipSuccess ? KT_OK
                                                                      : fail(KT_ERR_HIP, "hipSetDevice");
    return init_device(0);
}

Plan::~Plan() {
    if (h_ctr) (void)hipHostFree(h_ctr);
    if (h_ra) (void)hipHostFree(h_ra);
    for (auto& slot : ev_ra)
        for (auto& ev : slot)
            if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : ev_sw)
        if (ev) (void)hipEventDestroy(ev);
    for (auto* ev : {&ev_start, &ev_stop, &ev_c0, &ev_c1, &ev_m0, &ev_m1, &ev_a0, &ev_a1})
        if (*ev) (void)hipEventDestroy(*ev);
    if (stream) (void)hipStreamDestroy(stream);
    if (graph) graph_release(graph);
}

void Plan::release_timed(double* ms) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    auto lap = [&](int i) {
        const auto t1 = clk::now();
        if (ms) ms[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        t0 = t1;
    };
    if (stream) (void)hipStreamSynchronize(stream);
    for (int b = 0; b < 2; ++b)
        for (DBuf* x : {&ptr[b], &src[b], &dst[b], &eid[b], &sup[b]}) x->release();
    lap(0);
    for (DBuf* x : {&ptr0, &src0, &dst0}) x->release();
    lap(1);
    for (DBuf* x : {&state, &nflag, &wl, &tau, &tiles, &xs, &ctr, &rowlen, &feat_acc, &ctr_scratch, &feat_log, &stripes,
                    &stripes_x, &model, &sa, &sb, &sptr0, &snbr0, &seid0, &sa0, &sb0, &dl, &dl_alt, &dl3,
                    &sup_snap, &dscratch, &ctsum, &cxs, &ldeg})
        x->release();
    for (int b = 0; b < 2; ++b)
        for (DBuf* x : {&sptr_b[b], &snbr_b[b], &seid_b[b]}) x->release();
    delta_ready = false;
    lap(2);
}

// Tile size (tile_size_req, else ceil(n / tiles_per_dim), graph.cu:580) and
// destination classes (at most the tiles) of a tiled plan, and the rows of
// its stored layout.
int Plan::set_tiles(uint32_t tile_size_req, uint32_t tiles_per_dim, uint32_t classes) {
    if (tiled) {
        uint32_t ts = tile_size_req;
        if (ts == 0) {
            if (tiles_per_dim == 0) return fail(KT_ERR_ARG, "tiles_per_dim must be positive");
            ts = (n + tiles_per_dim - 1) / tiles_per_dim;
        }
        if (ts == 0) ts = 1;
        const uint32_t ntiles = (n + ts - 1) / ts;
        const uint32_t y = std::max<uint32_t>(1, std::min<uint32_t>(classes ? classes : 1, std::max<uint32_t>(ntiles, 1)));
        const uint64_t r = tiled_rows(n, ts, y);
        if (r >= 0xFFFFFFFFull)
            return fail(KT_ERR_CAPACITY, "tiled format: tiles x parallel tiles x tile size must stay below 2^32");
        tile_size = ts;
        ytiles = y;
        rows = (uint32_t)r;
    } else {
        tile_size = n ? n : 1;
        ytiles = 1;
        rows = n;
    }
    return KT_OK;
}

// The stored layout (pristine copy) of the plan's orientation and format,
// built on the device from the graph's edges; its row-pointer arrays are
// sized for the current rows.
int Plan::build_layout(const kt_graph* g) {
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); };
    int rc;
    const size_t Pb = ((size_t)n + 1) * sizeof(uint32_t);
    const size_t Rb = ((size_t)rows + 1) * sizeof(uint32_t);  // row pointers of the stored layout
    if (ptr0.size != Rb) {
        for (DBuf* x : {&ptr0, &ptr[0], &ptr[1]}) {
            x->release();
            if ((rc = x->alloc(Rb))) return rc;
        }
    }
    if (tiled && !directed && !rowlen.p && (rc = rowlen.alloc(std::max<size_t>((size_t)n * 4, 4)))) return rc;
    const uint32_t m = (uint32_t)g->m;
    const size_t Mb = (size_t)std::max<uint32_t>(m, 1) * sizeof(uint32_t);
    DBuf da, db, ddeg, dn2o, dpos, dmax;
    auto t0 = clk::now();
    if ((rc = da.alloc(Mb)) || (rc = db.alloc(Mb)) || (rc = ddeg.alloc(std::max<size_t>(Pb - 4, 4))) ||
        (rc = dn2o.alloc(std::max<size_t>(Pb - 4, 4))) || (rc = dpos.alloc(Mb)) || (rc = dmax.alloc(4)))
        return rc;
    if (m) {
        KT_HIP_TRY(hipMemcpyAsync(da.p, g->ua.data(), (size_t)m * 4, hipMemcpyHostToDevice, stream));
        KT_HIP_TRY(hipMemcpyAsync(db.p, g->ub.data(), (size_t)m * 4, hipMemcpyHostToDevice, stream));
    }
    if (n) KT_HIP_TRY(hipMemcpyAsync(ddeg.p, g->degree.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream));
    KT_HIP_TRY(hipStreamSynchronize(stream));
    stage_ms[1] = ms_since(t0);
    double build_ms[3] = {0, 0, 0};  // relabel, orientation, rows
    KT_HIP_TRY(dev::build_layout_device(n, m, da.as<uint32_t>(), db.as<uint32_t>(), v.directedness,
                                        ddeg.as<uint32_t>(), g->max_degree, ptr0.as<uint32_t>(),
                                        src0.as<uint32_t>(), dst0.as<uint32_t>(), dn2o.as<uint32_t>(),
                                        dpos.as<uint32_t>(), dmax.as<uint32_t>(), stream, build_ms,
                                        tiled ? tile_size : 0, tiled ? ytiles : 0));
    stage_ms[0] = build_ms[0];
    stage_ms[3] = build_ms[1];
    stage_ms[4] = build_ms[2];
    t0 = clk::now();
    new2old.resize(n);
    pos_of_edge.resize(m);
    if (n) KT_HIP_TRY(hipMemcpy(new2old.data(), dn2o.p, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (m) KT_HIP_TRY(hipMemcpy(pos_of_edge.data(), dpos.p, (size_t)m * 4, hipMemcpyDeviceToHost));
    KT_HIP_TRY(hipMemcpy(&max_row, dmax.p, 4, hipMemcpyDeviceToHost));
    stage_ms[4] += ms_since(t0);  // the rows' maps back to the edges
    t0 = clk::now();
    for (DBuf* b : {&da, &db, &ddeg, &dn2o, &dpos, &dmax}) b->release();
    stage_ms[5] = ms_since(t0);
    have_tau = false;
    return KT_OK;
}

int Plan::create(const kt_graph* g, const kt_variant* var) {
    int rc = ensure_device();
    if (rc) return rc;
    v = *var;
    if (v.directedness < 0 || v.directedness > 2 || v.format < 0 || v.format > 1 || v.recount < 0 ||
        v.recount > 1)
        return fail(KT_ERR_ARG, "variant enum out of range");
    if (v.block_size == 0) v.block_size = 512;
    if (v.block_size % 64 || v.block_size > 1024) return fail(KT_ERR_ARG, "block_size must be a multiple of 64 up to 1024");
    if (v.chunk == 0) v.chunk = 128;
    if (v.chunk % 64) return fail(KT_ERR_ARG, "chunk must be a multiple of 64");
    directed = v.directedness != KT_UNDIRECTED;
    tiled = v.format == KT_TILED_COOCSR;
    affected = v.recount == KT_RECOUNT_AFFECTED;
    // the stored layout of every orientation is built on the device
    // (kt_build.hip); tiled formats use the same rows
    {
        const uint64_t stored = directed ? g->m : 2 * g->m;
        if (stored >= (1ull << 32) - 1) return fail(KT_ERR_ARG, "too many stored entries");
        n = g->n;
        E0 = (uint32_t)stored;
    }
    if ((rc = set_tiles(v.tile_size, v.tiles_per_dim, v.num_parallel_tiles))) return rc;
    query_factor = 8;

    KT_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto* ev : {&ev_start, &ev_stop, &ev_c0, &ev_c1, &ev_m0, &ev_m1, &ev_a0, &ev_a1})
        KT_HIP_TRY(hipEventCreate(ev));
    KT_HIP_TRY(hipHostMalloc((void**)&h_ctr, sizeof(DevCounters), hipHostMallocDefault));
    std::memset(h_ctr, 0, sizeof(DevCounters));
    for (auto& ev : ev_sw) KT_HIP_TRY(hipEventCreate(&ev));
    for (auto& slot : ev_ra)
        for (auto& ev : slot) KT_HIP_TRY(hipEventCreate(&ev));
    // fine-grained: a kernel's system-scope store lands in host memory
    KT_HIP_TRY(hipHostMalloc((void**)&h_ra, (kAheadSlots + 2) * sizeof(uint32_t),
                             hipHostMallocMapped | hipHostMallocCoherent));
    KT_HIP_TRY(hipHostGetDevicePointer((void**)&d_ra, h_ra, 0));

    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); };
    std::fill(stage_ms, stage_ms + KT_PLAN_STAGES, 0.0);
    auto t0 = clk::now();
    const size_t Eb = (size_t)std::max<uint32_t>(E0, 1) * sizeof(uint32_t);
    if ((rc = src0.alloc(Eb)) || (rc = dst0.alloc(Eb))) return rc;
    for (int b = 0; b < 2; ++b)
        if ((rc = src[b].alloc(Eb)) || (rc = dst[b].alloc(Eb)) || (rc = eid[b].alloc(Eb)) || (rc = sup[b].alloc(Eb)))
            return rc;
    if ((rc = state.alloc(std::max<uint32_t>(E0, 1))) || (rc = nflag.alloc(std::max<uint32_t>(n, 1))) ||
        (rc = wl.alloc(Eb)) || (rc = tau.alloc(Eb)) ||
        (rc = tiles.alloc(((size_t)dev::compact_tiles(E0) + 2) * sizeof(uint32_t))) ||
        (rc = xs.alloc(Eb + sizeof(uint32_t))) || (rc = ctr.alloc(sizeof(DevCounters))) ||
        (rc = ctr_scratch.alloc(sizeof(DevCounters))) || (rc = stripes.alloc(kStripeBytes)) ||
        (rc = stripes_x.alloc(kStripeBytes)) ||
        (rc = feat_acc.alloc((size_t)KT_FEAT_BLOCKS * KT_FEAT_ACC * sizeof(unsigned long long))))
        return rc;
    stage_ms[2] = ms_since(t0);  // the peel state (Info, info.cu:22-35) and the layouts
    if ((rc = build_layout(g))) return rc;
    KT_HIP_TRY(dev::fill_u32(stripes.as<uint32_t>(), 0, kStripeBytes / 4, stream));
    KT_HIP_TRY(dev::fill_u32(stripes_x.as<uint32_t>(), 0, kStripeBytes / 4, stream));
    KT_HIP_TRY(hipStreamSynchronize(stream));
    // LDS slice per wave: the largest power of two up to the longest row that
    // keeps a workgroup within 64 KiB
    {
        kt_count_config c{v.count_kernel, v.block_size, v.cap_lds, v.chunk};
        if ((rc = resolve(c, fixed_cfg))) return rc;
        v.count_kernel = fixed_cfg.kernel;
        cap_lds = fixed_cfg.cap;
    }
    graph = graph_retain(g);
    return KT_OK;
}

int Plan::sync_counters(bool fold) {
    if (stripes_dirty) {
        KT_HIP_TRY(dev::stripes_fold(stripes.as<unsigned long long>(), ctr.as<DevCounters>(), stream));
        stripes_dirty = false;
    }
    if (fold && delta_ready) KT_HIP_TRY(dev::delta_stats_fold(dws, ctr.as<DevCounters>(), stream));
    KT_HIP_TRY(hipMemcpyAsync(h_ctr, ctr.p, sizeof(DevCounters), hipMemcpyDeviceToHost, stream));
    KT_HIP_TRY(hipStreamSynchronize(stream));
    if (h_ctr->void_del) return fail(KT_ERR_STATE, "void DELTA iteration deleted entries");
    if (h_ctr->mark_bad) return fail(KT_ERR_STATE, "the DELTA mark found a deleted entry below the threshold");
    return KT_OK;
}

int Plan::set_counter(size_t offset, uint32_t value) {
    KT_HIP_TRY(dev::fill_u32((uint32_t*)((char*)ctr.p + offset), value, 1, stream));
    return KT_OK;
}

int Plan::reset() {
    const size_t Rb = ((size_t)rows + 1) * sizeof(uint32_t);
    cur = 0;
    E = E0;
    live = E0;
    KT_HIP_TRY(hipMemcpyAsync(ptr[0].p, ptr0.p, Rb, hipMemcpyDeviceToDevice, stream));
    if (tiled && !directed)
        KT_HIP_TRY(dev::tiled_row_totals(ptr0.as<uint32_t>(), n, tile_size, ytiles, rowlen.as<uint32_t>(), stream));
    if (E0) {
        KT_HIP_TRY(hipMemcpyAsync(src[0].p, src0.p, (size_t)E0 * 4, hipMemcpyDeviceToDevice, stream));
        KT_HIP_TRY(hipMemcpyAsync(dst[0].p, dst0.p, (size_t)E0 * 4, hipMemcpyDeviceToDevice, stream));
    }
    KT_HIP_TRY(dev::iota_u32(eid[0].as<uint32_t>(), E0, stream));
    KT_HIP_TRY(dev::fill_u32(tau.as<uint32_t>(), 0, E0, stream));
    KT_HIP_TRY(dev::fill_u32(ctr.as<uint32_t>(), 0, sizeof(DevCounters) / 4, stream));
    KT_HIP_TRY(dev::fill_u32(stripes.as<uint32_t>(), 0, kStripeBytes / 4, stream));
    stripes_dirty = false;
    if (delta_ready) {
        const int rc = restore_delta();
        if (rc) return rc;
    }
    pending = 0;
    void_src = nullptr;
    KT_HIP_TRY(hipStreamSynchronize(stream));
    std::memset(h_ctr, 0, sizeof(DevCounters));
    trace.clear();
    phase_log.clear();
    feat_on.clear();
    stats = kt_stats{};
    have_tau = false;
    return KT_OK;
}

int Plan::compact_now() {
    const int o = 1 - cur;
    stats.compactions++;
    stats.compact_elems += E;
    KT_HIP_TRY(dev::compact(rows, E, ptr[cur].as<uint32_t>(), src[cur].as<uint32_t>(),
                            dst[cur].as<uint32_t>(), eid[cur].as<uint32_t>(),
                            sup[cur].as<uint32_t>(), ptr[o].as<uint32_t>(), src[o].as<uint32_t>(),
                            dst[o].as<uint32_t>(), eid[o].as<uint32_t>(), sup[o].as<uint32_t>(),
                            tiles.as<uint32_t>(), xs.as<uint32_t>(), ctr.as<DevCounters>(), stream));
    int rc = sync_counters();
    if (rc) return rc;
    if (h_ctr->live != live)
        return fail(KT_ERR_STATE, "compaction kept " + std::to_string(h_ctr->live) +
                                      " entries, expected " + std::to_string(live));
    cur = o;
    E = h_ctr->live;
    if (tiled && !directed)
        KT_HIP_TRY(dev::tiled_row_totals(ptr[cur].as<uint32_t>(), n, tile_size, ytiles, rowlen.as<uint32_t>(), stream));
    return KT_OK;
}

int Plan::resolve(const kt_count_config& in, CountCfg& out) const {
    if (in.count_kernel > KT_COUNT_DELTA) return fail(KT_ERR_ARG, "unknown count_kernel");
    if (in.count_kernel == KT_COUNT_DELTA && (!directed || tiled || affected))
        return fail(KT_ERR_ARG, "DELTA needs a directed COO+CSR variant with recount all");
    out.kernel = in.count_kernel == KT_COUNT_AUTO ? (uint32_t)KT_COUNT_ROW : in.count_kernel;
    if (!directed || tiled) out.kernel = (uint32_t)KT_COUNT_ROW;  // the other strategies exist for directed COO+CSR
    out.block = in.block_size ? in.block_size : 512;
    if (out.block % 64 || out.block > 1024)
        return fail(KT_ERR_ARG, "block_size must be a multiple of 64 up to 1024");
    // DELTA counts only in the full pass that opens a run, where 64-entry
    // tasks measured faster (R-MAT 20: 16.2 against 17.8 ms)
    out.chunk = in.chunk ? in.chunk : (out.kernel == KT_COUNT_DELTA ? 64u : 128u);
    if (out.chunk % 64) return fail(KT_ERR_ARG, "chunk must be a multiple of 64");
    if (in.cap_lds) {
        if (in.cap_lds < 64 || in.cap_lds > 4096 || (in.cap_lds & (in.cap_lds - 1)))
            return fail(KT_ERR_ARG, "cap_lds must be a power of two in [64, 4096]");
        out.cap = in.cap_lds;
    } else {
        out.cap = std::min<uint32_t>(4096, std::max<uint32_t>(64, (uint32_t)next_pow2(max_row)));
    }
    // the LDS slice per wave shrinks until a workgroup fits in 64 KiB
    while (out.cap > 64 && dev::count_lds_bytes(out.cap, out.block, directed, tiled, out.kernel) > 65536)
        out.cap >>= 1;
    return KT_OK;
}

int Plan::set_candidates(const kt_count_config* c, uint32_t n) {
    if (n > KT_MAX_CANDIDATES || (n && !c)) return fail(KT_ERR_ARG, "at most 16 candidates");
    std::vector<CountCfg> out(n);
    for (uint32_t i = 0; i < n; ++i) {
        const int rc = resolve(c[i], out[i]);
        if (rc) return rc;
    }
    cands = std::move(out);
    if (pick != KT_PICK_FIXED && cands.empty()) pick = KT_PICK_FIXED;
    return KT_OK;
}

int Plan::set_pick(uint32_t mode, const float* params, uint32_t nparams, uint32_t C,
                   const float* range) {
    if (mode > KT_PICK_MODEL) return fail(KT_ERR_ARG, "unknown pick mode");
    if (mode != KT_PICK_FIXED && cands.empty()) return fail(KT_ERR_STATE, "set candidates first");
    if (mode == KT_PICK_MODEL) {
        if (!params || !range || C != cands.size())
            return fail(KT_ERR_ARG, "the model must have one output per candidate");
        // parameters, then the input range (2 * KT_MLP_IN floats)
        int rc = model.alloc(((size_t)nparams + 2 * KT_MLP_IN) * sizeof(float));
        if (rc) return rc;
        KT_HIP_TRY(hipMemcpy(model.p, params, (size_t)nparams * sizeof(float), hipMemcpyDeviceToDevice));
        KT_HIP_TRY(hipMemcpy(model.as<float>() + nparams, range, 2 * KT_MLP_IN * sizeof(float),
                             hipMemcpyDeviceToDevice));
        model_C = C;
        model_n = nparams;
    }
    pick = mode;
    return KT_OK;
}

CountArgs Plan::count_args(uint32_t k, bool use_wl, uint32_t nitems, const CountCfg& cfg,
                           DevCounters* counters) {
    CountArgs a;
    a.ptr = ptr[cur].as<uint32_t>();
    a.src = src[cur].as<uint32_t>();
    a.dst = dst[cur].as<uint32_t>();
    a.sup = sup[cur].as<uint32_t>();
    a.wl = use_wl ? wl.as<uint32_t>() : nullptr;
    a.nitems = nitems;
    a.chunk = cfg.chunk;
    a.cap_lds = cfg.cap;
    a.cap = k - 2;
    a.tile_size = tile_size;
    a.ytiles = ytiles;
    a.query_factor = query_factor;
    a.prefilter = (!directed && !tiled) ? 1u : 0u;  // kernel_duc.cu:27; kernel_dut.cu has none
    // tiled undirected: the exact stored-length pre-filter of kernel_duc.cu:27
    // (rows shorter than k-2 hold fewer than k-2 triangles), from stored row lengths
    a.rowlen = tiled && !directed ? rowlen.as<uint32_t>() : nullptr;
    a.ctr = counters;
    if (counters == ctr.as<DevCounters>()) {
        a.stripes = stripes.as<unsigned long long>();
        stripes_dirty = true;
    } else {
        a.stripes = stripes_x.as<unsigned long long>();
    }
    return a;
}

// Phase features of the surviving graph (supports used only when `exact`)
// and, in MODEL mode, the cost model's pick among the candidates.
int Plan::phase_pick(uint32_t k, bool exact, CountCfg& cfg, uint32_t& chosen, bool defer, bool live_rows,
                     uint32_t* host_choice) {
    cfg = phase_fixed;
    chosen = 0;
    phase_has_feat = false;
    if (pick_now == KT_PICK_FIXED) return KT_OK;  // no pick to make: the phase log gets no features
    // the phase's features stay on the device (feat_log, one row per phase)
    // until the run ends
    const size_t ph = phase_log.size();
    if (ph >= feat_cap) {
        // phases without a pick leave the buffer as it is, so row ph may lie
        // more than one doubling beyond it
        uint32_t cap = std::max<uint32_t>(256, 2 * feat_cap);
        while (cap <= ph) cap *= 2;
        DBuf grown;
        int rc = grown.alloc((size_t)cap * KT_MLP_IN * sizeof(float));
        if (rc) return rc;
        if (feat_cap)
            KT_HIP_TRY(hipMemcpyAsync(grown.p, feat_log.p, (size_t)feat_cap * KT_MLP_IN * sizeof(float),
                                      hipMemcpyDeviceToDevice, stream));
        KT_HIP_TRY(hipStreamSynchronize(stream));
        std::swap(grown.p, feat_log.p);
        std::swap(grown.size, feat_log.size);
        feat_cap = cap;
    }
    const bool by_model = pick_now == KT_PICK_MODEL;
    KT_HIP_TRY(dev::phase_features(ptr[cur].as<uint32_t>(), live_rows ? ldeg.as<uint32_t>() : nullptr,
                                   dst[cur].as<uint32_t>(), sup[cur].as<uint32_t>(), rows, E, E0, k, exact,
                                   feat_acc.as<unsigned long long>(), feat_log.as<float>() + ph * KT_MLP_IN,
                                   by_model ? model.as<float>() : nullptr, model_C,
                                   by_model ? model.as<float>() + model_n : nullptr,
                                   &ctr.as<DevCounters>()->choice, stream, by_model ? host_choice : nullptr));
    phase_has_feat = true;
    if (by_model && !defer) {
        const int rc = sync_counters();
        if (rc) return rc;
        chosen = std::min<uint32_t>(h_ctr->choice, (uint32_t)cands.size() - 1);
        cfg = cands[chosen];
    }
    return KT_OK;
}

// One phase at k: the peel loop of kernel_ddc.cu:116-214 (and its tiled /
// undirected twins) on the current graph.
// The phase's pick mode and fixed configuration: a sampled SWEEP times every
// candidate only in the phases of index sweep_offset modulo sweep_every and
// runs candidate sweep_fill alone in the others.
void Plan::begin_phase() {
    pick_now = pick;
    phase_fixed = fixed_cfg;
    if (pick == KT_PICK_SWEEP && sweep_every > 1 && phase_log.size() % sweep_every != sweep_offset &&
        !cands.empty()) {
        pick_now = KT_PICK_FIXED;
        phase_fixed = cands[std::min<size_t>(sweep_fill, cands.size() - 1)];
    }
}

int Plan::run_phase(uint32_t& k, bool reuse, uint32_t schedule, bool& exact_out) {
    begin_phase();
    const bool jump = directed && (schedule & KT_SCHED_JUMP);
    uint32_t iter = 0;
    bool has_del = false, broke = false, changed = false;
    bool use_wl = false;   // false: items are all stored entries [0, E)
    uint32_t nitems = 0;   // worklist length when use_wl
    int rc;
    stats.phases++;
    // initInfoOnDevice (info.cu:55-60): every entry directly affected
    if (affected) KT_HIP_TRY(dev::fill_u8(state.as<uint8_t>(), KT_DIRECTLY_AFFECTED, E, stream));
    // phase features (exact supports only when handed over by REUSE) and the pick
    const uint32_t nc = (uint32_t)cands.size();
    std::array<float, KT_PHASE_REC> rec;
    rec.fill(-1.f);
    CountCfg cfg = phase_fixed;
    uint32_t chosen = 0;
    if ((rc = phase_pick(k, reuse, cfg, chosen))) return rc;
    std::vector<double> cand_ms(nc, 0.0);
    double phase_ms = 0.0;
    do {
        if ((rc = set_counter(offsetof(DevCounters, ndel), 0))) return rc;
        if (affected) KT_HIP_TRY(dev::fill_u8(nflag.as<uint8_t>(), KT_NOT_AFFECTED, n, stream));
        bool counted = false;
        const uint64_t rec_before = h_ctr->recount_edges;
        if (!(iter == 0 && reuse)) {
            const uint32_t items = use_wl ? nitems : E;
            if (pick_now == KT_PICK_SWEEP) {
                // every candidate from zeroed supports; the last one's result is kept
                for (uint32_t c = 0; c < nc; ++c) {
                    if (use_wl) KT_HIP_TRY(dev::zero_listed_supports(wl.as<uint32_t>(), nitems, sup[cur].as<uint32_t>(), stream));
                    else KT_HIP_TRY(dev::zero_live_supports(dst[cur].as<uint32_t>(), sup[cur].as<uint32_t>(), E, stream));
                    CountArgs a = count_args(k, use_wl, items, cands[c],
                                             c + 1 == nc ? ctr.as<DevCounters>() : ctr_scratch.as<DevCounters>());
                    KT_HIP_TRY(hipEventRecord(ev_sw[2 * c], stream));
                    KT_HIP_TRY(dev::count_supports(a, directed, tiled, cands[c].kernel, cands[c].block, stream));
                    KT_HIP_TRY(hipEventRecord(ev_sw[2 * c + 1], stream));
                }
            } else {
                if (!use_wl) KT_HIP_TRY(dev::zero_live_supports(dst[cur].as<uint32_t>(), sup[cur].as<uint32_t>(), E, stream));
                CountArgs a = count_args(k, use_wl, items, cfg, ctr.as<DevCounters>());
                KT_HIP_TRY(hipEventRecord(ev_c0, stream));
                KT_HIP_TRY(dev::count_supports(a, directed, tiled, cfg.kernel, cfg.block, stream));
                KT_HIP_TRY(hipEventRecord(ev_c1, stream));
            }
            counted = true;
            stats.count_passes++;
            stats.live_in_passes += live;
        }
        if (iter == 0 && jump) {
            if ((rc = set_counter(offsetof(DevCounters, min_sup), 0xFFFFFFFFu))) return rc;
            KT_HIP_TRY(dev::min_live_support(dst[cur].as<uint32_t>(), sup[cur].as<uint32_t>(), E,
                                             ctr.as<DevCounters>(), stream));
            if ((rc = sync_counters())) return rc;
            const uint32_t s = h_ctr->min_sup;
            if (s != 0xFFFFFFFFu && s + 3 > k) k = s + 3;
        }
        if (timing) KT_HIP_TRY(hipEventRecord(ev_m0, stream));
        KT_HIP_TRY(dev::mark_deleted(dst[cur].as<uint32_t>(), src[cur].as<uint32_t>(),
                                     eid[cur].as<uint32_t>(), sup[cur].as<uint32_t>(),
                                     state.as<uint8_t>(), nflag.as<uint8_t>(), tau.as<uint32_t>(),
                                     use_wl ? wl.as<uint32_t>() : nullptr, use_wl ? nitems : E, k, affected,
                                     ctr.as<DevCounters>(), stream));
        if (timing) KT_HIP_TRY(hipEventRecord(ev_m1, stream));
        if ((rc = sync_counters())) return rc;
        std::array<uint64_t, KT_ITER_REC> irec{};
        if (timing) {
            float mark_ms = 0;
            KT_HIP_TRY(hipEventElapsedTime(&mark_ms, ev_m0, ev_m1));
            irec[5] = (uint64_t)(mark_ms * 1e6);
            irec[10] |= KT_ITER_MARK_TIMED;
        }
        float count_ms = 0;
        if (counted && pick_now == KT_PICK_SWEEP) {
            float best = 0;
            for (uint32_t c = 0; c < nc; ++c) {
                float ms = 0;
                KT_HIP_TRY(hipEventElapsedTime(&ms, ev_sw[2 * c], ev_sw[2 * c + 1]));
                cand_ms[c] += ms;
                if (c == 0 || ms < best) best = ms;
            }
            count_ms = best;
            stats.ms_count += count_ms;
            stats.count_launches++;
        } else if (counted) {
            KT_HIP_TRY(hipEventElapsedTime(&count_ms, ev_c0, ev_c1));
            stats.ms_count += count_ms;
            stats.count_launches++;
        }
        phase_ms += count_ms;
        const uint32_t ndel = h_ctr->ndel;
        live -= ndel;
        stats.iterations++;
        irec[0] = k;
        irec[1] = iter;
        irec[2] = directed ? ndel : ndel / 2;
        irec[3] = h_ctr->recount_edges - rec_before;
        irec[4] = (uint64_t)(count_ms * 1e6);
        if (counted) irec[10] |= KT_ITER_COUNTED | KT_ITER_COUNT_TIMED;  // count events are always recorded here
        if (ndel && k - 1 > stats.k_max) stats.k_max = k - 1;
        changed = ndel > 0;
        if (changed) {
            if (iter < v.num_edge_remove_iter) {
                const auto t0 = std::chrono::steady_clock::now();
                if ((rc = compact_now())) return rc;
                irec[6] = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                              std::chrono::steady_clock::now() - t0).count();
                irec[7] = E;
                irec[10] |= KT_ITER_REMOVED;
                has_del = false;
            } else {
                has_del = true;
            }
            if (k == 3) {  // kernel_ddc.cu:183-186
                trace.push_back(irec);
                broke = true;
                break;
            }
            if (affected) {
                if (timing) KT_HIP_TRY(hipEventRecord(ev_a0, stream));
                KT_HIP_TRY(dev::mark_directly(src[cur].as<uint32_t>(), dst[cur].as<uint32_t>(),
                                              state.as<uint8_t>(), nflag.as<uint8_t>(), E, directed,
                                              stream));
                if ((rc = set_counter(offsetof(DevCounters, wl_size), 0))) return rc;
                KT_HIP_TRY(dev::build_worklist(src[cur].as<uint32_t>(), dst[cur].as<uint32_t>(),
                                               state.as<uint8_t>(), nflag.as<uint8_t>(),
                                               sup[cur].as<uint32_t>(), wl.as<uint32_t>(), E,
                                               directed, ctr.as<DevCounters>(), stream));
                if (timing) KT_HIP_TRY(hipEventRecord(ev_a1, stream));
                if ((rc = sync_counters())) return rc;
                if (timing) {
                    float aff_ms = 0;
                    KT_HIP_TRY(hipEventElapsedTime(&aff_ms, ev_a0, ev_a1));
                    irec[8] = (uint64_t)(aff_ms * 1e6);
                }
                irec[10] |= KT_ITER_AFFECTED;
                nitems = h_ctr->wl_size;
                use_wl = true;
            }
        }
        trace.push_back(irec);
        ++iter;
    } while (changed);
    if (has_del) {  // kernel_ddc.cu:206-212
        const auto t0 = std::chrono::steady_clock::now();
        if ((rc = compact_now())) return rc;
        if (!trace.empty()) {
            trace.back()[9] = (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                                  std::chrono::steady_clock::now() - t0).count();
            trace.back()[10] |= KT_ITER_FINAL_REMOVAL;
        }
    }
    exact_out = !broke && !changed;
    // phase record
    if (pick_now == KT_PICK_SWEEP && nc) {
        chosen = 0;
        for (uint32_t c = 1; c < nc; ++c)
            if (cand_ms[c] < cand_ms[chosen]) chosen = c;
        for (uint32_t c = 0; c < nc; ++c) rec[20 + c] = (float)cand_ms[c];
        phase_ms = cand_ms[chosen];
    }
    rec[0] = (float)k;
    rec[1] = (float)iter + (broke ? 1.f : 0.f);
    rec[2] = pick_now == KT_PICK_FIXED ? -1.f : (float)chosen;
    rec[3] = (float)phase_ms;
    phase_log.push_back(rec);  // features: filled in by flush_phase_features
    feat_on.push_back(phase_has_feat ? 1 : 0);
    stats.k_last = k;
    return KT_OK;
}

int Plan::decompose() {
    int rc = reset();
    if (rc) return rc;
    const bool use_delta = delta_capable();
    if (use_delta && (rc = ensure_delta())) return rc;
    KT_HIP_TRY(hipEventRecord(ev_start, stream));
    uint32_t k = 3;
    bool reuse = false;
    while (live > 0) {
        if (use_delta) {
            if ((rc = run_phase_delta(k, reuse, v.schedule))) return rc;
            reuse = (v.schedule & KT_SCHED_REUSE) != 0;
        } else {
            bool exact = false;
            if ((rc = run_phase(k, reuse, v.schedule, exact))) return rc;
            reuse = directed && !affected && (v.schedule & KT_SCHED_REUSE) && exact;
        }
        ++k;
    }
    if (use_delta && pending) {  // the last pass removed the last edges at k = 3
        KT_HIP_TRY(dev::finalize_deleted(dl.as<uint32_t>(), &ctr.as<DevCounters>()->ndel, pending,
                                         dst[cur].as<uint32_t>(), snbr_b[acur].as<uint32_t>(),
                                         sa.as<uint32_t>(), sb.as<uint32_t>(), stream));
        pending = 0;
    }
    if ((rc = take_void())) return rc;
    KT_HIP_TRY(hipEventRecord(ev_stop, stream));
    KT_HIP_TRY(hipEventSynchronize(ev_stop));
    float ms = 0;
    KT_HIP_TRY(hipEventElapsedTime(&ms, ev_start, ev_stop));
    if ((rc = sync_counters(true))) return rc;
    if ((rc = flush_phase_features())) return rc;
    stats.ms_total = ms;
    finish_stats();
    have_tau = true;
    return KT_OK;
}

// The phase features kept on the device go into the phase log.
int Plan::flush_phase_features() {
    const size_t np = phase_log.size();
    if (!np || !feat_cap) return KT_OK;
    // phases without a pick (a sampled SWEEP's fill phases) do not grow the
    // buffer: it may hold fewer rows than the log, and only those have features
    const size_t nrow = std::min<size_t>(np, feat_cap);
    std::vector<float> f(nrow * KT_MLP_IN);
    KT_HIP_TRY(hipMemcpy(f.data(), feat_log.p, f.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nrow; ++i)
        if (feat_on[i])
            for (int j = 0; j < KT_MLP_IN; ++j) phase_log[i][4 + j] = f[i * KT_MLP_IN + j];
    return KT_OK;
}

void Plan::finish_stats() {
    stats.merge_elems = h_ctr->merge_elems;
    stats.recount_edges = h_ctr->recount_edges;
    stats.delta_elems = h_ctr->delta_elems;
    stats.delta_live = h_ctr->delta_live;
    stats.delta_triangles = h_ctr->delta_tris;
    // DELTA's algorithmic bytes: candidates and longer-list windows read,
    // and per destroyed triangle two slot->entry reads, two flag reads and
    // two support read-modify-writes
    stats.balg_delta_bytes = 4.0 * (double)h_ctr->delta_elems + 4.0 * (double)h_ctr->delta_win +
                             32.0 * (double)h_ctr->delta_tris;
    stats.balg_count_bytes = 4.0 * (double)stats.merge_elems + 24.0 * (double)stats.recount_edges +
                             12.0 * (double)stats.live_in_passes;
}

int Plan::ktruss(uint32_t k, std::vector<uint32_t>& out_src, std::vector<uint32_t>& out_dst) {
    if (k < 2) return fail(KT_ERR_ARG, "k must be at least 2");
    int rc = reset();
    if (rc) return rc;
    KT_HIP_TRY(hipEventRecord(ev_start, stream));
    bool exact = false;
    uint32_t kk = k;
    const bool use_delta = delta_capable();
    if (use_delta && (rc = ensure_delta())) return rc;
    if (live > 0) {
        rc = use_delta ? run_phase_delta(kk, false, 0) : run_phase(kk, false, 0, exact);
        if (rc) return rc;
    }
    if ((rc = take_void())) return rc;
    KT_HIP_TRY(hipEventRecord(ev_stop, stream));
    KT_HIP_TRY(hipEventSynchronize(ev_stop));
    float ms = 0;
    KT_HIP_TRY(hipEventElapsedTime(&ms, ev_start, ev_stop));
    if ((rc = sync_counters(true))) return rc;
    if ((rc = flush_phase_features())) return rc;
    stats.ms_total = ms;
    finish_stats();
    // copy back, coocsr2coo, directed2undirectedCOO, unsort (kernel_d.cu:75-79,
    // kernel_ddc.cu:246-250, kernel_d.cu:57-61, kernel.cu:31-37), each timed
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); };
    auto t0 = clk::now();
    std::vector<uint32_t> s(E), d(E);
    if (E) {
        KT_HIP_TRY(hipMemcpy(s.data(), src[cur].p, (size_t)E * 4, hipMemcpyDeviceToHost));
        KT_HIP_TRY(hipMemcpy(d.data(), dst[cur].p, (size_t)E * 4, hipMemcpyDeviceToHost));
    }
    stage_ms[8] = ms_since(t0);
    // DELTA plans keep tombstones in place (and KT_DYING entries of a last
    // k == 3 pass); the others compacted them away
    t0 = clk::now();
    std::vector<uint32_t> keep;
    keep.reserve(E);
    for (uint32_t e = 0; e < E; ++e) {
        if (d[e] & (KT_DEAD | KT_DYING)) {
            if (!use_delta) return fail(KT_ERR_STATE, "tombstone left after final compaction");
            continue;
        }
        keep.push_back(e);
    }
    const size_t L = keep.size();
    const size_t out = directed ? 2 * L : L;
    out_src.resize(out);
    out_dst.resize(out);
    for (size_t i = 0; i < L; ++i) {
        out_src[i] = s[keep[i]];
        out_dst[i] = d[keep[i]];
    }
    stage_ms[6] = ms_since(t0);
    t0 = clk::now();
    if (directed) {
        for (size_t i = 0; i < L; ++i) {
            out_src[L + i] = out_dst[i];
            out_dst[L + i] = out_src[i];
        }
    }
    stage_ms[7] = ms_since(t0);
    t0 = clk::now();
    for (size_t i = 0; i < out; ++i) {
        out_src[i] = new2old[out_src[i]];
        out_dst[i] = new2old[out_dst[i]];
    }
    stage_ms[9] = ms_since(t0);
    return KT_OK;
}

// ---------------------------------------------------------------------------
// DELTA support maintenance

bool Plan::delta_capable() const {
    if (!directed || tiled || affected) return false;
    if (fixed_cfg.kernel == KT_COUNT_DELTA) return true;
    for (const CountCfg& c : cands)
        if (c.kernel == KT_COUNT_DELTA) return true;
    return false;
}

int Plan::ensure_delta() {
    if (delta_ready) return KT_OK;
    if (n >= KT_DYING) return fail(KT_ERR_CAPACITY, "DELTA needs fewer than 2^30 vertices");
    // adjacency slots (two per entry) are 32-bit indices
    if ((uint64_t)E0 * 2 >= (1ull << 32)) return fail(KT_ERR_CAPACITY, "DELTA needs fewer than 2^31 entries");
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    const size_t Eb = (size_t)std::max<uint32_t>(E0, 1) * 4, Pb = ((size_t)n + 1) * 4;
    const uint32_t ct = dev::compact_tiles_for(2 * E0) + 2;
    for (int b = 0; b < 2; ++b)
        // the update reads whole 64-slot blocks: the lists are padded to the last one
        if ((rc = sptr_b[b].alloc(Pb)) || (rc = snbr_b[b].alloc(2 * Eb + 256)) ||
            (rc = seid_b[b].alloc(2 * Eb)))
            return rc;
    if ((rc = sptr0.alloc(Pb)) || (rc = snbr0.alloc(2 * Eb)) || (rc = seid0.alloc(2 * Eb)) ||
        (rc = sa0.alloc(Eb)) || (rc = sb0.alloc(Eb)) || (rc = sa.alloc(Eb)) || (rc = sb.alloc(Eb)) ||
        (rc = dl.alloc(Eb)) || (rc = dl_alt.alloc(Eb)) || (rc = dl3.alloc(Eb)) || (rc = ldeg.alloc(std::max<size_t>(Pb - 4, 4))) ||
        (rc = gsmp.alloc(((size_t)2 * E0 + 63) / 64 * 4 + 4)) ||
        (rc = dscratch.alloc(dev::delta_scratch_bytes(E0))) ||
        (rc = ctsum.alloc((size_t)ct * 4)) || (rc = cxs.alloc(2 * Eb + 4)))
        return rc;
    dev::delta_scratch_carve(dscratch.p, E0, dws);
    // the full adjacency lists of the pristine oriented layout, built on the device
    KT_HIP_TRY(dev::build_adjacency_device(n, E0, src0.as<uint32_t>(), dst0.as<uint32_t>(), sptr0.as<uint32_t>(),
                                           snbr0.as<uint32_t>(), seid0.as<uint32_t>(), sa0.as<uint32_t>(),
                                           sb0.as<uint32_t>(), stream));
    if ((rc = restore_delta())) return rc;
    KT_HIP_TRY(hipStreamSynchronize(stream));
    delta_ready = true;
    stats.ms_setup = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return KT_OK;
}

// The working adjacency from its pristine copy.
int Plan::restore_delta() {
    const size_t Sb = 2 * (size_t)E0 * 4, Eb = (size_t)E0 * 4, Pb = ((size_t)n + 1) * 4;
    acur = 0;
    S_cur = 2 * E0;
    KT_HIP_TRY(hipMemcpyAsync(sptr_b[0].p, sptr0.p, Pb, hipMemcpyDeviceToDevice, stream));
    if (E0) {
        KT_HIP_TRY(hipMemcpyAsync(snbr_b[0].p, snbr0.p, Sb, hipMemcpyDeviceToDevice, stream));
        KT_HIP_TRY(hipMemcpyAsync(seid_b[0].p, seid0.p, Sb, hipMemcpyDeviceToDevice, stream));
        KT_HIP_TRY(hipMemcpyAsync(sa.p, sa0.p, Eb, hipMemcpyDeviceToDevice, stream));
        KT_HIP_TRY(hipMemcpyAsync(sb.p, sb0.p, Eb, hipMemcpyDeviceToDevice, stream));
    }
    KT_HIP_TRY(dev::delta_stats_reset(dws, stream));
    KT_HIP_TRY(dev::row_lengths(ptr0.as<uint32_t>(), n, ldeg.as<uint32_t>(), stream));
    KT_HIP_TRY(dev::block_samples(snbr_b[0].as<uint32_t>(), S_cur, gsmp.as<uint32_t>(), stream));
    return KT_OK;
}

// With no deletion pending (every entry is live or dead), drop the dead
// entries once they are a third of the stored ones: the oriented layout is
// compacted (stable, rows stay sorted) and the adjacency lists with it, their
// slot -> entry map following the entries to their new positions.  Live
// entries and slots (two per live edge) are counted on the host, so the new
// sizes need no read-back beyond the compaction's own check.
bool Plan::compaction_due() const { return live > 0 && 2 * (uint64_t)E > 3 * live && E > 1024; }

int Plan::compact_delta_state() {
    if (pending || !compaction_due()) return KT_OK;
    const uint32_t S = S_cur;
    int rc = compact_now();  // oriented layout: xs[old position] = new position
    if (rc) return rc;
    const int o = 1 - acur;
    const uint32_t keep = 2 * (uint32_t)live;
    KT_HIP_TRY(dev::compact_adjacency(n, S, sptr_b[acur].as<uint32_t>(), snbr_b[acur].as<uint32_t>(),
                                      seid_b[acur].as<uint32_t>(), xs.as<uint32_t>(),
                                      sptr_b[o].as<uint32_t>(), snbr_b[o].as<uint32_t>(),
                                      seid_b[o].as<uint32_t>(), keep, dst[cur].as<uint32_t>(),
                                      sa.as<uint32_t>(), sb.as<uint32_t>(), ctsum.as<uint32_t>(),
                                      cxs.as<uint32_t>(), stream));
    acur = o;
    S_cur = keep;
    KT_HIP_TRY(dev::block_samples(snbr_b[acur].as<uint32_t>(), S_cur, gsmp.as<uint32_t>(), stream));
    return KT_OK;
}

// Makes the supports exact for the live entries, as configuration `cfg`
// says: the DELTA update of the pending deletions, or a full recount (ROW
// when cfg is DELTA but nothing is pending).  The timed kernel is bracketed
// by ev_c0 / ev_c1.
int Plan::support_pass(const CountCfg& cfg, uint32_t k, float& ms) {
    (void)ms;
    uint32_t* dl_p = dl.as<uint32_t>();
    const uint32_t* nd_dev = &ctr.as<DevCounters>()->ndel;
    if (pending && cfg.kernel == KT_COUNT_DELTA) {
        KT_HIP_TRY(dev::delta_prepare(dl_p, nd_dev, pending, src[cur].as<uint32_t>(),
                                      dst[cur].as<uint32_t>(), sptr_b[acur].as<uint32_t>(), dws,
                                      ctr.as<DevCounters>(), stream, nullptr, nullptr, nullptr,
                                      sup[cur].as<uint32_t>()));
        KT_HIP_TRY(hipEventRecord(ev_c0, stream));
        KT_HIP_TRY(dev::delta_apply(dl_p, nd_dev, pending, snbr_b[acur].as<uint32_t>(),
                                    seid_b[acur].as<uint32_t>(), sup[cur].as<uint32_t>(), dws,
                                    ctr.as<DevCounters>(), stream, 0u, nullptr, nullptr,
                                    gsmp.as<uint32_t>()));
        KT_HIP_TRY(hipEventRecord(ev_c1, stream));
        last_pass_delta = true;
        KT_HIP_TRY(dev::finalize_deleted(dl_p, nd_dev, pending, dst[cur].as<uint32_t>(),
                                         snbr_b[acur].as<uint32_t>(), sa.as<uint32_t>(), sb.as<uint32_t>(),
                                         stream));
        stats.delta_passes++;
        stats.delta_edges += pending;
    } else {
        if (pending)
            KT_HIP_TRY(dev::finalize_deleted(dl_p, nd_dev, pending, dst[cur].as<uint32_t>(),
                                             snbr_b[acur].as<uint32_t>(), sa.as<uint32_t>(), sb.as<uint32_t>(),
                                             stream));
        CountCfg rcfg = cfg;
        if (rcfg.kernel == KT_COUNT_DELTA) rcfg.kernel = KT_COUNT_ROW;
        KT_HIP_TRY(dev::zero_live_supports(dst[cur].as<uint32_t>(), sup[cur].as<uint32_t>(), E, stream));
        CountArgs a = count_args(k, false, E, rcfg, ctr.as<DevCounters>());
        KT_HIP_TRY(hipEventRecord(ev_c0, stream));
        KT_HIP_TRY(dev::count_supports(a, directed, tiled, rcfg.kernel, rcfg.block, stream));
        KT_HIP_TRY(hipEventRecord(ev_c1, stream));
        stats.live_in_passes += live;
        last_pass_delta = false;
    }
    pending = 0;
    return KT_OK;
}

// SWEEP in a DELTA plan: DELTA candidates first (each from the supports of
// the previous pass, restored from a snapshot when there are several), then
// the recount candidates from zeroed supports.  Candidate c is bracketed by
// ev_sw[2c], ev_sw[2c+1].
int Plan::sweep_pass(uint32_t k, std::vector<double>& cand_ms, float& best_ms) {
    (void)cand_ms;
    (void)best_ms;
    const uint32_t nc = (uint32_t)cands.size();
    uint32_t* dl_p = dl.as<uint32_t>();
    const uint32_t* nd_dev = &ctr.as<DevCounters>()->ndel;
    uint32_t ndc = 0;
    for (const CountCfg& c : cands) ndc += c.kernel == KT_COUNT_DELTA;
    const bool had_pending = pending != 0;
    if (had_pending && ndc) {
        if (ndc > 1) {
            if (!sup_snap.p) {
                int rc = sup_snap.alloc((size_t)std::max<uint32_t>(E0, 1) * 4);
                if (rc) return rc;
            }
            KT_HIP_TRY(hipMemcpyAsync(sup_snap.p, sup[cur].p, (size_t)E * 4, hipMemcpyDeviceToDevice, stream));
        }
        bool first = true;
        for (uint32_t c = 0; c < nc; ++c) {
            if (cands[c].kernel != KT_COUNT_DELTA) continue;
            if (!first)
                KT_HIP_TRY(hipMemcpyAsync(sup[cur].p, sup_snap.p, (size_t)E * 4, hipMemcpyDeviceToDevice, stream));
            first = false;
            KT_HIP_TRY(hipEventRecord(ev_sw[2 * c], stream));
            KT_HIP_TRY(dev::delta_prepare(dl_p, nd_dev, pending, src[cur].as<uint32_t>(),
                                          dst[cur].as<uint32_t>(), sptr_b[acur].as<uint32_t>(), dws,
                                          ctr.as<DevCounters>(), stream, nullptr, nullptr, nullptr,
                                          sup[cur].as<uint32_t>()));
            KT_HIP_TRY(dev::delta_apply(dl_p, nd_dev, pending, snbr_b[acur].as<uint32_t>(),
                                        seid_b[acur].as<uint32_t>(), sup[cur].as<uint32_t>(), dws,
                                        ctr.as<DevCounters>(), stream, 0u, nullptr, nullptr,
                                        gsmp.as<uint32_t>()));
            KT_HIP_TRY(hipEventRecord(ev_sw[2 * c + 1], stream));
        }
    }
    if (had_pending)
        KT_HIP_TRY(dev::finalize_deleted(dl_p, nd_dev, pending, dst[cur].as<uint32_t>(),
                                         snbr_b[acur].as<uint32_t>(), sa.as<uint32_t>(), sb.as<uint32_t>(),
                                         stream));
    uint32_t last_recount = nc;
    for (uint32_t c = 0; c < nc; ++c)
        if (!(had_pending && cands[c].kernel == KT_COUNT_DELTA)) last_recount = c;
    for (uint32_t c = 0; c < nc; ++c) {
        if (had_pending && cands[c].kernel == KT_COUNT_DELTA) continue;
        CountCfg rcfg = cands[c];
        if (rcfg.kernel == KT_COUNT_DELTA) rcfg.kernel = KT_COUNT_ROW;
        KT_HIP_TRY(dev::zero_live_supports(dst[cur].as<uint32_t>(), sup[cur].as<uint32_t>(), E, stream));
        CountArgs a = count_args(k, false, E, rcfg,
                                 c == last_recount ? ctr.as<DevCounters>() : ctr_scratch.as<DevCounters>());
        KT_HIP_TRY(hipEventRecord(ev_sw[2 * c], stream));
        KT_HIP_TRY(dev::count_supports(a, directed, tiled, rcfg.kernel, rcfg.block, stream));
        KT_HIP_TRY(hipEventRecord(ev_sw[2 * c + 1], stream));
    }
    pending = 0;
    return KT_OK;
}

// Bookkeeping of one finished peel iteration (its deletions are read back).
int Plan::record_iteration(uint32_t k, uint32_t iter, uint32_t ndel, uint64_t recount, float count_ms,
                           uint32_t flags) {
    live -= ndel;
    stats.iterations++;
    std::array<uint64_t, KT_ITER_REC> r{};
    r[0] = k;
    r[1] = iter;
    r[2] = ndel;
    r[3] = recount;
    r[4] = (uint64_t)(count_ms * 1e6);
    r[10] = flags;
    trace.push_back(r);
    if (ndel && k - 1 > stats.k_max) stats.k_max = k - 1;
    return KT_OK;
}

int Plan::poll_pinned(uint32_t i, uint32_t& v) {
    for (uint32_t spin = 1;; ++spin) {
        if (__atomic_load_n(h_ra + i, __ATOMIC_ACQUIRE) != kCountPending) break;
        if (spin % 256 == 0) {
            const hipError_t q = hipStreamQuery(stream);
            if (q == hipSuccess && __atomic_load_n(h_ra + i, __ATOMIC_ACQUIRE) == kCountPending)
                return fail(KT_ERR_STATE, "a value the device was to store in pinned memory never arrived");
            if (q != hipSuccess && q != hipErrorNotReady) KT_HIP_TRY(q);
        }
        std::this_thread::yield();
    }
    v = h_ra[i];
    return KT_OK;
}

int Plan::take_void(bool zero_ndel, bool reset_min) {
    if (!void_src && !zero_ndel && !reset_min) return KT_OK;
    KT_HIP_TRY(dev::phase_start(void_src, ctr.as<DevCounters>(), zero_ndel, reset_min, stream));
    void_src = nullptr;
    return KT_OK;
}

// The remaining iterations of a DELTA phase at k (k > 3), queued one ahead of
// their read-backs.  Every live entry has support k - 2 or more before an
// update (the smaller ones are the pending deletions), so the next deletions
// are the entries whose support the update takes from k - 2 to k - 3.  An
// iteration is heavy or light, chosen from the last two counts read back:
//   heavy: task list, update, mark (a pass over the stored supports that
//     flags and lists the next deletions and finalizes the ones just
//     updated);
//   light: task list (which also flags the pending list and finalizes the one
//     before it when a light update made them), update (whose decrements
//     return the old support and list the entries that cross, so no pass over
//     the stored supports is needed).  Returning decrements wait for their
//     turn at L2 under contention on hub edges' supports, so heavy updates
//     keep the plain ones.
// The task list also clears the count of the list the iteration fills and
// writes the count of its own list into pinned host memory with a
// system-scope store.  Iteration j's kernels are queued before iteration
// j - 1's count comes back, so the device does not wait for the host
// between iterations; grids are sized from upper bounds and kernels read the
// true counts on the device.  The iteration queued after the one that
// deletes nothing only finishes the last update and finalization (its list
// is empty); it is not an iteration of the peel and is not recorded, and its
// count is checked at the next read-back (take_void) instead of waiting for
// it now.  Returns with `ended` when the phase is over (nothing pending,
// every deletion finalized), else with the last iteration's deletions
// flagged and pending in dl / ctr->ndel and `iter` the next index, for a
// compaction.  On entry `pending` holds the deletions of iteration iter - 1
// (> 0, flagged by the mark pass), or, with `head`, iteration iter's mark
// pass is queued into dl / ctr->ndel and its count is unread.
int Plan::run_ahead_delta(uint32_t k, uint32_t& iter, double& phase_ms, bool& ended, bool head) {
    DevCounters* c = ctr.as<DevCounters>();
    // three deletion lists rotate: iteration r (counted from entry) updates
    // list r % 3 and fills list (r + 1) % 3; list (r + 2) % 3, updated by the
    // iteration before, is the one a task list may finalize; the pending
    // deletions on entry are in dl / ctr->ndel (list 0)
    DBuf* bufs[3] = {&dl, &dl_alt, &dl3};
    uint32_t* lists[3] = {dl.as<uint32_t>(), dl_alt.as<uint32_t>(), dl3.as<uint32_t>()};
    unsigned int* counts[3] = {&c->ndel, &c->ndel2, &c->ndel3};
    // deletions the next update consumes, at most
    uint32_t bound = head ? (uint32_t)std::min<uint64_t>(live, E) : pending;
    uint32_t last_nd = bound;  // deletions of the last recorded iteration
    uint32_t prev_nd = bound;  // and of the one before
    // queued iterations whose counts are unread, oldest first
    struct Queued {
        int slot;
        uint32_t it;
        bool updated;  // an update produced its deletions
    };
    std::deque<Queued> open;
    ended = false;
    uint32_t r = 0;  // iterations queued since entry
    int prev_slot = -1;
    if (head) {  // the first task list writes the head mark's count
        prev_slot = (int)(iter % kAheadSlots);
        __atomic_store_n(h_ra + prev_slot, kCountPending, __ATOMIC_RELEASE);
        open.push_back({prev_slot, iter, false});
    }
    // the last queued iteration was light: the list it filled is not flagged
    // and the list it updated is not finalized
    bool unflagged = false;
    dev::ListMark lm{};
    lm.dst = dst[cur].as<uint32_t>();
    lm.eid = eid[cur].as<uint32_t>();
    lm.sup = sup[cur].as<uint32_t>();
    lm.tau = tau.as<uint32_t>();
    lm.snbr = snbr_b[acur].as<uint32_t>();
    lm.sa = sa.as<uint32_t>();
    lm.sb = sb.as<uint32_t>();
    lm.ldeg = live_rows_ptr();
    // task list of list rr % 3 (its grid follows `hint`: it strides over the
    // true counts), with the flagging and finalization a light iteration left
    auto prepare = [&](uint32_t rr, uint32_t hint, unsigned int* zero_cnt, uint32_t* host_cnt) -> int {
        const uint32_t in = rr % 3, prev = (rr + 2) % 3;
        dev::ListMark m = lm;
        m.k = unflagged ? k : 0u;
        m.fin_dl = unflagged ? lists[prev] : nullptr;
        m.fin_nd = counts[prev];
        KT_HIP_TRY(dev::delta_prepare(lists[in], counts[in], std::max<uint32_t>(hint, 1u),
                                      src[cur].as<uint32_t>(), dst[cur].as<uint32_t>(),
                                      sptr_b[acur].as<uint32_t>(), dws, c, stream, zero_cnt, host_cnt, &m,
                                      sup[cur].as<uint32_t>()));
        return KT_OK;
    };
    // iteration r; the task list's hint is the last count read back (the size
    // of the list it may finalize), the update's grid follows the upper bound
    auto queue = [&](int s) -> int {
        const uint32_t in = r % 3, out = (r + 1) % 3;
        const bool light = std::max(last_nd, prev_nd) <= kLightDeletions;
        // a list a mark made carries its tasks: the update reads them
        // directly and takes over the task-list kernel's other duties
        // (clearing the next list's count, publishing this one's)
        const bool direct = !light && !unflagged;
        __atomic_store_n(h_ra + s, kCountPending, __ATOMIC_RELEASE);  // the successor overwrites it
        uint32_t* host_cnt = prev_slot >= 0 ? d_ra + prev_slot : nullptr;
        if (!direct) {
            const int rc = prepare(r, last_nd, counts[out], host_cnt);
            if (rc) return rc;
        }
        const dev::DirectTasks dt{counts[in] + 1, (unsigned long long*)counts[out], host_cnt};
        if (timing) KT_HIP_TRY(hipEventRecord(ev_ra[s][0], stream));
        KT_HIP_TRY(dev::delta_apply(lists[in], counts[in], bound, snbr_b[acur].as<uint32_t>(),
                                    seid_b[acur].as<uint32_t>(), sup[cur].as<uint32_t>(), dws, c, stream,
                                    light ? k - 2 : 0u, light ? lists[out] : nullptr,
                                    light ? counts[out] : nullptr, gsmp.as<uint32_t>(), direct ? &dt : nullptr));
        if (timing) KT_HIP_TRY(hipEventRecord(ev_ra[s][1], stream));
        if (!light) {
            const dev::MarkTasks mt{(unsigned long long*)counts[out], sptr_b[acur].as<uint32_t>(), dws};
            KT_HIP_TRY(dev::mark_dying(dst[cur].as<uint32_t>(), eid[cur].as<uint32_t>(),
                                       sup[cur].as<uint32_t>(), tau.as<uint32_t>(), lists[out], E, k,
                                       snbr_b[acur].as<uint32_t>(), sa.as<uint32_t>(), sb.as<uint32_t>(),
                                       counts[out], lists[in], counts[in], bound, src[cur].as<uint32_t>(),
                                       live_rows_ptr(), &c->mark_bad, stream, &mt));
        }
        unflagged = light;
        prev_slot = s;
        ++r;
        return KT_OK;
    };
    // deletion count of the i-th queued iteration: written by its successor's
    // task list, or, for the youngest, copied back now
    auto read_count = [&](size_t i, uint32_t& nd) -> int {
        const int s = open[i].slot;
        if (i + 1 < open.size()) {
            // the successor's task list stores the count into pinned memory:
            // wait for it there (an event between the kernels would cost the
            // device a gap)
            return poll_pinned((uint32_t)s, nd);
        } else {
            KT_HIP_TRY(hipMemcpyAsync(h_ra + s, counts[r % 3], sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
            KT_HIP_TRY(hipEventRecord(ev_ra[s][2], stream));
            KT_HIP_TRY(hipEventSynchronize(ev_ra[s][2]));
        }
        nd = h_ra[s];
        return KT_OK;
    };
    // reads the oldest queued iteration back and records it
    auto finish = [&](uint32_t& nd) -> int {
        const Queued q = open.front();
        int rc = read_count(0, nd);
        if (rc) return rc;
        open.pop_front();
        float ms = 0;
        if (q.updated) {
            if (timing) {
                KT_HIP_TRY(hipEventSynchronize(ev_ra[q.slot][1]));
                KT_HIP_TRY(hipEventElapsedTime(&ms, ev_ra[q.slot][0], ev_ra[q.slot][1]));
            }
            stats.ms_delta += ms;
            phase_ms += ms;
            stats.count_passes++;
            stats.delta_passes++;
            stats.delta_edges += last_nd;  // its update consumed the previous iteration's deletions
            last_pass_delta = true;
        }
        prev_nd = last_nd;
        last_nd = nd;
        iter = q.it + 1;
        const uint32_t fl = q.updated ? (KT_ITER_COUNTED | KT_ITER_DELTA | (timing ? (uint32_t)KT_ITER_COUNT_TIMED : 0u)) : 0u;
        return record_iteration(k, q.it, nd, 0, ms, fl);
    };
    // the phase ended at the iteration just read: the younger queued ones are
    // void; the youngest's count is left for the next read-back to check.
    // Without a younger one, the list a light youngest updated is finalized.
    auto drain_void = [&]() -> int {
        for (size_t i = 0; i + 1 < open.size(); ++i) {
            uint32_t nd = 0;
            const int rc = read_count(i, nd);
            if (rc) return rc;
            if (nd) return fail(KT_ERR_STATE, "void DELTA iteration deleted entries");
        }
        if (!open.empty()) {
            void_src = counts[r % 3];
        } else if (unflagged) {
            const uint32_t upd = (r + 2) % 3;
            KT_HIP_TRY(dev::finalize_deleted(lists[upd], counts[upd], std::max<uint32_t>(prev_nd, 1u),
                                             dst[cur].as<uint32_t>(), snbr_b[acur].as<uint32_t>(),
                                             sa.as<uint32_t>(), sb.as<uint32_t>(), stream));
        }
        open.clear();
        pending = 0;
        ended = true;
        return KT_OK;
    };
    uint32_t next = head ? iter + 1 : iter;  // the next iteration to queue
    for (;;) {
        const int s = (int)(next % kAheadSlots);
        int rc = queue(s);
        if (rc) return rc;
        open.push_back({s, next++, true});
        bound = (uint32_t)std::min<uint64_t>(live, E);  // a queued iteration deletes live entries only
        const bool leave = compaction_due();
        while (open.size() > (leave ? 0u : kAhead)) {
            uint32_t nd = 0;
            if ((rc = finish(nd))) return rc;
            if (nd == 0) return drain_void();
        }
        if (leave) {
            // everything read back: after a light youngest its list is flagged
            // and the one before finalized (a task list without an update);
            // the pending list becomes dl / ctr->ndel and the compaction
            // follows in the caller
            if (unflagged && (rc = prepare(r, std::max(prev_nd, last_nd), nullptr, nullptr))) return rc;
            pending = last_nd;
            std::swap(dl.p, bufs[r % 3]->p);
            return set_counter(offsetof(DevCounters, ndel), pending);
        }
    }
}

// One phase at k with DELTA available: the mark pass defers its deletions
// (KT_DYING + list), and the support pass that opens the next iteration
// either removes their triangles (DELTA) or recounts.  The dead entries are
// compacted away once they are a third of the stored ones.
int Plan::run_phase_delta(uint32_t& k, bool reuse, uint32_t schedule) {
    begin_phase();
    // marks into dl / ctr->ndel list the DELTA tasks of their deletions (over
    // the adjacency lists current at the mark)
    dev::MarkTasks mt0{(unsigned long long*)&ctr.as<DevCounters>()->ndel, nullptr, dws};
    const bool jump = (schedule & KT_SCHED_JUMP) != 0;
    uint32_t iter = 0;
    bool broke = false, changed = false;
    int rc;
    stats.phases++;
    const uint32_t nc = (uint32_t)cands.size();
    std::array<float, KT_PHASE_REC> rec;
    rec.fill(-1.f);
    CountCfg cfg = phase_fixed;
    uint32_t chosen = 0;
    // the model's choice is read from pinned memory at an early phase start
    // (below), else with the next counter read-back (the JUMP's or, when
    // deletions are pending, right before the support pass)
    bool unresolved = pick_now == KT_PICK_MODEL;
    // with nothing pending the deletion count is free to clear now; the JUMP
    // minimum is reset here too (one kernel instead of three)
    const bool start_clear = pending == 0;
    if ((rc = take_void(start_clear, jump))) return rc;
    // a compaction that fell due in the last phase happens before the phase's
    // features are taken (they describe the stored rows)
    if ((rc = compact_delta_state())) return rc;
    // With nothing pending and no support pass to open the phase, the JUMP
    // minimum is taken before the features and stored into pinned memory by
    // its last workgroup, so the host queues the phase's first mark while the
    // features and the model's pick run; the choice arrives the same way.  The
    // phase then starts without a counter read-back.
    const bool early = jump && start_clear && reuse && pick_now != KT_PICK_SWEEP && k > 3 &&
                       (pick_now == KT_PICK_MODEL || phase_fixed.kernel == KT_COUNT_DELTA);
    if (early) {
        __atomic_store_n(h_ra + kSlotMin, kCountPending, __ATOMIC_RELEASE);
        __atomic_store_n(h_ra + kSlotChoice, kCountPending, __ATOMIC_RELEASE);
        KT_HIP_TRY(dev::min_live_delta(E, sup[cur].as<uint32_t>(), ctr.as<DevCounters>(), stream,
                                       d_ra + kSlotMin));
    }
    if ((rc = phase_pick(k, reuse && !pending, cfg, chosen, true, true, early ? d_ra + kSlotChoice : nullptr)))
        return rc;
    auto resolve_choice = [&]() {
        chosen = std::min<uint32_t>(h_ctr->choice, nc - 1);
        cfg = cands[chosen];
        unresolved = false;
    };
    std::vector<double> cand_ms(nc, 0.0);
    double phase_ms = 0.0;
    bool peel = true;  // the loop below runs
    if (early) {
        uint32_t mv = 0;
        if ((rc = poll_pinned(kSlotMin, mv))) return rc;
        if (mv != dev::kNoLiveSupport && mv + 3 > k) k = mv + 3;
        mt0.sptr = sptr_b[acur].as<uint32_t>();
        KT_HIP_TRY(dev::mark_dying(dst[cur].as<uint32_t>(), eid[cur].as<uint32_t>(), sup[cur].as<uint32_t>(),
                                   tau.as<uint32_t>(), dl.as<uint32_t>(), E, k, snbr_b[acur].as<uint32_t>(),
                                   sa.as<uint32_t>(), sb.as<uint32_t>(), &ctr.as<DevCounters>()->ndel, nullptr,
                                   nullptr, 0, src[cur].as<uint32_t>(), live_rows_ptr(),
                                   &ctr.as<DevCounters>()->mark_bad, stream, &mt0));
        if (unresolved) {
            uint32_t ch = 0;
            if ((rc = poll_pinned(kSlotChoice, ch))) return rc;
            chosen = std::min<uint32_t>(ch, nc - 1);
            cfg = cands[chosen];
            unresolved = false;
        }
        if (cfg.kernel == KT_COUNT_DELTA && !compaction_due()) {
            // the first mark heads the run-ahead queue
            bool ended = false;
            if ((rc = run_ahead_delta(k, iter, phase_ms, ended, true))) return rc;
            changed = !ended;
        } else {  // the first mark's count, read back as the loop does
            if ((rc = sync_counters())) return rc;
            const uint32_t ndel = h_ctr->ndel;
            pending = ndel;
            if ((rc = record_iteration(k, iter, ndel, 0, 0.f))) return rc;
            changed = ndel > 0;
            ++iter;
        }
        peel = changed;
    }
    if (peel) do {
        bool counted = false;
        const uint64_t rec_before = h_ctr->recount_edges;
        float dummy = 0;
        // the pending count is ctr->ndel of the last mark: reset only afterwards
        if (pending || !(iter == 0 && reuse)) {
            if (unresolved) {
                if ((rc = sync_counters())) return rc;
                resolve_choice();
            }
            if (pick_now == KT_PICK_SWEEP) rc = sweep_pass(k, cand_ms, dummy);
            else rc = support_pass(cfg, k, dummy);
            if (rc) return rc;
            counted = true;
            stats.count_passes++;
        }
        if (!(iter == 0 && start_clear) &&
            ((rc = set_counter(offsetof(DevCounters, ndel), 0)) || (rc = set_counter(offsetof(DevCounters, ntask), 0))))
            return rc;
        if ((rc = compact_delta_state())) return rc;
        if (iter == 0 && jump) {
            KT_HIP_TRY(dev::min_live_delta(E, sup[cur].as<uint32_t>(), ctr.as<DevCounters>(), stream));
            if ((rc = sync_counters())) return rc;
            if (unresolved) resolve_choice();
            const uint32_t s = h_ctr->min_sup;
            if (s != 0xFFFFFFFFu && s + 3 > k) k = s + 3;
            // with DELTA settled for the phase, iteration 0's mark (it deletes
            // the minimum) opens the run-ahead queue: its count comes back
            // with the next iteration's
            if (s != 0xFFFFFFFFu && k > 3 && !counted && pick_now != KT_PICK_SWEEP && !unresolved &&
                cfg.kernel == KT_COUNT_DELTA && !compaction_due()) {
                mt0.sptr = sptr_b[acur].as<uint32_t>();
                KT_HIP_TRY(dev::mark_dying(dst[cur].as<uint32_t>(), eid[cur].as<uint32_t>(),
                                           sup[cur].as<uint32_t>(), tau.as<uint32_t>(), dl.as<uint32_t>(), E,
                                           k, snbr_b[acur].as<uint32_t>(), sa.as<uint32_t>(),
                                           sb.as<uint32_t>(), &ctr.as<DevCounters>()->ndel, nullptr, nullptr,
                                           0, src[cur].as<uint32_t>(), live_rows_ptr(),
                                   &ctr.as<DevCounters>()->mark_bad, stream, &mt0));
                bool ended = false;
                if ((rc = run_ahead_delta(k, iter, phase_ms, ended, true))) return rc;
                changed = !ended;
                continue;
            }
        }
        mt0.sptr = sptr_b[acur].as<uint32_t>();
        KT_HIP_TRY(dev::mark_dying(dst[cur].as<uint32_t>(), eid[cur].as<uint32_t>(),
                                   sup[cur].as<uint32_t>(), tau.as<uint32_t>(), dl.as<uint32_t>(), E, k,
                                   snbr_b[acur].as<uint32_t>(), sa.as<uint32_t>(), sb.as<uint32_t>(),
                                   &ctr.as<DevCounters>()->ndel, nullptr, nullptr, 0, src[cur].as<uint32_t>(),
                                   live_rows_ptr(), &ctr.as<DevCounters>()->mark_bad, stream, &mt0));
        if ((rc = sync_counters())) return rc;
        if (unresolved) resolve_choice();
        float count_ms = 0;
        if (counted && pick_now == KT_PICK_SWEEP) {
            float best = 0;
            for (uint32_t c = 0; c < nc; ++c) {
                float ms = 0;
                KT_HIP_TRY(hipEventElapsedTime(&ms, ev_sw[2 * c], ev_sw[2 * c + 1]));
                cand_ms[c] += ms;
                if (c == 0 || ms < best) best = ms;
            }
            count_ms = best;
        } else if (counted) {
            KT_HIP_TRY(hipEventElapsedTime(&count_ms, ev_c0, ev_c1));
        }
        if (counted && pick_now != KT_PICK_SWEEP && last_pass_delta) {
            stats.ms_delta += count_ms;
        } else if (counted) {
            stats.ms_count += count_ms;
            stats.count_launches++;
        }
        phase_ms += count_ms;
        const uint32_t ndel = h_ctr->ndel;
        pending = ndel;
        const uint32_t fl = counted ? (KT_ITER_COUNTED | KT_ITER_COUNT_TIMED |
                                       (pick_now != KT_PICK_SWEEP && last_pass_delta ? (uint32_t)KT_ITER_DELTA : 0u))
                                    : 0u;
        if ((rc = record_iteration(k, iter, ndel, h_ctr->recount_edges - rec_before, count_ms, fl))) return rc;
        changed = ndel > 0;
        if (changed && k == 3) {  // kernel_ddc.cu:183-186
            broke = true;
            break;
        }
        ++iter;
        // with DELTA settled for the phase, the next iterations are queued
        // ahead of their read-backs until the phase ends or a compaction is due
        if (changed && pick_now != KT_PICK_SWEEP && !unresolved && cfg.kernel == KT_COUNT_DELTA &&
            !compaction_due()) {
            bool ended = false;
            if ((rc = run_ahead_delta(k, iter, phase_ms, ended))) return rc;
            if (ended) changed = false;
        }
    } while (changed);
    if (pick_now == KT_PICK_SWEEP && nc) {
        chosen = 0;
        for (uint32_t c = 1; c < nc; ++c)
            if (cand_ms[c] < cand_ms[chosen]) chosen = c;
        for (uint32_t c = 0; c < nc; ++c) rec[20 + c] = (float)cand_ms[c];
        phase_ms = cand_ms[chosen];
    }
    rec[0] = (float)k;
    rec[1] = (float)iter + (broke ? 1.f : 0.f);
    rec[2] = pick_now == KT_PICK_FIXED ? -1.f : (float)chosen;
    rec[3] = (float)phase_ms;
    phase_log.push_back(rec);  // features: filled in by flush_phase_features
    feat_on.push_back(phase_has_feat ? 1 : 0);
    stats.k_last = k;
    return KT_OK;
}

// ---------------------------------------------------------------------------
// Autotuner

// Times trial t (its ms filled in): the layout is rebuilt when t's tile shape
// is not the current one; one support count of the pristine layout at k,
// best of `repeats`.
int Plan::tune_trial(kt_tune_trial& t, uint32_t k, uint32_t repeats) {
    int rc;
    if (tiled && (t.tile_size != tile_size || t.num_parallel_tiles != ytiles)) {
        if ((rc = set_tiles(t.tile_size, 0, t.num_parallel_tiles))) return rc;
        if ((rc = build_layout(graph))) return rc;
        t.tile_size = tile_size;
        t.num_parallel_tiles = ytiles;
    }
    CountCfg c;
    const kt_count_config in{t.count_kernel == KT_COUNT_DELTA ? (uint32_t)KT_COUNT_ROW : t.count_kernel,
                             t.block_size, t.cap_lds, t.chunk};
    if ((rc = resolve(in, c))) return rc;
    if ((rc = reset())) return rc;
    float best = -1.f;
    for (uint32_t r = 0; r < std::max<uint32_t>(repeats, 1); ++r) {
        KT_HIP_TRY(dev::zero_live_supports(dst[cur].as<uint32_t>(), sup[cur].as<uint32_t>(), E, stream));
        CountArgs a = count_args(k, false, E, c, ctr_scratch.as<DevCounters>());
        KT_HIP_TRY(hipEventRecord(ev_c0, stream));
        KT_HIP_TRY(dev::count_supports(a, directed, tiled, c.kernel, c.block, stream));
        KT_HIP_TRY(hipEventRecord(ev_c1, stream));
        KT_HIP_TRY(hipEventSynchronize(ev_c1));
        float ms = 0;
        KT_HIP_TRY(hipEventElapsedTime(&ms, ev_c0, ev_c1));
        if (best < 0 || ms < best) best = ms;
    }
    t.ms = best;
    tune_log.push_back(t);
    return KT_OK;
}

int Plan::autotune(uint32_t k, uint32_t repeats) {
    if (k < 2) return fail(KT_ERR_ARG, "k must be at least 2");
    tune_log.clear();
    int rc;
    kt_tune_trial best{};
    best.tile_size = tiled ? tile_size : 0;
    best.num_parallel_tiles = ytiles;
    best.count_kernel = fixed_cfg.kernel;
    best.block_size = fixed_cfg.block;
    best.cap_lds = v.cap_lds;
    best.chunk = fixed_cfg.chunk;
    if ((rc = tune_trial(best, k, repeats))) return rc;
    // the plan's per-phase candidates first (a DELTA candidate as the count
    // that opens its runs): the log then holds every candidate's cost on the
    // whole graph, a sample for the cost model
    if (!tiled) {
        for (const CountCfg& c : cands) {
            kt_tune_trial t = best;
            t.count_kernel = c.kernel;
            t.block_size = c.block;
            t.cap_lds = c.cap;
            t.chunk = c.chunk;
            if ((rc = tune_trial(t, k, repeats))) return rc;
            if (t.ms < best.ms && (c.kernel != KT_COUNT_DELTA || fixed_cfg.kernel == KT_COUNT_DELTA)) best = t;
        }
    }
    // one axis at a time: each value tried against the best so far
    auto axis = [&](uint32_t kt_tune_trial::*field, std::vector<uint32_t> values) -> int {
        for (uint32_t x : values) {
            if (best.*field == x) continue;
            kt_tune_trial t = best;
            t.*field = x;
            int r = tune_trial(t, k, repeats);
            if (r == KT_ERR_ARG || r == KT_ERR_CAPACITY) continue;  // not a valid shape for this plan
            if (r) return r;
            if (t.ms < best.ms) best = t;
        }
        return KT_OK;
    };
    if (tiled) {
        std::vector<uint32_t> ts;
        for (uint32_t x : {tile_size, (n + 7) / 8, 65536u, 16384u, 4096u})
            if (x > 0 && x <= std::max<uint32_t>(n, 1) && std::find(ts.begin(), ts.end(), x) == ts.end()) ts.push_back(x);
        if ((rc = axis(&kt_tune_trial::tile_size, ts))) return rc;
        if ((rc = axis(&kt_tune_trial::num_parallel_tiles, {1u, 2u, 4u, 8u, 16u}))) return rc;
        if ((rc = axis(&kt_tune_trial::block_size, {256u, 512u, 1024u}))) return rc;
    } else {
        if (directed && fixed_cfg.kernel != KT_COUNT_DELTA &&
            (rc = axis(&kt_tune_trial::count_kernel, {(uint32_t)KT_COUNT_ROW, (uint32_t)KT_COUNT_SEARCH,
                                                      (uint32_t)KT_COUNT_MERGE})))
            return rc;
        if ((rc = axis(&kt_tune_trial::block_size, {256u, 512u, 1024u}))) return rc;
        if ((rc = axis(&kt_tune_trial::cap_lds, {0u, 512u, 1024u, 2048u}))) return rc;
        if ((rc = axis(&kt_tune_trial::chunk, {64u, 128u, 256u}))) return rc;
    }
    // apply the best: its tile shape (layout rebuilt if the last trial left
    // another) and its count configuration
    if (tiled && (best.tile_size != tile_size || best.num_parallel_tiles != ytiles)) {
        if ((rc = set_tiles(best.tile_size, 0, best.num_parallel_tiles))) return rc;
        if ((rc = build_layout(graph))) return rc;
    }
    const kt_count_config in{best.count_kernel, best.block_size, best.cap_lds, best.chunk};
    if ((rc = resolve(in, fixed_cfg))) return rc;
    v.count_kernel = fixed_cfg.kernel;
    v.block_size = fixed_cfg.block;
    v.cap_lds = best.cap_lds;
    v.chunk = fixed_cfg.chunk;
    cap_lds = fixed_cfg.cap;
    if (tiled) {
        v.tile_size = tile_size;
        v.num_parallel_tiles = ytiles;
    }
    return reset();
}

int Plan::truss_numbers(uint32_t* out, uint64_t count) {
    if (!have_tau) return fail(KT_ERR_STATE, "no decomposition has run on this plan");
    if (count != graph->m_input)
        return fail(KT_ERR_CAPACITY, "need " + std::to_string(graph->m_input) + " entries");
    std::vector<uint32_t> t(E0);
    if (E0) KT_HIP_TRY(hipMemcpy(t.data(), tau.p, (size_t)E0 * 4, hipMemcpyDeviceToHost));
    const uint32_t* uid = graph->uid.data();
    const uint32_t* pos = pos_of_edge.data();
    parallel_for(count, 1 << 16, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) out[i] = t[pos[uid[i]]];
    });
    return KT_OK;
}

// ---------------------------------------------------------------------------
// graph reference counting (a plan keeps its graph alive)

const kt_graph* graph_retain(const kt_graph* g) {
    const_cast<kt_graph*>(g)->refs.fetch_add(1);
    return g;
}

void graph_release(const kt_graph* g) {
    if (g && const_cast<kt_graph*>(g)->refs.fetch_sub(1) == 1) delete g;
}

}  // namespace kt
