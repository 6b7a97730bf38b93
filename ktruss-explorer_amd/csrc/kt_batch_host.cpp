// kt_batch_host.cpp -- host side of the batched decomposition: the union
// layout, the captured peel steps and their replay loop, and the autotuner.
//
// The host never decides anything per peel step.  It launches a pre-step
// (first full count and the first JUMP), then replays a HIP graph holding
// `steps_per_replay` peel steps and reads the 64-byte counter block once per
// replay; steps after the last graph finished find empty task lists and exit
// at once.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>

#include "kt_batch_host.h"
#include "kt_host.h"

namespace kt {

namespace {

constexpr uint32_t kStateArrays = 9;  // gk gflag glive gndel gmin giter gphase gkmax gpass

uint32_t next_pow2(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

bool same_tuning(const kt_batch_tuning& a, const kt_batch_tuning& b) {
    return a.block_size == b.block_size && a.chunk == b.chunk && a.cap_lds == b.cap_lds &&
           a.blocks_per_cu == b.blocks_per_cu && a.mark_chunk == b.mark_chunk &&
           a.steps_per_replay == b.steps_per_replay;
}

}  // namespace

Batch::~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (h_bc) (void)hipHostFree(h_bc);
    if (h_rb) (void)hipHostFree(h_rb);
    for (auto* ev : {&ev_start, &ev_stop, &ev_t0, &ev_t1, &ev_rb[0], &ev_rb[1]})
        if (*ev) (void)hipEventDestroy(*ev);
    for (hipEvent_t ev : tev) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
    for (const kt_graph* g : graphs) graph_release(g);
}

int Batch::create(const kt_graph* const* gs, uint32_t count, const kt_variant* var) {
    int rc = ensure_device();
    if (rc) return rc;
    if (!count) return fail(KT_ERR_ARG, "empty batch");
    v = *var;
    if (v.directedness != KT_DIRECTED_BY_DEGREE && v.directedness != KT_DIRECTED_BY_INDEX)
        return fail(KT_ERR_ARG, "batched decomposition supports the directed variants");
    if (v.format != KT_COOCSR || v.recount != KT_RECOUNT_ALL)
        return fail(KT_ERR_ARG, "batched decomposition supports COO+CSR with recount all");
    if (v.count_kernel > KT_COUNT_DELTA) return fail(KT_ERR_ARG, "unknown count_kernel");
    delta = v.count_kernel == KT_COUNT_DELTA;
    for (uint32_t i = 0; i < count; ++i)
        if (!gs[i]) return fail(KT_ERR_ARG, "graph " + std::to_string(i) + " is NULL");
    G = count;
    vbeg.assign(G + 1, 0);
    gbeg.assign(G + 1, 0);
    uint64_t nv = 0, ne = 0;
    for (uint32_t i = 0; i < G; ++i) {
        vbeg[i] = (uint32_t)std::min<uint64_t>(nv, 0xFFFFFFFFu);
        gbeg[i] = (uint32_t)std::min<uint64_t>(ne, 0xFFFFFFFFu);
        nv += gs[i]->n;
        ne += gs[i]->m;
        if (nv >= KT_DEAD || ne >= KT_DEAD)
            return fail(KT_ERR_CAPACITY, "a batch holds fewer than 2^31 vertices and entries");
    }
    vbeg[G] = (uint32_t)nv;
    gbeg[G] = (uint32_t)ne;
    N = (uint32_t)nv;
    E = (uint32_t)ne;

    if (delta && N >= KT_DYING) return fail(KT_ERR_CAPACITY, "DELTA needs fewer than 2^30 vertices");
    // per-graph layouts (relabel, orientation, sorted rows) written into the union,
    // and with DELTA the graphs' full adjacency lists (graph g owns the list
    // slots [2 gbeg[g], 2 gbeg[g + 1]))
    std::vector<uint32_t> hptr((size_t)N + 1), hsrc(E), hdst(E);
    std::vector<uint32_t> hs_ptr, hs_nbr, hs_eid, hs_a, hs_b;
    if (delta) {
        hs_ptr.resize((size_t)N + 1);
        hs_nbr.resize(2 * (size_t)E);
        hs_eid.resize(2 * (size_t)E);
        hs_a.resize(E);
        hs_b.resize(E);
    }
    pos_of_edge.assign(G, {});
    std::atomic<uint32_t> mrow{0};
    std::mutex mu;
    int first_rc = KT_OK;
    std::string first_msg;
    parallel_for(G, 1, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            HostLayout L;
            const int r = build_layout(gs[i], v.directedness, L);
            if (r || L.E != gbeg[i + 1] - gbeg[i]) {
                std::lock_guard<std::mutex> lk(mu);
                if (!first_rc) {
                    first_rc = r ? r : KT_ERR_STATE;
                    first_msg = "graph " + std::to_string(i) + ": " +
                                (r ? std::string(last_error()) : std::string("layout size mismatch"));
                }
                continue;
            }
            const uint32_t vb = vbeg[i], eb = gbeg[i];
            for (uint32_t x = 0; x < L.n; ++x) hptr[(size_t)vb + x] = eb + L.ptr[x];
            for (uint32_t x = 0; x < L.E; ++x) {
                hsrc[(size_t)eb + x] = vb + L.src[x];
                hdst[(size_t)eb + x] = vb + L.dst[x];
            }
            uint32_t cur = mrow.load();
            while (L.max_row > cur && !mrow.compare_exchange_weak(cur, L.max_row)) {
            }
            if (delta) {
                std::vector<uint32_t> sp, sn, se, a, bb;
                const int r2 = build_adjacency(L.n, L.E, L.ptr.data(), L.src.data(), L.dst.data(), sp,
                                               sn, se, a, bb);
                if (r2) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (!first_rc) {
                        first_rc = r2;
                        first_msg = "graph " + std::to_string(i) + ": " + last_error();
                    }
                    continue;
                }
                const uint64_t sbase = 2 * (uint64_t)eb;
                for (uint32_t x = 0; x < L.n; ++x) hs_ptr[(size_t)vb + x] = (uint32_t)(sbase + sp[x]);
                for (size_t q = 0; q < sn.size(); ++q) {
                    hs_nbr[sbase + q] = vb + sn[q];
                    hs_eid[sbase + q] = eb + se[q];
                }
                for (uint32_t x = 0; x < L.E; ++x) {
                    hs_a[(size_t)eb + x] = (uint32_t)(sbase + a[x]);
                    hs_b[(size_t)eb + x] = (uint32_t)(sbase + bb[x]);
                }
            }
            pos_of_edge[i] = std::move(L.pos_of_edge);
        }
    });
    if (first_rc) return fail(first_rc, first_msg);
    hptr[N] = E;
    if (delta) {
        if (2 * (uint64_t)E >= 0xFFFFFFFFull)
            return fail(KT_ERR_CAPACITY, "DELTA needs fewer than 2^31 entries");
        hs_ptr[N] = 2 * E;
    }
    max_row = mrow.load();

    hipDeviceProp_t prop;
    int dev = 0;
    KT_HIP_TRY(hipGetDevice(&dev));
    KT_HIP_TRY(hipGetDeviceProperties(&prop, dev));
    num_cu = prop.multiProcessorCount > 0 ? (uint32_t)prop.multiProcessorCount : 256;

    KT_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto* ev : {&ev_start, &ev_stop, &ev_t0, &ev_t1, &ev_rb[0], &ev_rb[1]})
        KT_HIP_TRY(hipEventCreate(ev));
    KT_HIP_TRY(hipHostMalloc((void**)&h_bc, sizeof(BatchCounters), hipHostMallocDefault));
    KT_HIP_TRY(hipHostMalloc((void**)&h_rb, 2 * sizeof(BatchCounters), hipHostMallocDefault));
    std::memset(h_bc, 0, sizeof(BatchCounters));

    const size_t Eb = (size_t)std::max<uint32_t>(E, 1) * 4;
    // task descriptors: at most E / 64 + G per list (chunks are multiples of 64)
    const size_t maxtasks = (size_t)E / 64 + G + 1;
    if ((rc = ptr.alloc(((size_t)N + 1) * 4)) || (rc = src.alloc(Eb)) || (rc = dst0.alloc(Eb)) ||
        (rc = dst.alloc(Eb)) || (rc = sup.alloc(Eb)) || (rc = tau.alloc(Eb)) ||
        (rc = d_gbeg.alloc(((size_t)G + 1) * 4)) ||
        (rc = gstate.alloc((size_t)kStateArrays * G * 4)) ||
        (rc = lists.alloc(4 * ((size_t)G + 1) * 4)) || (rc = ddesc.alloc(maxtasks * 16)) ||
        (rc = adesc.alloc(maxtasks * 16)) || (rc = bc.alloc(sizeof(BatchCounters))) ||
        (rc = stripes.alloc(kStripeBytes)))
        return rc;
    KT_HIP_TRY(hipMemcpyAsync(ptr.p, hptr.data(), ((size_t)N + 1) * 4, hipMemcpyHostToDevice, stream));
    if (E) {
        KT_HIP_TRY(hipMemcpyAsync(src.p, hsrc.data(), (size_t)E * 4, hipMemcpyHostToDevice, stream));
        KT_HIP_TRY(hipMemcpyAsync(dst0.p, hdst.data(), (size_t)E * 4, hipMemcpyHostToDevice, stream));
    }
    KT_HIP_TRY(hipMemcpyAsync(d_gbeg.p, gbeg.data(), ((size_t)G + 1) * 4, hipMemcpyHostToDevice, stream));
    KT_HIP_TRY(dev::fill_u32(gstate.as<uint32_t>(), 0, (size_t)kStateArrays * G, stream));
    if (delta) {
        if ((rc = sptr.alloc(((size_t)N + 1) * 4)) || (rc = snbr0.alloc(2 * Eb)) ||
            (rc = snbr.alloc(2 * Eb)) || (rc = seid.alloc(2 * Eb)) || (rc = sa.alloc(Eb)) ||
            (rc = sb.alloc(Eb)) || (rc = dl.alloc(Eb)) || (rc = dl2.alloc(Eb)) ||
            (rc = dscratch.alloc(dev::delta_scratch_bytes(E))))
            return rc;
        dev::delta_scratch_carve(dscratch.p, E, dws);
        KT_HIP_TRY(hipMemcpyAsync(sptr.p, hs_ptr.data(), ((size_t)N + 1) * 4, hipMemcpyHostToDevice, stream));
        if (E) {
            KT_HIP_TRY(hipMemcpyAsync(snbr0.p, hs_nbr.data(), 2 * (size_t)E * 4, hipMemcpyHostToDevice, stream));
            KT_HIP_TRY(hipMemcpyAsync(seid.p, hs_eid.data(), 2 * (size_t)E * 4, hipMemcpyHostToDevice, stream));
            KT_HIP_TRY(hipMemcpyAsync(sa.p, hs_a.data(), (size_t)E * 4, hipMemcpyHostToDevice, stream));
            KT_HIP_TRY(hipMemcpyAsync(sb.p, hs_b.data(), (size_t)E * 4, hipMemcpyHostToDevice, stream));
        }
    }
    KT_HIP_TRY(hipStreamSynchronize(stream));

    uint32_t* st = gstate.as<uint32_t>();
    bd.G = G;
    bd.gbeg = d_gbeg.as<uint32_t>();
    bd.gend = d_gbeg.as<uint32_t>() + 1;
    bd.gk = st;
    bd.gflag = st + (size_t)G;
    bd.glive = st + 2 * (size_t)G;
    bd.gndel = st + 3 * (size_t)G;
    bd.gmin = st + 4 * (size_t)G;
    bd.giter = st + 5 * (size_t)G;
    bd.gphase = st + 6 * (size_t)G;
    bd.gkmax = st + 7 * (size_t)G;
    bd.gpass = st + 8 * (size_t)G;
    uint32_t* ls = lists.as<uint32_t>();
    bd.dlist = ls;
    bd.dscan = ls + ((size_t)G + 1);
    bd.alist = ls + 2 * ((size_t)G + 1);
    bd.ascan = ls + 3 * ((size_t)G + 1);
    bd.ddesc = ddesc.as<uint4>();
    bd.adesc = adesc.as<uint4>();
    bd.bc = bc.as<BatchCounters>();
    bd.schedule = v.schedule;
    bd.delta = delta ? 1u : 0u;
    graphs.assign(gs, gs + G);
    for (const kt_graph* g : graphs) graph_retain(g);
    kt_batch_tuning t{};
    t.block_size = v.block_size;
    t.chunk = v.chunk;
    t.cap_lds = v.cap_lds;
    return set_tuning(t);
}

int Batch::resolve(const kt_batch_tuning& in, kt_batch_tuning& out, uint32_t& cap_out) const {
    out.block_size = in.block_size ? in.block_size : 512;
    out.chunk = in.chunk ? in.chunk : 128;
    out.mark_chunk = in.mark_chunk ? in.mark_chunk : 1024;
    out.blocks_per_cu = in.blocks_per_cu ? in.blocks_per_cu : 2;
    out.steps_per_replay = in.steps_per_replay ? in.steps_per_replay : 16;
    // DELTA steps alternate between two deletion lists: a captured replay
    // holds an even number of steps so every replay starts on the first
    out.steps_per_replay += out.steps_per_replay & 1u;
    if (out.block_size % 64 || out.block_size > 1024)
        return fail(KT_ERR_ARG, "block_size must be a multiple of 64 up to 1024");
    if (out.chunk % 64 || out.mark_chunk % 64) return fail(KT_ERR_ARG, "chunks must be multiples of 64");
    if (out.blocks_per_cu > 32) return fail(KT_ERR_ARG, "blocks_per_cu must be at most 32");
    if (out.steps_per_replay > 4096) return fail(KT_ERR_ARG, "steps_per_replay must be at most 4096");
    uint32_t c;
    if (in.cap_lds) {
        if (in.cap_lds < 64 || in.cap_lds > 4096 || (in.cap_lds & (in.cap_lds - 1)))
            return fail(KT_ERR_ARG, "cap_lds must be a power of two in [64, 4096]");
        c = in.cap_lds;
    } else {
        c = std::min<uint32_t>(4096, std::max<uint32_t>(64, next_pow2(max_row)));
    }
    while (c > 64 && dev::count_lds_bytes(c, out.block_size, true, false, KT_COUNT_ROW) > 65536) c >>= 1;
    out.cap_lds = c;
    cap_out = c;
    return KT_OK;
}

int Batch::apply(const kt_batch_tuning& t) {
    kt_batch_tuning r{};
    uint32_t c = 0;
    int rc = resolve(t, r, c);
    if (rc) return rc;
    if (exec && !same_tuning(r, tune)) {
        KT_HIP_TRY(hipGraphExecDestroy(exec));
        exec = nullptr;
    }
    tune = r;
    cap = c;
    bd.cchunk = r.chunk;
    bd.mchunk = r.mark_chunk;
    return KT_OK;
}

int Batch::set_tuning(const kt_batch_tuning& t) { return apply(t); }

uint32_t Batch::count_blocks() const { return num_cu * tune.blocks_per_cu; }

int Batch::reset() {
    parity = 0;  // the first step marks into the first deletion list
    if (E) KT_HIP_TRY(hipMemcpyAsync(dst.p, dst0.p, (size_t)E * 4, hipMemcpyDeviceToDevice, stream));
    if (delta && E)
        KT_HIP_TRY(hipMemcpyAsync(snbr.p, snbr0.p, 2 * (size_t)E * 4, hipMemcpyDeviceToDevice, stream));
    KT_HIP_TRY(dev::fill_u32(tau.as<uint32_t>(), 0, E, stream));
    KT_HIP_TRY(dev::fill_u32(bc.as<uint32_t>(), 0, sizeof(BatchCounters) / 4, stream));
    KT_HIP_TRY(dev::fill_u32(stripes.as<uint32_t>(), 0, kStripeBytes / 4, stream));
    if (delta) KT_HIP_TRY(dev::delta_stats_reset(dws, stream));
    std::memset(h_bc, 0, sizeof(BatchCounters));
    have_tau = false;
    stats = kt_batch_stats{};
    return KT_OK;
}

int Batch::event_pair(size_t i, hipEvent_t& a, hipEvent_t& b) {
    while (tev.size() < 2 * (i + 1)) {
        hipEvent_t ev = nullptr;
        KT_HIP_TRY(hipEventCreate(&ev));
        tev.push_back(ev);
    }
    a = tev[2 * i];
    b = tev[2 * i + 1];
    return KT_OK;
}

// First full count and the first JUMP.
int Batch::prestep(hipEvent_t c0, hipEvent_t c1) {
    const TaskList dtl{bd.ddesc, &bd.bc->dtasks}, atl{bd.adesc, &bd.bc->atasks};
    CountArgs a{};
    a.ptr = ptr.as<uint32_t>();
    a.src = src.as<uint32_t>();
    a.dst = dst.as<uint32_t>();
    a.sup = sup.as<uint32_t>();
    a.nitems = E;  // the layout's size (the kernel walks task lists): picks the count kernel
    a.chunk = tune.chunk;
    a.cap_lds = cap;
    a.ytiles = 1;
    a.ctr = &bd.bc->count;
    a.stripes = stripes.as<unsigned long long>();
    KT_HIP_TRY(dev::batch_control(bd, 0, stream));
    KT_HIP_TRY(dev::batch_expand(bd, stream_blocks(), stream));
    KT_HIP_TRY(dev::batch_zero(dtl, dst.as<uint32_t>(), sup.as<uint32_t>(), stream_blocks(), stream));
    if (c0) KT_HIP_TRY(hipEventRecord(c0, stream));
    KT_HIP_TRY(dev::count_supports_batch(a, dtl, tune.block_size, count_blocks(), stream));
    if (c1) KT_HIP_TRY(hipEventRecord(c1, stream));
    KT_HIP_TRY(dev::batch_min(atl, dst.as<uint32_t>(), sup.as<uint32_t>(), bd.gmin, stream_blocks(), stream));
    KT_HIP_TRY(dev::batch_control(bd, 1, stream));
    KT_HIP_TRY(dev::batch_expand(bd, stream_blocks(), stream));
    return KT_OK;
}

// mark -> control -> expand -> zero -> count
int Batch::launch_step(hipEvent_t c0, hipEvent_t c1) {
    const TaskList dtl{bd.ddesc, &bd.bc->dtasks}, atl{bd.adesc, &bd.bc->atasks};
    CountArgs a{};
    a.ptr = ptr.as<uint32_t>();
    a.src = src.as<uint32_t>();
    a.dst = dst.as<uint32_t>();
    a.sup = sup.as<uint32_t>();
    a.nitems = E;  // the layout's size (the kernel walks task lists): picks the count kernel
    a.chunk = tune.chunk;
    a.cap_lds = cap;
    a.ytiles = 1;
    a.ctr = &bd.bc->count;
    a.stripes = stripes.as<unsigned long long>();
    if (delta) {
        // mark (which finalizes the previous step's deletions) -> DELTA update
        // of this step's deletions -> control (clears the next mark's count) -> expand
        const uint32_t p = parity;
        parity ^= 1u;
        uint32_t* lists[2] = {dl.as<uint32_t>(), dl2.as<uint32_t>()};
        unsigned int* counts[2] = {&bd.bc->count.ndel, &bd.bc->count.ndel2};
        KT_HIP_TRY(dev::batch_mark(atl, dst.as<uint32_t>(), sup.as<uint32_t>(), tau.as<uint32_t>(), bd.gk,
                                   bd.gndel, bd.gmin, lists[p], counts[p], snbr.as<uint32_t>(),
                                   sa.as<uint32_t>(), sb.as<uint32_t>(), lists[p ^ 1u], counts[p ^ 1u],
                                   src.as<uint32_t>(), &bd.bc->count.mark_bad, stream_blocks(), stream));
        KT_HIP_TRY(dev::delta_prepare(lists[p], counts[p], E, src.as<uint32_t>(), dst.as<uint32_t>(),
                                      sptr.as<uint32_t>(), dws, &bd.bc->count, stream));
        if (c0) KT_HIP_TRY(hipEventRecord(c0, stream));
        KT_HIP_TRY(dev::delta_apply(lists[p], counts[p], E, snbr.as<uint32_t>(), seid.as<uint32_t>(),
                                    sup.as<uint32_t>(), dws, &bd.bc->count, stream));
        if (c1) KT_HIP_TRY(hipEventRecord(c1, stream));
        BatchDev b2 = bd;
        b2.zero_cnt = counts[p ^ 1u];
        KT_HIP_TRY(dev::batch_control(b2, 2, stream));
        KT_HIP_TRY(dev::batch_expand(bd, stream_blocks(), stream));
        return KT_OK;
    }
    KT_HIP_TRY(dev::batch_mark(atl, dst.as<uint32_t>(), sup.as<uint32_t>(), tau.as<uint32_t>(), bd.gk,
                               bd.gndel, bd.gmin, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nullptr, stream_blocks(), stream));
    KT_HIP_TRY(dev::batch_control(bd, 2, stream));
    KT_HIP_TRY(dev::batch_expand(bd, stream_blocks(), stream));
    KT_HIP_TRY(dev::batch_zero(dtl, dst.as<uint32_t>(), sup.as<uint32_t>(), stream_blocks(), stream));
    if (c0) KT_HIP_TRY(hipEventRecord(c0, stream));
    KT_HIP_TRY(dev::count_supports_batch(a, dtl, tune.block_size, count_blocks(), stream));
    if (c1) KT_HIP_TRY(hipEventRecord(c1, stream));
    return KT_OK;
}

int Batch::capture(uint32_t steps) {
    if (exec && exec_steps == steps) return KT_OK;
    if (exec) {
        KT_HIP_TRY(hipGraphExecDestroy(exec));
        exec = nullptr;
    }
    KT_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = KT_OK;
    parity = 0;  // every replay starts on the first deletion list (steps is even)
    for (uint32_t s = 0; s < steps && !rc; ++s) rc = launch_step();
    parity = 0;
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(KT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        exec = nullptr;
        return fail(KT_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    exec_steps = steps;
    return KT_OK;
}

int Batch::decompose() {
    int rc = reset();
    if (rc) return rc;
    if (!timed && (rc = capture(tune.steps_per_replay))) return rc;
    // a graph of E_g entries needs at most E_g steps that delete and one
    // phase-closing step per distinct truss number
    uint64_t max_e = 0;
    for (uint32_t i = 0; i < G; ++i) max_e = std::max<uint64_t>(max_e, gbeg[i + 1] - gbeg[i]);
    const uint64_t step_bound = 2 * max_e + max_row + 64;
    KT_HIP_TRY(hipEventRecord(ev_start, stream));
    size_t nev = 0;
    hipEvent_t c0 = nullptr, c1 = nullptr;
    if (timed && (rc = event_pair(nev++, c0, c1))) return rc;
    if ((rc = prestep(c0, c1))) return rc;
    uint64_t replays = 0;
    if (timed) {
        while (true) {
            for (uint32_t s = 0; s < tune.steps_per_replay; ++s) {
                if ((rc = event_pair(nev++, c0, c1))) return rc;
                if ((rc = launch_step(c0, c1))) return rc;
            }
            ++replays;
            KT_HIP_TRY(hipMemcpyAsync(h_bc, bc.p, sizeof(BatchCounters), hipMemcpyDeviceToHost, stream));
            KT_HIP_TRY(hipStreamSynchronize(stream));
            if (h_bc->nactive == 0) break;
            if (h_bc->steps > step_bound)
                return fail(KT_ERR_STATE, "batched peel exceeded " + std::to_string(step_bound) + " steps");
        }
    } else {
        // replays run one ahead of their counter read-backs, so the device
        // does not wait for the host between replays; the replay queued after
        // the last graph finished finds empty task lists and changes nothing
        int open = -1;  // read-back slot of the previous replay
        while (true) {
            KT_HIP_TRY(hipGraphLaunch(exec, stream));
            const int slot = (int)(replays & 1u);
            ++replays;
            KT_HIP_TRY(hipMemcpyAsync(h_rb + slot, bc.p, sizeof(BatchCounters), hipMemcpyDeviceToHost, stream));
            KT_HIP_TRY(hipEventRecord(ev_rb[slot], stream));
            if (open >= 0) {
                KT_HIP_TRY(hipEventSynchronize(ev_rb[open]));
                if (h_rb[open].nactive == 0) {
                    --replays;  // the one just queued is void
                    break;
                }
                if (h_rb[open].steps > step_bound)
                    return fail(KT_ERR_STATE, "batched peel exceeded " + std::to_string(step_bound) + " steps");
            }
            open = slot;
        }
    }
    KT_HIP_TRY(hipEventRecord(ev_stop, stream));
    KT_HIP_TRY(dev::stripes_fold(stripes.as<unsigned long long>(), &bd.bc->count, stream));
    if (delta) KT_HIP_TRY(dev::delta_stats_fold(dws, &bd.bc->count, stream));
    KT_HIP_TRY(hipMemcpyAsync(h_bc, bc.p, sizeof(BatchCounters), hipMemcpyDeviceToHost, stream));
    KT_HIP_TRY(hipEventSynchronize(ev_stop));
    float ms = 0;
    KT_HIP_TRY(hipEventElapsedTime(&ms, ev_start, ev_stop));

    tau_on_host = false;  // the truss numbers stay on the device until read
    std::vector<uint32_t> st((size_t)kStateArrays * G);
    KT_HIP_TRY(hipMemcpy(st.data(), gstate.p, st.size() * 4, hipMemcpyDeviceToHost));
    h_ginfo.assign((size_t)4 * G, 0);
    stats = kt_batch_stats{};
    for (uint32_t i = 0; i < G; ++i) {
        const uint32_t kmax = st[7 * (size_t)G + i], it = st[5 * (size_t)G + i];
        const uint32_t ph = st[6 * (size_t)G + i], pa = st[8 * (size_t)G + i];
        if (!(st[(size_t)G + i] & KT_BF_DONE) || st[2 * (size_t)G + i] != 0)
            return fail(KT_ERR_STATE, "graph " + std::to_string(i) + " did not finish");
        h_ginfo[4 * (size_t)i] = kmax;
        h_ginfo[4 * (size_t)i + 1] = it;
        h_ginfo[4 * (size_t)i + 2] = ph;
        h_ginfo[4 * (size_t)i + 3] = pa;
        stats.k_max = std::max(stats.k_max, kmax);
        stats.iterations += it;
        stats.phases += ph;
        stats.count_passes += pa;
    }
    if (h_bc->count.mark_bad)
        return fail(KT_ERR_STATE, "the DELTA mark found a deleted entry below the threshold");
    stats.graphs = G;
    stats.entries = E;
    stats.steps = h_bc->steps;
    stats.replays = replays;
    stats.recount_edges = h_bc->count.recount_edges;
    stats.merge_elems = h_bc->count.merge_elems;
    stats.delta_elems = h_bc->count.delta_elems;
    stats.delta_triangles = h_bc->count.delta_tris;
    stats.balg_delta_bytes = 4.0 * (double)h_bc->count.delta_elems +
                             4.0 * (double)h_bc->count.delta_win +
                             32.0 * (double)h_bc->count.delta_tris;
    stats.live_in_passes = h_bc->live_in_passes;
    stats.ms_total = ms;
    stats.balg_count_bytes = 4.0 * (double)stats.merge_elems + 24.0 * (double)stats.recount_edges +
                             12.0 * (double)stats.live_in_passes;
    stats.count_launches = 1 + replays * tune.steps_per_replay;
    // timing mode: pair 0 brackets the first full count, the others the
    // steps' count launches (recount mode) or DELTA updates (DELTA mode)
    for (size_t i = 0; timed && i < nev; ++i) {
        float x = 0;
        KT_HIP_TRY(hipEventElapsedTime(&x, tev[2 * i], tev[2 * i + 1]));
        if (i > 0 && delta) stats.ms_delta += x;
        else stats.ms_count += x;
    }
    if (delta) stats.count_launches = 1;
    have_tau = true;
    return KT_OK;
}

int Batch::trial(uint32_t steps, float& ms) {
    int rc = reset();
    if (rc) return rc;
    KT_HIP_TRY(hipEventRecord(ev_t0, stream));
    if ((rc = prestep())) return rc;
    for (uint32_t s = 0; s < steps; ++s)
        if ((rc = launch_step())) return rc;
    KT_HIP_TRY(hipEventRecord(ev_t1, stream));
    KT_HIP_TRY(hipEventSynchronize(ev_t1));
    KT_HIP_TRY(hipEventElapsedTime(&ms, ev_t0, ev_t1));
    return KT_OK;
}

int Batch::autotune(uint32_t trial_steps, uint32_t repeats) {
    if (repeats == 0) repeats = 1;
    tune_log_cfg.clear();
    tune_log_ms.clear();
    auto measure = [&](const kt_batch_tuning& t, float& ms) -> int {
        for (size_t i = 0; i < tune_log_cfg.size(); ++i)
            if (same_tuning(tune_log_cfg[i], t)) {
                ms = tune_log_ms[i];
                return KT_OK;
            }
        int rc = apply(t);
        if (rc) return rc;
        ms = 1e30f;
        for (uint32_t r = 0; r < repeats; ++r) {
            float x = 0;
            if ((rc = trial(trial_steps, x))) return rc;
            ms = std::min(ms, x);
        }
        tune_log_cfg.push_back(tune);
        tune_log_ms.push_back(ms);
        return KT_OK;
    };
    kt_batch_tuning best = tune;
    float best_ms = 0;
    int rc = measure(best, best_ms);
    if (rc) return rc;
    const uint32_t blocks[] = {256, 512, 1024};
    const uint32_t caps[] = {4096, 2048, 1024, 512, 256};
    const uint32_t chunks[] = {64, 128, 256, 512};
    const uint32_t bpcs[] = {1, 2, 4, 8};
    const uint32_t mchunks[] = {256, 1024, 4096};
    const uint32_t auto_cap = std::min<uint32_t>(4096, std::max<uint32_t>(64, next_pow2(max_row)));
    for (int pass = 0; pass < 3; ++pass) {
        bool improved = false;
        for (int p = 0; p < 5; ++p) {
            const uint32_t* vals = p == 0 ? blocks : p == 1 ? caps : p == 2 ? chunks : p == 3 ? bpcs : mchunks;
            const int nv = p == 0 ? 3 : p == 1 ? 5 : p == 2 ? 4 : p == 3 ? 4 : 3;
            for (int i = 0; i < nv; ++i) {
                kt_batch_tuning c = best;
                if (p == 0) c.block_size = vals[i];
                if (p == 1) {
                    if (vals[i] > auto_cap) continue;  // slices beyond the longest row buy nothing
                    c.cap_lds = vals[i];
                }
                if (p == 2) c.chunk = vals[i];
                if (p == 3) c.blocks_per_cu = vals[i];
                if (p == 4) c.mark_chunk = vals[i];
                kt_batch_tuning r{};
                uint32_t cc = 0;
                if (resolve(c, r, cc)) continue;
                if (same_tuning(r, best)) continue;
                float ms = 0;
                if ((rc = measure(r, ms))) return rc;
                if (ms < best_ms * 0.98f) {
                    best = r;
                    best_ms = ms;
                    improved = true;
                }
            }
        }
        if (!improved) break;
    }
    return apply(best);
}

int Batch::truss_numbers(uint32_t g, uint32_t* out, uint64_t count) {
    if (!have_tau) return fail(KT_ERR_STATE, "no decomposition has run on this batch");
    if (g >= G) return fail(KT_ERR_ARG, "graph index out of range");
    if (count != graphs[g]->m_input)
        return fail(KT_ERR_CAPACITY, "need " + std::to_string(graphs[g]->m_input) + " entries");
    if (!tau_on_host) {  // the first read after a run downloads them all once
        h_tau.resize(E);
        if (E) KT_HIP_TRY(hipMemcpy(h_tau.data(), tau.p, (size_t)E * 4, hipMemcpyDeviceToHost));
        tau_on_host = true;
    }
    const uint32_t* uid = graphs[g]->uid.data();
    const uint32_t* pos = pos_of_edge[g].data();
    const uint32_t* t = h_tau.data() + gbeg[g];
    for (uint64_t i = 0; i < count; ++i) out[i] = t[pos[uid[i]]];
    return KT_OK;
}

int Batch::graph_info(uint32_t g, uint32_t* info) {
    if (!have_tau) return fail(KT_ERR_STATE, "no decomposition has run on this batch");
    if (g >= G) return fail(KT_ERR_ARG, "graph index out of range");
    std::memcpy(info, &h_ginfo[4 * (size_t)g], 16);
    return KT_OK;
}

}  // namespace kt
