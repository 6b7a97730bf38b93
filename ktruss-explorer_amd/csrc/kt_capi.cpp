// kt_capi.cpp -- the extern "C" boundary declared in include/ktruss.h.
// Every entry point converts C++ exceptions into status codes; nothing
// throws across the ABI.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "kt_batch_host.h"
#include "kt_host.h"
#include "kt_internal.h"
#include "kt_mlp.h"
#include "kt_plan.h"
#include "ktruss.h"

#define KT_GUARD(...)                                                     \
    try {                                                                 \
        __VA_ARGS__                                                       \
    } catch (const std::bad_alloc&) {                                     \
        return kt::fail(KT_ERR_NOMEM, "host allocation failed");          \
    } catch (const std::exception& ex) {                                  \
        return kt::fail(KT_ERR_STATE, ex.what());                         \
    }

extern "C" {

int kt_abi_version(void) { return KT_ABI_VERSION; }

const char* kt_last_error(void) { return kt::last_error(); }

const char* kt_status_string(int status) {
    switch (status) {
        case KT_OK: return "ok";
        case KT_ERR_ARG: return "invalid argument";
        case KT_ERR_SELF_LOOP: return "self-loop";
        case KT_ERR_ASYMMETRIC: return "asymmetric input";
        case KT_ERR_DUPLICATE: return "duplicate entry";
        case KT_ERR_RANGE: return "vertex id out of range";
        case KT_ERR_NOMEM: return "out of memory";
        case KT_ERR_CAPACITY: return "buffer too small";
        case KT_ERR_IO: return "i/o error";
        case KT_ERR_HIP: return "HIP error";
        case KT_ERR_NO_DEVICE: return "no gfx950 device";
        case KT_ERR_STATE: return "invalid state";
        default: return "unknown status";
    }
}

int kt_init(int device) { KT_GUARD(return kt::init_device(device);) }

int kt_device_synchronize(void) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return kt::fail(KT_ERR_HIP, std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
    return KT_OK;
}

void kt_free(void* p) { std::free(p); }

void kt_variant_default(kt_variant* v) {
    if (!v) return;
    std::memset(v, 0, sizeof(*v));
    v->directedness = KT_DIRECTED_BY_DEGREE;  // config.h:192-198
    v->format = KT_COOCSR;
    v->recount = KT_RECOUNT_ALL;
    v->num_edge_remove_iter = 0;
    v->block_size = 512;
    v->tiles_per_dim = 8;
    v->num_parallel_tiles = 8;
    v->tile_size = 0;
    v->chunk = 0;
    v->schedule = KT_SCHED_JUMP | KT_SCHED_REUSE;
}

int kt_graph_create(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                    uint32_t flags, kt_graph** out) {
    KT_GUARD(return kt::graph_create(n, m, src, dst, flags, out);)
}

void kt_graph_destroy(kt_graph* g) { kt::graph_release(g); }

int kt_graph_info(const kt_graph* g, uint32_t* n, uint64_t* m_undirected, uint64_t* m_input,
                  uint32_t* max_degree) {
    if (!g) return kt::fail(KT_ERR_ARG, "graph is NULL");
    if (n) *n = g->n;
    if (m_undirected) *m_undirected = g->m;
    if (m_input) *m_input = g->m_input;
    if (max_degree) *max_degree = g->max_degree;
    return KT_OK;
}

int kt_read_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst) {
    KT_GUARD(return kt::read_tsv(path, n, m, src, dst);)
}

int kt_write_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst) {
    KT_GUARD(return kt::write_tsv(path, m, src, dst);)
}

int kt_read_truss_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst,
                      uint32_t** tau) {
    KT_GUARD(return kt::read_truss_tsv(path, n, m, src, dst, tau);)
}

int kt_write_truss_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst,
                       const uint32_t* tau) {
    KT_GUARD(return kt::write_truss_tsv(path, m, src, dst, tau);)
}

int kt_write_edges_bin(const char* path, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                       uint32_t flags) {
    KT_GUARD(return kt::write_edges_bin(path, n, m, src, dst, flags);)
}

int kt_read_edges_bin(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst,
                      uint32_t* flags) {
    KT_GUARD(return kt::read_edges_bin(path, n, m, src, dst, flags);)
}

int kt_is_edges_bin(const char* path) { KT_GUARD(return kt::is_edges_bin(path);) }

int kt_verify_edges(uint64_t m_a, const uint32_t* src_a, const uint32_t* dst_a, uint64_t m_b,
                    const uint32_t* src_b, const uint32_t* dst_b) {
    if ((m_a && (!src_a || !dst_a)) || (m_b && (!src_b || !dst_b)))
        return kt::fail(KT_ERR_ARG, "NULL edge arrays");
    KT_GUARD(return kt::verify_edges(m_a, src_a, dst_a, m_b, src_b, dst_b);)
}

int kt_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                     uint64_t seed, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst) {
    KT_GUARD(return kt::generate_rmat(scale, edge_factor, a, b, c, seed, n, m, src, dst);)
}

int kt_generate_chung_lu(uint32_t n, double gamma, double avg_degree, double max_weight,
                         uint32_t num_cliques, uint32_t clique_min, uint32_t clique_max,
                         uint64_t seed, uint64_t* m, uint32_t** src, uint32_t** dst) {
    KT_GUARD(return kt::generate_chung_lu(n, gamma, avg_degree, max_weight, num_cliques,
                                          clique_min, clique_max, seed, m, src, dst);)
}

int kt_generate_erdos_renyi(uint32_t n, double avg_degree, uint64_t seed, uint64_t* m,
                            uint32_t** src, uint32_t** dst) {
    KT_GUARD(return kt::generate_erdos_renyi(n, avg_degree, seed, m, src, dst);)
}

int kt_plan_create(const kt_graph* g, const kt_variant* v, kt_plan** out) {
    if (!g || !out) return kt::fail(KT_ERR_ARG, "graph or out is NULL");
    *out = nullptr;
    kt_variant dv;
    kt_variant_default(&dv);
    KT_GUARD({
        auto* p = new kt_plan();
        const int rc = p->impl.create(g, v ? v : &dv);
        if (rc) {
            delete p;
            return rc;
        }
        *out = p;
        return KT_OK;
    })
}

void kt_plan_destroy(kt_plan* p) { delete p; }

void kt_plan_destroy_timed(kt_plan* p, double* ms) {
    if (!p) return;
    p->impl.release_timed(ms);
    delete p;
}

int kt_plan_info(const kt_plan* p, uint64_t* entries, uint32_t* max_row) {
    if (!p) return kt::fail(KT_ERR_ARG, "plan is NULL");
    if (entries) *entries = p->impl.E0;
    if (max_row) *max_row = p->impl.max_row;
    return KT_OK;
}

int kt_plan_ktruss(kt_plan* p, uint32_t k, uint32_t* out_src, uint32_t* out_dst, uint64_t cap,
                   uint64_t* out_m, kt_stats* stats) {
    if (!p || !out_m) return kt::fail(KT_ERR_ARG, "plan or out_m is NULL");
    KT_GUARD({
        std::vector<uint32_t> s, d;
        const int rc = p->impl.ktruss(k, s, d);
        if (rc) return rc;
        if (stats) *stats = p->impl.stats;
        *out_m = s.size();
        if (s.size() > cap) return kt::fail(KT_ERR_CAPACITY, "output needs " + std::to_string(s.size()) + " entries");
        if (!s.empty() && (!out_src || !out_dst)) return kt::fail(KT_ERR_ARG, "output arrays NULL");
        if (!s.empty()) {
            std::memcpy(out_src, s.data(), s.size() * 4);
            std::memcpy(out_dst, d.data(), d.size() * 4);
        }
        return KT_OK;
    })
}

int kt_ktruss_noalloc(uint32_t n, uint32_t E, const uint32_t* ptr, const uint32_t* src, uint32_t* dst,
                      uint32_t k, void* stream, uint64_t* out_live) {
    KT_GUARD(return kt::ktruss_noalloc(n, E, ptr, src, dst, k, (hipStream_t)stream, out_live);)
}

int kt_plan_decompose(kt_plan* p, kt_stats* stats) {
    if (!p) return kt::fail(KT_ERR_ARG, "plan is NULL");
    KT_GUARD({
        const int rc = p->impl.decompose();
        if (rc) return rc;
        if (stats) *stats = p->impl.stats;
        return KT_OK;
    })
}

int kt_plan_truss_numbers(kt_plan* p, uint32_t* tau, uint64_t count) {
    if (!p || (count && !tau)) return kt::fail(KT_ERR_ARG, "NULL argument");
    KT_GUARD(return p->impl.truss_numbers(tau, count);)
}

int kt_plan_trace(const kt_plan* p, uint64_t* recs, uint64_t cap, uint64_t* len) {
    if (!p || !len) return kt::fail(KT_ERR_ARG, "NULL argument");
    const auto& tr = p->impl.trace;
    *len = tr.size();
    if (recs) {
        const uint64_t c = std::min<uint64_t>(cap, tr.size());
        for (uint64_t i = 0; i < c; ++i)
            for (int j = 0; j < 5; ++j) recs[5 * i + j] = tr[i][j];
    }
    return KT_OK;
}

int kt_plan_iteration_log(const kt_plan* p, uint64_t* recs, uint64_t cap, uint64_t* len) {
    if (!p || !len) return kt::fail(KT_ERR_ARG, "NULL argument");
    const auto& tr = p->impl.trace;
    *len = tr.size();
    if (recs) {
        const uint64_t c = std::min<uint64_t>(cap, tr.size());
        for (uint64_t i = 0; i < c; ++i)
            for (int j = 0; j < KT_ITER_REC; ++j) recs[(size_t)KT_ITER_REC * i + j] = tr[i][j];
    }
    return KT_OK;
}

int kt_plan_stage_times(const kt_plan* p, double* ms, uint32_t n) {
    if (!p || (n && !ms)) return kt::fail(KT_ERR_ARG, "NULL argument");
    for (uint32_t i = 0; i < n; ++i) ms[i] = i < KT_PLAN_STAGES ? p->impl.stage_ms[i] : 0.0;
    return KT_OK;
}

int kt_plan_autotune(kt_plan* p, uint32_t k, uint32_t repeats, kt_variant* best) {
    if (!p) return kt::fail(KT_ERR_ARG, "plan is NULL");
    KT_GUARD({
        const int rc = p->impl.autotune(k, repeats);
        if (rc) return rc;
        if (best) *best = p->impl.v;
        return KT_OK;
    })
}

int kt_plan_tune_log(const kt_plan* p, kt_tune_trial* recs, uint64_t cap, uint64_t* len) {
    if (!p || !len) return kt::fail(KT_ERR_ARG, "NULL argument");
    const auto& lg = p->impl.tune_log;
    *len = lg.size();
    if (recs)
        for (uint64_t i = 0; i < std::min<uint64_t>(cap, lg.size()); ++i) recs[i] = lg[i];
    return KT_OK;
}

int kt_plan_set_sweep_sampling(kt_plan* p, uint32_t every, uint32_t offset, uint32_t fill) {
    if (!p || every == 0 || offset >= every) return kt::fail(KT_ERR_ARG, "plan is NULL, every is 0 or offset >= every");
    p->impl.sweep_every = every;
    p->impl.sweep_offset = offset;
    p->impl.sweep_fill = fill;
    return KT_OK;
}

int kt_plan_set_candidates(kt_plan* p, const kt_count_config* cands, uint32_t n) {
    if (!p) return kt::fail(KT_ERR_ARG, "plan is NULL");
    KT_GUARD(return p->impl.set_candidates(cands, n);)
}

int kt_plan_set_timing(kt_plan* p, int on) {
    if (!p) return kt::fail(KT_ERR_ARG, "plan is NULL");
    p->impl.timing = on != 0;
    return KT_OK;
}

int kt_plan_set_pick(kt_plan* p, uint32_t mode, const kt_mlp* model) {
    if (!p) return kt::fail(KT_ERR_ARG, "plan is NULL");
    if (mode == KT_PICK_MODEL && !model) return kt::fail(KT_ERR_ARG, "model is NULL");
    KT_GUARD({
        if (mode == KT_PICK_MODEL)
            return p->impl.set_pick(mode, model->impl.P, model->impl.total, model->impl.C,
                                    model->impl.R);
        return p->impl.set_pick(mode, nullptr, 0, 0, nullptr);
    })
}

int kt_plan_phase_log(const kt_plan* p, float* recs, uint64_t cap, uint64_t* len) {
    if (!p || !len) return kt::fail(KT_ERR_ARG, "NULL argument");
    const auto& log = p->impl.phase_log;
    *len = log.size();
    if (recs) {
        const uint64_t c = std::min<uint64_t>(cap, log.size());
        for (uint64_t i = 0; i < c; ++i)
            std::memcpy(recs + i * KT_PHASE_REC, log[i].data(), KT_PHASE_REC * sizeof(float));
    }
    return KT_OK;
}

int kt_mlp_create(uint32_t n_out, uint64_t seed, kt_mlp** out) {
    if (!out) return kt::fail(KT_ERR_ARG, "out is NULL");
    *out = nullptr;
    KT_GUARD({
        int rc = kt::init_device(0);
        if (rc) return rc;
        auto* m = new kt_mlp();
        rc = m->impl.create(n_out, seed);
        if (rc) {
            delete m;
            return rc;
        }
        *out = m;
        return KT_OK;
    })
}

void kt_mlp_destroy(kt_mlp* m) { delete m; }

int kt_mlp_num_params(const kt_mlp* m, uint32_t* n) {
    if (!m || !n) return kt::fail(KT_ERR_ARG, "NULL argument");
    *n = m->impl.total;
    return KT_OK;
}

int kt_mlp_get_params(const kt_mlp* m, float* out, uint32_t n) {
    if (!m || !out) return kt::fail(KT_ERR_ARG, "NULL argument");
    KT_GUARD(return m->impl.get(out, n);)
}

int kt_mlp_set_params(kt_mlp* m, const float* in, uint32_t n) {
    if (!m || !in) return kt::fail(KT_ERR_ARG, "NULL argument");
    KT_GUARD(return m->impl.set(in, n);)
}

int kt_mlp_set_input_range(kt_mlp* m, const float* lo, const float* hi) {
    if (!m) return kt::fail(KT_ERR_ARG, "model is NULL");
    KT_GUARD(return m->impl.set_range(lo, hi);)
}

int kt_mlp_get_input_range(const kt_mlp* m, float* lo, float* hi) {
    if (!m) return kt::fail(KT_ERR_ARG, "model is NULL");
    KT_GUARD(return m->impl.get_range(lo, hi);)
}

int kt_mlp_forward(kt_mlp* m, const float* x, uint32_t batch, float* y) {
    if (!m || (batch && (!x || !y))) return kt::fail(KT_ERR_ARG, "NULL argument");
    KT_GUARD(return m->impl.forward(x, batch, y);)
}

int kt_mlp_loss_grad(kt_mlp* m, const float* x, const float* t, const float* mask, const float* w,
                     uint32_t batch, float* loss, float* grad) {
    if (!m || !x || !t) return kt::fail(KT_ERR_ARG, "NULL argument");
    KT_GUARD(return m->impl.loss_grad(x, t, mask, w, batch, loss, grad);)
}

int kt_mlp_train(kt_mlp* m, const float* x, const float* t, const float* mask, const float* w,
                 uint32_t n, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                 float* loss_out) {
    if (!m || !x || !t) return kt::fail(KT_ERR_ARG, "NULL argument");
    KT_GUARD(return m->impl.train(x, t, mask, w, n, epochs, batch, lr, seed, loss_out);)
}

// ---- batched decomposition --------------------------------------------------

int kt_batch_create(const kt_graph* const* graphs, uint32_t count, const kt_variant* v,
                    kt_batch** out) {
    if (!graphs || !out) return kt::fail(KT_ERR_ARG, "graphs or out is NULL");
    *out = nullptr;
    kt_variant dv;
    kt_variant_default(&dv);
    KT_GUARD({
        auto* b = new kt_batch();
        const int rc = b->impl.create(graphs, count, v ? v : &dv);
        if (rc) {
            delete b;
            return rc;
        }
        *out = b;
        return KT_OK;
    })
}

void kt_batch_destroy(kt_batch* b) { delete b; }

int kt_batch_info(const kt_batch* b, uint32_t* graphs, uint64_t* entries, uint32_t* max_row) {
    if (!b) return kt::fail(KT_ERR_ARG, "batch is NULL");
    if (graphs) *graphs = b->impl.G;
    if (entries) *entries = b->impl.E;
    if (max_row) *max_row = b->impl.max_row;
    return KT_OK;
}

int kt_batch_set_tuning(kt_batch* b, const kt_batch_tuning* t) {
    if (!b || !t) return kt::fail(KT_ERR_ARG, "NULL argument");
    KT_GUARD(return b->impl.set_tuning(*t);)
}

int kt_batch_get_tuning(const kt_batch* b, kt_batch_tuning* t) {
    if (!b || !t) return kt::fail(KT_ERR_ARG, "NULL argument");
    *t = b->impl.tune;
    return KT_OK;
}

int kt_batch_autotune(kt_batch* b, uint32_t trial_steps, uint32_t repeats, kt_batch_tuning* best) {
    if (!b) return kt::fail(KT_ERR_ARG, "batch is NULL");
    KT_GUARD({
        const int rc = b->impl.autotune(trial_steps, repeats);
        if (rc) return rc;
        if (best) *best = b->impl.tune;
        return KT_OK;
    })
}

int kt_batch_tune_log(const kt_batch* b, kt_batch_tuning* cfg, float* ms, uint64_t cap,
                      uint64_t* len) {
    if (!b || !len) return kt::fail(KT_ERR_ARG, "NULL argument");
    const auto& c = b->impl.tune_log_cfg;
    *len = c.size();
    const uint64_t nn = std::min<uint64_t>(cap, c.size());
    for (uint64_t i = 0; i < nn; ++i) {
        if (cfg) cfg[i] = c[i];
        if (ms) ms[i] = b->impl.tune_log_ms[i];
    }
    return KT_OK;
}

int kt_batch_set_timing(kt_batch* b, int on) {
    if (!b) return kt::fail(KT_ERR_ARG, "batch is NULL");
    b->impl.timed = on != 0;
    return KT_OK;
}

int kt_batch_decompose(kt_batch* b, kt_batch_stats* stats) {
    if (!b) return kt::fail(KT_ERR_ARG, "batch is NULL");
    KT_GUARD({
        const int rc = b->impl.decompose();
        if (rc) return rc;
        if (stats) *stats = b->impl.stats;
        return KT_OK;
    })
}

int kt_batch_truss_numbers(kt_batch* b, uint32_t graph, uint32_t* tau, uint64_t count) {
    if (!b || (count && !tau)) return kt::fail(KT_ERR_ARG, "NULL argument");
    KT_GUARD(return b->impl.truss_numbers(graph, tau, count);)
}

int kt_batch_graph_info(kt_batch* b, uint32_t graph, uint32_t* info) {
    if (!b || !info) return kt::fail(KT_ERR_ARG, "NULL argument");
    return b->impl.graph_info(graph, info);
}

}  // extern "C"
