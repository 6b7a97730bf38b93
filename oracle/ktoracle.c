/*
 * ktoracle.c -- CPU oracle for the k-truss peel path.
 *
 * TEST INFRASTRUCTURE ONLY (see ktoracle.h): the checker, never the product.
 * Single-threaded plain C restating the reference's Host path ("-p cpu") plus
 * the k++ outer loop.  Each function names the reference lines it follows.
 * Differences from the reference that do not change any result:
 *   - sortByDegree (graph.cu:152-185) orders equal-degree vertices with an
 *     unstable quicksort; here ties go to the lower id.  Truss membership does
 *     not depend on vertex labels, only the orientation of equal-degree pairs
 *     (and so the work counters) does.
 *   - compaction also carries a per-entry edge id (to map deletions to truss
 *     numbers) and the support array (for the REUSE schedule).
 *   - the templates' CHECK_IF_DELETED=0 instances skip the tombstone test when
 *     a graph holds no tombstone; testing always gives the same result.
 */
#include "ktoracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NOT_AFFECTED 0        /* info.h:10 */
#define DIRECTLY_AFFECTED 1   /* info.h:11 */
#define INDIRECTLY_AFFECTED 2 /* info.h:12 */

/* ---------------------------------------------------------------------- */
/* Graph containers: COOGraph / COOCSRGraph / TiledCOOCSRGraph (graph.h:10-76)
 * folded into one struct; ptr is srcPtr[n+1] or tileSrcPtr[T*T*S+1]. */
typedef struct {
    uint32_t n;
    uint32_t E;      /* numEdges (stored entries)          */
    uint32_t* src;   /* srcIdx                             */
    uint32_t* dst;   /* dstIdx, KTO_DELETED = tombstone    */
    uint32_t* eid;   /* unique-edge id of each entry       */
    uint32_t* tri;   /* Info.numTriangles                  */
    int tiled;
    uint32_t T, S;   /* tilesPerDim, tileSize              */
    uint64_t nkeys;  /* n or T*T*S                         */
    uint32_t* ptr;   /* nkeys + 1                          */
} kg;

/* Info (info.h:14-19); node/edge states kept as bytes. */
typedef struct {
    int changed;
    uint8_t* node_aff;
    uint8_t* edge_aff;
} kinfo;

double kto_balg_bytes(const kto_stats* s) {
    return 4.0 * (double)s->merge_elems + 24.0 * (double)s->recount_edges +
           12.0 * (double)s->live_in_passes + 16.0 * (double)s->compact_elems;
}

void kto_free(void* p) { free(p); }

/* Row key of an entry: src (COO+CSR) or the tiled key of graph.cu:674-676. */
static inline uint64_t row_key(const kg* g, uint32_t s, uint32_t d) {
    if (!g->tiled) return s;
    uint64_t st = s / g->S, dt = d / g->S;
    return (st * g->T + dt) * (uint64_t)g->S + s % g->S;
}

/* ---------------------------------------------------------------------- */
/* Input validation and unique-edge numbering.  The reference assumes a
 * symmetric graph without self-loops or duplicates (graph.cu:54,
 * kernel_h.cu:24 halves numEdges); here each assumption is checked. */
typedef struct { uint32_t a, b; uint32_t dir; uint64_t idx; } pair_rec;

static int cmp_pair(const void* x, const void* y) {
    const pair_rec* p = (const pair_rec*)x;
    const pair_rec* q = (const pair_rec*)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    if (p->dir != q->dir) return p->dir < q->dir ? -1 : 1;
    return 0;
}

/* uid[i] for every entry; *nu unique edges; ua/ub their endpoints (a < b). */
static int number_edges(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                        uint32_t* uid, uint32_t* nu, uint32_t** ua, uint32_t** ub) {
    pair_rec* r = (pair_rec*)malloc((m ? m : 1) * sizeof(pair_rec));
    if (!r) return KTO_ERR_NOMEM;
    for (uint64_t i = 0; i < m; ++i) {
        uint32_t s = src[i], d = dst[i];
        if (s >= n || d >= n) { free(r); return KTO_ERR_RANGE; }
        if (s == d) { free(r); return KTO_ERR_SELF_LOOP; }
        r[i].a = s < d ? s : d;
        r[i].b = s < d ? d : s;
        r[i].dir = s < d ? 0u : 1u;
        r[i].idx = i;
    }
    qsort(r, m, sizeof(pair_rec), cmp_pair);
    if (m % 2) { free(r); return KTO_ERR_ASYMMETRIC; }
    uint32_t count = 0;
    *ua = (uint32_t*)malloc((m / 2 + 1) * sizeof(uint32_t));
    *ub = (uint32_t*)malloc((m / 2 + 1) * sizeof(uint32_t));
    if (!*ua || !*ub) { free(r); return KTO_ERR_NOMEM; }
    uint64_t i = 0;
    while (i < m) {
        uint64_t j = i;
        while (j < m && r[j].a == r[i].a && r[j].b == r[i].b) ++j;
        uint64_t c = j - i;
        if (c == 2 && r[i].dir == 0 && r[i + 1].dir == 1) {
            (*ua)[count] = r[i].a;
            (*ub)[count] = r[i].b;
            uid[r[i].idx] = count;
            uid[r[i + 1].idx] = count;
            ++count;
        } else {
            int dup = (c > 2) || (c == 2 && r[i].dir == r[i + 1].dir);
            free(r); free(*ua); free(*ub); *ua = *ub = NULL;
            return dup ? KTO_ERR_DUPLICATE : KTO_ERR_ASYMMETRIC;
        }
        i = j;
    }
    free(r);
    *nu = count;
    return KTO_OK;
}

/* ---------------------------------------------------------------------- */
/* sortByDegree (graph.cu:152-185): ascending degree, ties by id.  Degree is
 * the number of entries with the vertex as source (graph.cu:159-162). */
typedef struct { uint32_t deg, id; } degrec;
static int cmp_deg(const void* x, const void* y) {
    const degrec* p = (const degrec*)x;
    const degrec* q = (const degrec*)y;
    if (p->deg != q->deg) return p->deg < q->deg ? -1 : 1;
    return p->id < q->id ? -1 : (p->id > q->id);
}

static int relabel_by_degree(uint32_t n, uint32_t nu, const uint32_t* ua, const uint32_t* ub,
                             uint32_t* old2new) {
    degrec* d = (degrec*)calloc(n ? n : 1, sizeof(degrec));
    if (!d) return KTO_ERR_NOMEM;
    for (uint32_t v = 0; v < n; ++v) d[v].id = v;
    for (uint32_t e = 0; e < nu; ++e) { d[ua[e]].deg++; d[ub[e]].deg++; }
    qsort(d, n, sizeof(degrec), cmp_deg);
    for (uint32_t i = 0; i < n; ++i) old2new[d[i].id] = i;
    free(d);
    return KTO_OK;
}

/* ---------------------------------------------------------------------- */
/* Build the stored graph from entry lists, sorted by (row key, dst):
 * coo2coocsr (graph.cu:372-412) / coo2tiledcoocsr (graph.cu:656-707).  The
 * histogram + scan + binning + per-row quicksort of the reference leaves each
 * row sorted by dst; sorting entries by (key, dst) gives the same arrays. */
typedef struct { uint64_t key; uint32_t s, d, id; } entry_rec;
static int cmp_entry(const void* x, const void* y) {
    const entry_rec* p = (const entry_rec*)x;
    const entry_rec* q = (const entry_rec*)y;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return p->d < q->d ? -1 : (p->d > q->d);
}

static void kg_free(kg* g) {
    free(g->src); free(g->dst); free(g->eid); free(g->tri); free(g->ptr);
    memset(g, 0, sizeof(*g));
}

static void rebuild_ptr(kg* g) {
    /* histogram + exclusive scan, as in removeCOOCSRDeletedEdges
     * (graph.cu:494-507) and removeTiledCOOCSRDeletedEdges (graph.cu:796-816) */
    memset(g->ptr, 0, (g->nkeys + 1) * sizeof(uint32_t));
    for (uint32_t e = 0; e < g->E; ++e) g->ptr[row_key(g, g->src[e], g->dst[e])]++;
    uint32_t sum = 0;
    for (uint64_t k = 0; k < g->nkeys; ++k) {
        uint32_t c = g->ptr[k];
        g->ptr[k] = sum;
        sum += c;
    }
    g->ptr[g->nkeys] = sum;
}

static int kg_build(kg* g, uint32_t n, uint32_t E, const uint32_t* s, const uint32_t* d,
                    const uint32_t* id, int tiled, uint32_t tiles_per_dim) {
    memset(g, 0, sizeof(*g));
    g->n = n;
    g->E = E;
    g->tiled = tiled;
    if (tiled) {
        if (tiles_per_dim == 0) return KTO_ERR_ARG;
        g->T = tiles_per_dim;
        g->S = (n + tiles_per_dim - 1) / tiles_per_dim; /* graph.cu:580 */
        if (g->S == 0) g->S = 1;
        g->nkeys = (uint64_t)g->T * g->T * g->S;          /* graph.cu:581 */
        if (g->nkeys > (1ull << 31)) return KTO_ERR_ARG;
    } else {
        g->nkeys = n;
    }
    size_t ecap = E ? E : 1;
    g->src = (uint32_t*)malloc(ecap * sizeof(uint32_t));
    g->dst = (uint32_t*)malloc(ecap * sizeof(uint32_t));
    g->eid = (uint32_t*)malloc(ecap * sizeof(uint32_t));
    g->tri = (uint32_t*)calloc(ecap, sizeof(uint32_t));
    g->ptr = (uint32_t*)malloc((g->nkeys + 1) * sizeof(uint32_t));
    entry_rec* r = (entry_rec*)malloc(ecap * sizeof(entry_rec));
    if (!g->src || !g->dst || !g->eid || !g->tri || !g->ptr || !r) {
        free(r); kg_free(g); return KTO_ERR_NOMEM;
    }
    for (uint32_t e = 0; e < E; ++e) {
        r[e].key = row_key(g, s[e], d[e]);
        r[e].s = s[e]; r[e].d = d[e]; r[e].id = id[e];
    }
    qsort(r, E, sizeof(entry_rec), cmp_entry);
    for (uint32_t e = 0; e < E; ++e) { g->src[e] = r[e].s; g->dst[e] = r[e].d; g->eid[e] = r[e].id; }
    free(r);
    rebuild_ptr(g);
    return KTO_OK;
}

/* removeCOOCSRDeletedEdges (graph.cu:480-509) and
 * removeTiledCOOCSRDeletedEdges (graph.cu:782-818): stable removal of
 * tombstoned entries, then pointer rebuild. */
static void kg_compact(kg* g, kto_stats* st) {
    if (st) { st->compactions++; st->compact_elems += g->E; }
    uint32_t w = 0;
    for (uint32_t e = 0; e < g->E; ++e) {
        if (g->dst[e] != KTO_DELETED) {
            g->src[w] = g->src[e];
            g->dst[w] = g->dst[e];
            g->eid[w] = g->eid[e];
            g->tri[w] = g->tri[e];
            ++w;
        }
    }
    g->E = w;
    rebuild_ptr(g);
}

/* ---------------------------------------------------------------------- */
/* Support counting.  Every count_* tallies R_p and L_u+L_v for B_alg. */

static inline uint32_t row_begin(const kg* g, uint32_t u) { return g->ptr[u]; }
static inline uint32_t row_end(const kg* g, uint32_t u) { return g->ptr[u + 1]; }

/* Directed COO+CSR: kernel_hdc.cu:11-49.  Each triangle u<v<w is found once,
 * on edge (u,v), and credited to (u,v), (u,w), (v,w). */
static void count_dc(kg* g, const kinfo* in, int all, kto_stats* st) {
    for (uint32_t e = 0; e < g->E; ++e) {
        if (!all && in->edge_aff[e] != DIRECTLY_AFFECTED && in->edge_aff[e] != INDIRECTLY_AFFECTED)
            continue;
        uint32_t v = g->dst[e];
        if (v == KTO_DELETED) continue;
        uint32_t u = g->src[e];
        uint32_t e1 = row_begin(g, u), end1 = row_end(g, u);
        uint32_t e2 = row_begin(g, v), end2 = row_end(g, v);
        st->recount_edges++;
        st->merge_elems += (uint64_t)(end1 - e1) + (end2 - e2);
        while (e1 < end1 && e2 < end2) {
            uint32_t a = g->dst[e1];
            if (a == KTO_DELETED) { ++e1; continue; }
            uint32_t b = g->dst[e2];
            if (b == KTO_DELETED) { ++e2; continue; }
            if (a < b) ++e1;
            else if (a > b) ++e2;
            else { g->tri[e]++; g->tri[e1]++; g->tri[e2]++; ++e1; ++e2; }
        }
    }
}

/* Directed tiled: kernel_hdt.cu:11-62.  Edge (u,v) in tile block
 * (tile(u), tile(v)) merges the row segments of u and v inside each x-tile
 * from tile(v) on (w > v lies in a tile >= tile(v)). */
static void count_dt(kg* g, const kinfo* in, int all, kto_stats* st) {
    const uint64_t T = g->T, S = g->S;
    for (uint32_t e = 0; e < g->E; ++e) {
        if (!all && in->edge_aff[e] != DIRECTLY_AFFECTED && in->edge_aff[e] != INDIRECTLY_AFFECTED)
            continue;
        uint32_t v = g->dst[e];
        if (v == KTO_DELETED) continue;
        uint32_t u = g->src[e];
        uint64_t ut = u / S, vt = v / S;
        st->recount_edges++;
        for (uint64_t x = vt; x < T; ++x) {
            uint64_t k1 = (ut * T + x) * S + u % S;
            uint64_t k2 = (vt * T + x) * S + v % S;
            uint32_t e1 = g->ptr[k1], end1 = g->ptr[k1 + 1];
            uint32_t e2 = g->ptr[k2], end2 = g->ptr[k2 + 1];
            st->merge_elems += (uint64_t)(end1 - e1) + (end2 - e2);
            while (e1 < end1 && e2 < end2) {
                uint32_t a = g->dst[e1];
                if (a == KTO_DELETED) { ++e1; continue; }
                uint32_t b = g->dst[e2];
                if (b == KTO_DELETED) { ++e2; continue; }
                if (a < b) ++e1;
                else if (a > b) ++e2;
                else { g->tri[e]++; g->tri[e1]++; g->tri[e2]++; ++e1; ++e2; }
            }
        }
    }
}

/* Undirected COO+CSR: kernel_huc.cu:11-49.  |N(u) n N(v)| capped at k-2 with
 * early exit, after a stored-length pre-filter (huc:20). */
static void count_uc(kg* g, uint32_t k, const kinfo* in, int all, kto_stats* st) {
    const uint32_t cap = k - 2;
    for (uint32_t e = 0; e < g->E; ++e) {
        if (!all && in->edge_aff[e] != DIRECTLY_AFFECTED) continue;
        uint32_t v = g->dst[e];
        if (v == KTO_DELETED) continue;
        uint32_t u = g->src[e];
        uint32_t e1 = row_begin(g, u), end1 = row_end(g, u);
        uint32_t e2 = row_begin(g, v), end2 = row_end(g, v);
        st->recount_edges++;
        st->merge_elems += (uint64_t)(end1 - e1) + (end2 - e2);
        if (end1 - e1 >= cap && end2 - e2 >= cap) {
            while (e1 < end1 && e2 < end2 && g->tri[e] < cap) {
                uint32_t a = g->dst[e1];
                if (a == KTO_DELETED) { ++e1; continue; }
                uint32_t b = g->dst[e2];
                if (b == KTO_DELETED) { ++e2; continue; }
                if (a < b) ++e1;
                else if (a > b) ++e2;
                else { ++e1; ++e2; g->tri[e]++; }
            }
        }
    }
}

/* Undirected tiled: kernel_hut.cu:11-60.  Same capped count, accumulated over
 * all x-tiles; no stored-length pre-filter. */
static void count_ut(kg* g, uint32_t k, const kinfo* in, int all, kto_stats* st) {
    const uint32_t cap = k - 2;
    const uint64_t T = g->T, S = g->S;
    for (uint32_t e = 0; e < g->E; ++e) {
        if (!all && in->edge_aff[e] != DIRECTLY_AFFECTED) continue;
        uint32_t v = g->dst[e];
        if (v == KTO_DELETED) continue;
        uint32_t u = g->src[e];
        uint64_t ut = u / S, vt = v / S;
        st->recount_edges++;
        for (uint64_t x = 0; x < T; ++x) {
            uint64_t k1 = (ut * T + x) * S + u % S;
            uint64_t k2 = (vt * T + x) * S + v % S;
            uint32_t e1 = g->ptr[k1], end1 = g->ptr[k1 + 1];
            uint32_t e2 = g->ptr[k2], end2 = g->ptr[k2 + 1];
            st->merge_elems += (uint64_t)(end1 - e1) + (end2 - e2);
            while (e1 < end1 && e2 < end2 && g->tri[e] < cap) {
                uint32_t a = g->dst[e1];
                if (a == KTO_DELETED) { ++e1; continue; }
                uint32_t b = g->dst[e2];
                if (b == KTO_DELETED) { ++e2; continue; }
                if (a < b) ++e1;
                else if (a > b) ++e2;
                else { ++e1; ++e2; g->tri[e]++; }
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* mark_deleted_edges: identical in hdc:51-70, hdt:64-83, huc:51-70,
 * hut:62-81.  Records truss number k-1 for every deleted entry's edge. */
static uint64_t mark_deleted(kg* g, uint32_t k, kinfo* in, int all, uint32_t* tau_u) {
    uint64_t deleted = 0;
    for (uint32_t e = 0; e < g->E; ++e) {
        if (!all && in->edge_aff[e] != DIRECTLY_AFFECTED) continue;
        uint32_t v = g->dst[e];
        if (v == KTO_DELETED) continue;
        if (g->tri[e] < k - 2) {
            g->dst[e] = KTO_DELETED;
            in->changed = 1;
            ++deleted;
            if (tau_u) tau_u[g->eid[e]] = k - 1;
            if (!all) {
                uint32_t u = g->src[e];
                in->node_aff[u] = DIRECTLY_AFFECTED;
                in->node_aff[v] = DIRECTLY_AFFECTED;
            }
        }
    }
    return deleted;
}

/* mark_directly_affected_edges: directed hdc:72-90 / hdt:85-103 also flag the
 * other endpoint INDIRECTLY; undirected huc:72-85 / hut:83-96 do not. */
static void mark_directly(kg* g, kinfo* in, int directed) {
    for (uint32_t e = 0; e < g->E; ++e) {
        uint8_t st = NOT_AFFECTED;
        uint32_t v = g->dst[e];
        if (v != KTO_DELETED) {
            uint32_t u = g->src[e];
            if (in->node_aff[u] == DIRECTLY_AFFECTED || in->node_aff[v] == DIRECTLY_AFFECTED) {
                st = DIRECTLY_AFFECTED;
                if (directed) {
                    if (in->node_aff[u] != DIRECTLY_AFFECTED) in->node_aff[u] = INDIRECTLY_AFFECTED;
                    else if (in->node_aff[v] != DIRECTLY_AFFECTED) in->node_aff[v] = INDIRECTLY_AFFECTED;
                }
            }
        }
        in->edge_aff[e] = st;
    }
}

/* mark_indirectly_affected_edges: hdc:92-105 / hdt:105-118. */
static void mark_indirectly(kg* g, kinfo* in) {
    for (uint32_t e = 0; e < g->E; ++e) {
        uint32_t v = g->dst[e];
        if (v == KTO_DELETED) continue;
        if (in->edge_aff[e] != DIRECTLY_AFFECTED) {
            uint32_t u = g->src[e];
            if (in->node_aff[u] == INDIRECTLY_AFFECTED || in->node_aff[v] == INDIRECTLY_AFFECTED)
                in->edge_aff[e] = INDIRECTLY_AFFECTED;
        }
    }
}

static uint64_t live_entries(const kg* g) {
    uint64_t c = 0;
    for (uint32_t e = 0; e < g->E; ++e) c += g->dst[e] != KTO_DELETED;
    return c;
}

static int trace_push(kto_trace* tr, uint32_t k, uint32_t iter, uint64_t del, uint64_t rec) {
    if (!tr) return KTO_OK;
    if (tr->len >= tr->cap) { tr->len++; return KTO_ERR_CAPACITY; }
    tr->recs[tr->len].k = k;
    tr->recs[tr->len].iter = iter;
    tr->recs[tr->len].deleted = del;
    tr->recs[tr->len].recounted = rec;
    tr->len++;
    return KTO_OK;
}

static void do_count(kg* g, uint32_t k, const kinfo* in, int directed, int all, kto_stats* st) {
    st->count_passes++;
    st->live_in_passes += live_entries(g);
    if (directed) {
        if (g->tiled) count_dt(g, in, all, st); else count_dc(g, in, all, st);
    } else {
        if (g->tiled) count_ut(g, k, in, all, st); else count_uc(g, k, in, all, st);
    }
}

/* One phase at k: the peel loop ktruss(COOCSRGraph* or TiledCOOCSRGraph*, Info,
 * Config) of hdc:107-201 / hdt:120-214 / huc:87-179 / hut:98-190.
 *   *k_io       in: k; out: k after a JUMP
 *   reuse       the support array already holds exact supports of g
 *   *exact_out  the phase ended on a count with no deletion (supports exact)
 */
static int peel_phase(kg* g, uint32_t* k_io, const kto_variant* var, kinfo* in, int reuse,
                      uint32_t schedule, uint32_t* tau_u, kto_stats* st, kto_trace* tr,
                      int* exact_out) {
    const int directed = var->directedness != KTO_UNDIRECTED;
    const int all = var->recount == KTO_RECOUNT_ALL;
    uint32_t k = *k_io;
    uint32_t iter = 0;
    int has_del = 0;
    int changed;
    int rc = KTO_OK;
    int ended_by_break = 0;
    /* initInfo (info.cu:37-45): every edge directly affected at the start */
    if (!all) memset(in->edge_aff, DIRECTLY_AFFECTED, g->E ? g->E : 1);
    st->phases++;
    do {
        uint64_t recounted = 0;
        /* clearIterInfo (info.cu:62-70) */
        in->changed = 0;
        if (!all) memset(in->node_aff, NOT_AFFECTED, g->n ? g->n : 1);
        if (!(iter == 0 && reuse)) {
            memset(g->tri, 0, (g->E ? g->E : 1) * sizeof(uint32_t));
            uint64_t before = st->recount_edges;
            do_count(g, k, in, directed, all, st);
            recounted = st->recount_edges - before;
            if (directed && st->count_passes == 1) {
                uint64_t t = 0;
                for (uint32_t e = 0; e < g->E; ++e) t += g->tri[e];
                st->triangles = t / 3;
            }
        }
        if (iter == 0 && directed && (schedule & KTO_SCHED_JUMP)) {
            /* supports are exact here: phases k..s+2 would delete nothing */
            uint32_t s = 0xFFFFFFFFu;
            for (uint32_t e = 0; e < g->E; ++e)
                if (g->dst[e] != KTO_DELETED && g->tri[e] < s) s = g->tri[e];
            if (s != 0xFFFFFFFFu && s + 3 > k) k = s + 3;
        }
        uint64_t del = mark_deleted(g, k, in, all, tau_u);
        if (!directed) del /= 2;
        st->iterations++;
        int trc = trace_push(tr, k, iter, del, recounted);
        if (trc != KTO_OK) rc = trc;
        changed = in->changed;
        if (changed) {
            if (iter < var->num_edge_remove_iter) {
                kg_compact(g, st);
                has_del = 0;
            } else {
                has_del = 1;
            }
            if (k == 3) { ended_by_break = 1; break; } /* hdc:170-172 */
            if (!all) {
                mark_directly(g, in, directed);
                if (directed) mark_indirectly(g, in);
            }
        }
        ++iter;
    } while (changed);
    if (has_del) kg_compact(g, st); /* hdc:195-200 */
    *exact_out = !ended_by_break && !changed;
    *k_io = k;
    return rc;
}

/* Common set-up: validate, number edges, relabel, orient, build. */
typedef struct {
    uint32_t nu;
    uint32_t *ua, *ub;   /* unique edges, original labels, a < b */
    uint32_t* uid;       /* per input entry */
    uint32_t* old2new;
    uint32_t* new2old;
    kg g;
    kinfo in;
} setup_t;

static void setup_free(setup_t* s) {
    free(s->ua); free(s->ub); free(s->uid); free(s->old2new); free(s->new2old);
    kg_free(&s->g);
    free(s->in.node_aff); free(s->in.edge_aff);
    memset(s, 0, sizeof(*s));
}

static int setup(setup_t* s, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                 const kto_variant* v) {
    memset(s, 0, sizeof(*s));
    if ((m && (!src || !dst)) || !v) return KTO_ERR_ARG;
    if (v->directedness < 0 || v->directedness > 2 || v->format < 0 || v->format > 1 ||
        v->recount < 0 || v->recount > 1)
        return KTO_ERR_ARG;
    if (m > 0xFFFFFFFEull) return KTO_ERR_ARG;
    s->uid = (uint32_t*)malloc((m ? m : 1) * sizeof(uint32_t));
    s->old2new = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    s->new2old = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
    if (!s->uid || !s->old2new || !s->new2old) { setup_free(s); return KTO_ERR_NOMEM; }
    int rc = number_edges(n, m, src, dst, s->uid, &s->nu, &s->ua, &s->ub);
    if (rc != KTO_OK) { setup_free(s); return rc; }
    /* kernel.cu:13-20: relabel only for directed-by-degree */
    if (v->directedness == KTO_DIRECTED_BY_DEGREE) {
        rc = relabel_by_degree(n, s->nu, s->ua, s->ub, s->old2new);
        if (rc != KTO_OK) { setup_free(s); return rc; }
    } else {
        for (uint32_t i = 0; i < n; ++i) s->old2new[i] = i;
    }
    for (uint32_t i = 0; i < n; ++i) s->new2old[s->old2new[i]] = i;
    /* kernel_h.cu:19-25 + graph.cu:194-206: directed keeps src < dst */
    uint32_t E = v->directedness == KTO_UNDIRECTED ? 2 * s->nu : s->nu;
    uint32_t* es = (uint32_t*)malloc((E ? E : 1) * sizeof(uint32_t));
    uint32_t* ed = (uint32_t*)malloc((E ? E : 1) * sizeof(uint32_t));
    uint32_t* ei = (uint32_t*)malloc((E ? E : 1) * sizeof(uint32_t));
    if (!es || !ed || !ei) { free(es); free(ed); free(ei); setup_free(s); return KTO_ERR_NOMEM; }
    for (uint32_t e = 0; e < s->nu; ++e) {
        uint32_t a = s->old2new[s->ua[e]], b = s->old2new[s->ub[e]];
        if (v->directedness == KTO_UNDIRECTED) {
            es[2 * e] = a; ed[2 * e] = b; ei[2 * e] = e;
            es[2 * e + 1] = b; ed[2 * e + 1] = a; ei[2 * e + 1] = e;
        } else {
            es[e] = a < b ? a : b; ed[e] = a < b ? b : a; ei[e] = e;
        }
    }
    rc = kg_build(&s->g, n, E, es, ed, ei, v->format == KTO_TILED_COOCSR, v->tiles_per_dim);
    free(es); free(ed); free(ei);
    if (rc != KTO_OK) { setup_free(s); return rc; }
    s->in.node_aff = (uint8_t*)calloc(n ? n : 1, 1);
    s->in.edge_aff = (uint8_t*)calloc(E ? E : 1, 1);
    if (!s->in.node_aff || !s->in.edge_aff) { setup_free(s); return KTO_ERR_NOMEM; }
    return KTO_OK;
}

int kto_ktruss(uint32_t n, uint64_t m_sym, const uint32_t* src, const uint32_t* dst,
               uint32_t k, const kto_variant* v,
               uint32_t* out_src, uint32_t* out_dst, uint64_t cap, uint64_t* out_m,
               kto_stats* stats, kto_trace* trace) {
    if (k < 2 || !out_m) return KTO_ERR_ARG;
    kto_stats local;
    memset(&local, 0, sizeof(local));
    kto_stats* st = stats ? stats : &local;
    setup_t s;
    int rc = setup(&s, n, m_sym, src, dst, v);
    if (rc != KTO_OK) return rc;
    int exact;
    uint32_t kk = k;
    rc = peel_phase(&s.g, &kk, v, &s.in, 0, 0, NULL, st, trace, &exact);
    /* coocsr2coo (graph.cu:548-554), directed2undirectedCOO (graph.cu:239-252),
     * unsort (graph.cu:187-192) */
    const int directed = v->directedness != KTO_UNDIRECTED;
    uint64_t need = directed ? 2ull * s.g.E : s.g.E;
    *out_m = need;
    if (need > cap || (need && (!out_src || !out_dst))) { setup_free(&s); return KTO_ERR_CAPACITY; }
    for (uint32_t e = 0; e < s.g.E; ++e) {
        out_src[e] = s.new2old[s.g.src[e]];
        out_dst[e] = s.new2old[s.g.dst[e]];
    }
    if (directed) {
        for (uint32_t e = 0; e < s.g.E; ++e) {
            out_src[s.g.E + e] = s.new2old[s.g.dst[e]];
            out_dst[s.g.E + e] = s.new2old[s.g.src[e]];
        }
    }
    setup_free(&s);
    return rc;
}

int kto_decompose(uint32_t n, uint64_t m_sym, const uint32_t* src, const uint32_t* dst,
                  const kto_variant* v, uint32_t schedule,
                  uint32_t* tau, kto_stats* stats, kto_trace* trace) {
    if (m_sym && !tau) return KTO_ERR_ARG;
    kto_stats local;
    memset(&local, 0, sizeof(local));
    kto_stats* st = stats ? stats : &local;
    setup_t s;
    int rc = setup(&s, n, m_sym, src, dst, v);
    if (rc != KTO_OK) return rc;
    const int directed = v->directedness != KTO_UNDIRECTED;
    uint32_t* tau_u = (uint32_t*)calloc(s.nu ? s.nu : 1, sizeof(uint32_t));
    if (!tau_u) { setup_free(&s); return KTO_ERR_NOMEM; }
    uint32_t k = 3;
    int reuse = 0;
    while (s.g.E > 0) {
        int exact = 0;
        int prc = peel_phase(&s.g, &k, v, &s.in, reuse, schedule, tau_u, st, trace, &exact);
        if (prc != KTO_OK) rc = prc;
        reuse = directed && v->recount == KTO_RECOUNT_ALL && (schedule & KTO_SCHED_REUSE) && exact;
        ++k;
    }
    uint32_t kmax = 0;
    for (uint32_t e = 0; e < s.nu; ++e) if (tau_u[e] > kmax) kmax = tau_u[e];
    st->k_max = kmax;
    for (uint64_t i = 0; i < m_sym; ++i) tau[i] = tau_u[s.uid[i]];
    free(tau_u);
    setup_free(&s);
    return rc;
}

/* ---------------------------------------------------------------------- */
/* Brute force: one-at-a-time deletion on a bitset adjacency matrix. */
int kto_bruteforce_tau(uint32_t n, uint64_t m_sym, const uint32_t* src, const uint32_t* dst,
                       uint32_t* tau) {
    if (n > 4096 || (m_sym && (!src || !dst || !tau))) return KTO_ERR_ARG;
    uint32_t nu = 0, *ua = NULL, *ub = NULL;
    uint32_t* uid = (uint32_t*)malloc((m_sym ? m_sym : 1) * sizeof(uint32_t));
    if (!uid) return KTO_ERR_NOMEM;
    int rc = number_edges(n, m_sym, src, dst, uid, &nu, &ua, &ub);
    if (rc != KTO_OK) { free(uid); return rc; }
    const uint32_t W = (n + 63) / 64;
    uint64_t* adj = (uint64_t*)calloc((size_t)(n ? n : 1) * (W ? W : 1), sizeof(uint64_t));
    uint8_t* alive = (uint8_t*)malloc(nu ? nu : 1);
    uint32_t* tu = (uint32_t*)calloc(nu ? nu : 1, sizeof(uint32_t));
    if (!adj || !alive || !tu) { free(uid); free(ua); free(ub); free(adj); free(alive); free(tu); return KTO_ERR_NOMEM; }
    for (uint32_t e = 0; e < nu; ++e) {
        adj[(size_t)ua[e] * W + ub[e] / 64] |= 1ull << (ub[e] % 64);
        adj[(size_t)ub[e] * W + ua[e] / 64] |= 1ull << (ua[e] % 64);
        alive[e] = 1;
    }
    uint32_t remaining = nu;
    for (uint32_t k = 3; remaining > 0; ++k) {
        int any = 1;
        while (any) {
            any = 0;
            for (uint32_t e = 0; e < nu; ++e) {
                if (!alive[e]) continue;
                const uint64_t* ra = adj + (size_t)ua[e] * W;
                const uint64_t* rb = adj + (size_t)ub[e] * W;
                uint32_t c = 0;
                for (uint32_t w = 0; w < W; ++w) c += (uint32_t)__builtin_popcountll(ra[w] & rb[w]);
                if (c < k - 2) {
                    alive[e] = 0;
                    tu[e] = k - 1;
                    adj[(size_t)ua[e] * W + ub[e] / 64] &= ~(1ull << (ub[e] % 64));
                    adj[(size_t)ub[e] * W + ua[e] / 64] &= ~(1ull << (ua[e] % 64));
                    --remaining;
                    any = 1;
                }
            }
        }
    }
    for (uint64_t i = 0; i < m_sym; ++i) tau[i] = tu[uid[i]];
    free(uid); free(ua); free(ub); free(adj); free(alive); free(tu);
    return KTO_OK;
}

/* ---------------------------------------------------------------------- */
/* createCOOFromFile (graph.cu:37-79): whitespace-separated "dst src w"
 * triples, 1-based; n = largest id seen. */
int kto_read_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst) {
    if (!path || !n || !m || !src || !dst) return KTO_ERR_ARG;
    FILE* fp = fopen(path, "r");
    if (!fp) return KTO_ERR_IO;
    uint64_t cap = 1u << 20, cnt = 0;
    uint32_t* s = (uint32_t*)malloc(cap * sizeof(uint32_t));
    uint32_t* d = (uint32_t*)malloc(cap * sizeof(uint32_t));
    uint32_t nn = 0, a, b, x;
    int rc = KTO_OK;
    while (s && d && fscanf(fp, "%u", &b) == 1) {
        if (fscanf(fp, "%u", &a) != 1 || fscanf(fp, "%u", &x) != 1) { rc = KTO_ERR_IO; break; }
        if (a == b) { rc = KTO_ERR_SELF_LOOP; break; }
        if (a == 0 || b == 0) { rc = KTO_ERR_RANGE; break; }
        if (cnt == cap) {
            cap *= 2;
            uint32_t* s2 = (uint32_t*)realloc(s, cap * sizeof(uint32_t));
            uint32_t* d2 = (uint32_t*)realloc(d, cap * sizeof(uint32_t));
            if (s2) s = s2;
            if (d2) d = d2;
            if (!s2 || !d2) { rc = KTO_ERR_NOMEM; break; }
        }
        s[cnt] = a - 1;
        d[cnt] = b - 1;
        ++cnt;
        if (a > nn) nn = a;
        if (b > nn) nn = b;
    }
    fclose(fp);
    if (!s || !d) rc = KTO_ERR_NOMEM;
    if (rc != KTO_OK) { free(s); free(d); return rc; }
    *n = nn; *m = cnt; *src = s; *dst = d;
    return KTO_OK;
}

/* writeCOOGraphToFile (graph.cu:96-104) */
int kto_write_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst) {
    FILE* fp = fopen(path, "w");
    if (!fp) return KTO_ERR_IO;
    for (uint64_t e = 0; e < m; ++e) fprintf(fp, "%u\t%u\t%u\n", dst[e] + 1, src[e] + 1, 1u);
    fclose(fp);
    return KTO_OK;
}

typedef struct { uint32_t s, d; } sd_t;
static int cmp_sd(const void* x, const void* y) {
    const sd_t* p = (const sd_t*)x;
    const sd_t* q = (const sd_t*)y;
    if (p->s != q->s) return p->s < q->s ? -1 : 1;
    return p->d < q->d ? -1 : (p->d > q->d);
}

/* verify (verify.cu:12-40): same count, then both put in COO+CSR order
 * (rows by source, each row sorted) and compared entry by entry. */
int kto_verify(uint32_t n, uint64_t m_a, const uint32_t* src_a, const uint32_t* dst_a,
               uint64_t m_b, const uint32_t* src_b, const uint32_t* dst_b) {
    (void)n;
    if (m_a != m_b) return 1;
    sd_t* a = (sd_t*)malloc((m_a ? m_a : 1) * sizeof(sd_t));
    sd_t* b = (sd_t*)malloc((m_b ? m_b : 1) * sizeof(sd_t));
    if (!a || !b) { free(a); free(b); return KTO_ERR_NOMEM; }
    for (uint64_t i = 0; i < m_a; ++i) { a[i].s = src_a[i]; a[i].d = dst_a[i]; }
    for (uint64_t i = 0; i < m_b; ++i) { b[i].s = src_b[i]; b[i].d = dst_b[i]; }
    qsort(a, m_a, sizeof(sd_t), cmp_sd);
    qsort(b, m_b, sizeof(sd_t), cmp_sd);
    int diff = 0;
    for (uint64_t i = 0; i < m_a && !diff; ++i) diff = a[i].s != b[i].s || a[i].d != b[i].d;
    free(a); free(b);
    return diff;
}
