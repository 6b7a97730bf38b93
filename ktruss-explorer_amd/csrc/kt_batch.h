// kt_batch.h -- batched decomposition of many graphs (device side).
//
// Device layout of a batch of G graphs (uint32 unless noted):
//   ptr[N+1], src[E], dst[E]   disjoint union of the graphs' oriented COO+CSR
//                              layouts; graph g owns vertices [vbeg[g], ...)
//                              and entries [gbeg[g], gend[g])
//   sup[E], tau[E]             support and truss number per entry position
//   gk, gflag, glive, gndel,   per-graph peel state: current k, flags,
//   gmin                       live entries, deletions of the last mark, min
//                              surviving support of the last mark
//   giter, gphase, gkmax,      per-graph statistics
//   gpass
//   dlist/dscan, alist/ascan   graphs to recount (dirty) / still peeling
//                              (active) and exclusive scans of their tasks
//   ddesc, adesc (uint4)       task descriptors {begin, end, graph, 0}
//   sptr, snbr, seid, sa, sb   DELTA only: the union of the graphs' full
//                              adjacency lists (see kt_delta.hip), dl the
//                              entries deleted by the current step
#pragma once

#include "kt_internal.h"

#define KT_BF_DIRTY 1u  // supports must be recounted before the next mark
#define KT_BF_DONE 2u   // no live entry left

namespace kt {

struct BatchCounters {
    DevCounters count;                 // the count kernel's merge / recount totals
    unsigned int ndirty, nactive;      // list lengths for the next step
    unsigned int dtasks, atasks;       // task counts of the two lists
    unsigned int steps;                // mark passes run (control mode 2)
    unsigned int pad[3];
    unsigned long long live_in_passes; // live entries over all count passes
};

struct BatchDev {
    uint32_t G;
    const uint32_t* gbeg;
    const uint32_t* gend;
    uint32_t *gk, *gflag, *glive, *gndel, *gmin;
    uint32_t *giter, *gphase, *gkmax, *gpass;
    uint32_t *dlist, *dscan, *alist, *ascan;
    uint4 *ddesc, *adesc;
    BatchCounters* bc;
    uint32_t cchunk;    // entries per count / zero task
    uint32_t mchunk;    // entries per mark task
    uint32_t schedule;  // KT_SCHED_* bits
    uint32_t delta;     // supports maintained by DELTA updates: never recount
    unsigned int* zero_cnt;  // DELTA: the deletion count the next mark fills (control clears it)
};

namespace dev {
hipError_t batch_control(const BatchDev& b, uint32_t mode, hipStream_t s);
hipError_t batch_expand(const BatchDev& b, uint32_t blocks, hipStream_t s);
hipError_t batch_zero(const TaskList& tl, const uint32_t* dst, uint32_t* sup, uint32_t blocks,
                      hipStream_t s);
// dl != nullptr: DELTA marking (KT_DYING on the entry and its two adjacency
// slots snbr[sa[e]], snbr[sb[e]]; deleted entries appended to dl, count *dn;
// src: the entries' rows; *bad counts entries below the threshold that were
// flagged already, none while deleted entries hold KT_SUP_GONE);
// fin_dl / fin_nd: the previous step's deletions, whose update is done,
// become KT_DEAD in the same launch (workgroups beyond `blocks`).
hipError_t batch_mark(const TaskList& tl, uint32_t* dst, uint32_t* sup, uint32_t* tau,
                      const uint32_t* gk, uint32_t* gndel, uint32_t* gmin, uint32_t* dl,
                      unsigned int* dn, uint32_t* snbr, const uint32_t* sa, const uint32_t* sb,
                      const uint32_t* fin_dl, const unsigned int* fin_nd, const uint32_t* src,
                      unsigned int* bad, uint32_t blocks, hipStream_t s);
hipError_t batch_min(const TaskList& tl, const uint32_t* dst, const uint32_t* sup, uint32_t* gmin,
                     uint32_t blocks, hipStream_t s);
}  // namespace dev

}  // namespace kt
