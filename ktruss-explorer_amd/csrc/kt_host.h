// kt_host.h -- host-side helpers of the k-truss engine.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "ktruss.h"

struct kt_graph;

namespace kt {

const char* last_error();
unsigned num_threads();
void parallel_for(size_t n, size_t grain, const std::function<void(size_t, size_t)>& body);

// Stored COO+CSR layout of one orientation, built on the host.
struct HostLayout {
    uint32_t n = 0;
    uint32_t E = 0;                  // stored entries
    bool directed = true;
    uint32_t max_row = 0;
    std::vector<uint32_t> ptr;       // n + 1
    std::vector<uint32_t> src, dst;  // E, rows sorted by dst
    std::vector<uint32_t> old2new, new2old;
    std::vector<uint32_t> pos_of_edge;  // undirected edge id -> one stored position
};

int graph_create(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, uint32_t flags,
                 kt_graph** out);
int build_layout(const kt_graph* g, int directedness, HostLayout& L);
int build_adjacency(uint32_t n, uint32_t E, const uint32_t* ptr, const uint32_t* src,
                    const uint32_t* dst, std::vector<uint32_t>& sptr, std::vector<uint32_t>& snbr,
                    std::vector<uint32_t>& seid, std::vector<uint32_t>& sa,
                    std::vector<uint32_t>& sb);

int read_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst);
int write_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst);
int read_truss_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst, uint32_t** tau);
int write_truss_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst, const uint32_t* tau);
int write_edges_bin(const char* path, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                    uint32_t flags);
int read_edges_bin(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst, uint32_t* flags);
int is_edges_bin(const char* path);
int verify_edges(uint64_t ma, const uint32_t* sa, const uint32_t* da, uint64_t mb,
                 const uint32_t* sb, const uint32_t* db);

int generate_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed,
                  uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst);
int generate_erdos_renyi(uint32_t n, double avg_degree, uint64_t seed, uint64_t* m,
                         uint32_t** src, uint32_t** dst);
int generate_chung_lu(uint32_t n, double gamma, double avg_degree, double max_weight,
                      uint32_t num_cliques, uint32_t cmin, uint32_t cmax, uint64_t seed,
                      uint64_t* m, uint32_t** src, uint32_t** dst);

}  // namespace kt
