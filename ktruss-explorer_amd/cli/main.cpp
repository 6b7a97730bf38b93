// main.cpp -- the ktruss command-line tool on the MI355X engine.
//
// Flow of the reference's main (main.cu:13-53): set up (read graph), print
// the configuration, compute, optionally write the result, verify it against
// the -c file, clean up -- with the same printed labels.  The computation
// goes through the C ABI (include/ktruss.h) only.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "config.h"
#include "ktruss.h"
#include "timer.h"

using namespace cli;

namespace {

int report(int rc, const char* what) {
    std::fprintf(stderr, "%s failed: %s\n", what, kt_last_error());
    return rc;
}

std::vector<std::string> split(const char* s, char sep) {
    std::vector<std::string> out;
    std::string cur;
    for (const char* p = s; *p; ++p) {
        if (*p == sep) {
            out.push_back(cur);
            cur.clear();
        } else {
            cur += *p;
        }
    }
    out.push_back(cur);
    return out;
}

// Synthetic input: returns unique pairs (KT_INPUT_PAIRS).
int make_synthetic(const char* spec, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst) {
    auto f = split(spec, ':');
    if (f[0] == "rmat" && f.size() == 4) {
        return kt_generate_rmat((uint32_t)std::stoul(f[1]), (uint32_t)std::stoul(f[2]), 0.57, 0.19,
                                0.19, std::stoull(f[3]), n, m, src, dst);
    }
    if (f[0] == "er" && f.size() == 4) {
        *n = (uint32_t)std::stoul(f[1]);
        return kt_generate_erdos_renyi(*n, std::stod(f[2]), std::stoull(f[3]), m, src, dst);
    }
    if (f[0] == "cl" && (f.size() == 5 || f.size() == 9)) {
        *n = (uint32_t)std::stoul(f[1]);
        double maxw = 0;
        uint32_t cliques = 0, cmin = 0, cmax = 0;
        if (f.size() == 9) {
            maxw = std::stod(f[5]);
            cliques = (uint32_t)std::stoul(f[6]);
            cmin = (uint32_t)std::stoul(f[7]);
            cmax = (uint32_t)std::stoul(f[8]);
        }
        return kt_generate_chung_lu(*n, std::stod(f[2]), std::stod(f[3]), maxw, cliques, cmin, cmax,
                                    std::stoull(f[4]), m, src, dst);
    }
    std::fprintf(stderr, "Unrecognized -s spec: %s\n", spec);
    return KT_ERR_ARG;
}

}  // namespace

// The -v 2 lines of one run (kernel_ddc.cu:126-198, kernel_ddt.cu:135-207,
// kernel_duc.cu:105-174, kernel_dut.cu:116-186; the undirected COO+CSR
// variant says "Mark deleted edges time" and prints no "# edges remaining"),
// from the plan's iteration log.  A decomposition heads each phase with its
// k.  A DELTA iteration's support pass is its update; its mark is fused with
// the next kernels and has no line.
static void print_iterations(const kt_plan* plan, const Config& c) {
    uint64_t len = 0;
    kt_plan_iteration_log(plan, nullptr, 0, &len);
    std::vector<uint64_t> lg((size_t)KT_ITER_REC * len);
    kt_plan_iteration_log(plan, lg.data(), len, &len);
    const bool duc = c.directedness == KT_UNDIRECTED && c.inputFormat == KT_COOCSR;
    for (uint64_t r = 0; r < len; ++r) {
        const uint64_t* x = &lg[(size_t)KT_ITER_REC * r];
        const uint64_t fl = x[10];
        if (c.decompose && x[1] == 0) std::printf("        k = %llu\n", (unsigned long long)x[0]);
        std::printf("        Iteration %llu\n", (unsigned long long)x[1]);
        if (fl & KT_ITER_COUNTED) Timer::print(x[4] * 1e-6, "            Triangle counting time");
        if (fl & KT_ITER_MARK_TIMED)
            Timer::print(x[5] * 1e-6, duc ? "            Mark deleted edges time" : "            Mark deleted edges");
        if (fl & KT_ITER_REMOVED) {
            Timer::print(x[6] * 1e-6, "            Remove deleted edges");
            if (!duc) std::printf("                # edges remaining = %llu\n", (unsigned long long)x[7]);
        }
        if (fl & KT_ITER_AFFECTED) Timer::print(x[8] * 1e-6, "            Mark affected edges");
        if (fl & KT_ITER_FINAL_REMOVAL) Timer::print(x[9] * 1e-6, "        Remove deleted edges");
    }
}

// The -v 1 lines after the runs (kernel_ddc.cu:249-254 and twins,
// kernel_d.cu:60-84, kernel.cu:36): plan stage times sm, frees fm.
static void print_conversion_stages(const double* sm, const double* fm, const Config& c) {
    const bool directed = c.directedness != KT_UNDIRECTED;
    const bool tiled = c.inputFormat == KT_TILED_COOCSR;
    if (!c.decompose) Timer::print(sm[6], tiled ? "    Convert tiled COOCSR to COO" : "    Convert COO/CSR to COO");
    Timer::print(fm[0], tiled ? "    Deallocate tiled COOCSR" : "    Deallocate COOCSR");
    if (directed) {
        if (!c.decompose) Timer::print(sm[7], "    Convert directed COO to undirected COO");
        Timer::print(fm[1], "    Deallocate directed COO");
    }
    Timer::print(directed ? fm[2] : fm[1] + fm[2], "    Deallocate intermediate data structures");
    if (!c.decompose) Timer::print(sm[8], "    Copy COO truss from GPU");
    Timer::print(sm[5], "    Deallocate COO graph and truss from GPU");
    if (c.directedness == KT_DIRECTED_BY_DEGREE && !c.decompose)
        Timer::print(sm[9], "    Restore original vertex labels (not optimized)");
}

// Cost model written by learn.export_model_bin, little-endian: "KTMODEL1",
// u32 candidates C, u32 parameter count P, C x (kernel, block, cap, chunk)
// u32, P float32 parameters, KT_MLP_IN lower then upper input bounds.  Sets
// the plan's candidates and its per-phase pick; the model is returned for
// kt_mlp_destroy.
static int load_model(const char* path, kt_plan* plan, kt_mlp** out) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) {
        std::fprintf(stderr, "cannot open %s\n", path);
        return KT_ERR_IO;
    }
    char magic[8];
    uint32_t hdr[2] = {0, 0};
    bool ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, "KTMODEL1", 8) == 0 &&
              std::fread(hdr, 4, 2, fp) == 2;
    const uint32_t C = hdr[0], P = hdr[1];
    ok = ok && C >= 1 && C <= KT_MAX_CANDIDATES && C <= KT_MLP_MAX_OUT && P < (1u << 24);
    std::vector<uint32_t> raw(ok ? 4 * (size_t)C : 0);
    std::vector<float> params(ok ? P : 0), lo(KT_MLP_IN), hi(KT_MLP_IN);
    ok = ok && std::fread(raw.data(), 4, raw.size(), fp) == raw.size() &&
         std::fread(params.data(), 4, P, fp) == P && std::fread(lo.data(), 4, KT_MLP_IN, fp) == KT_MLP_IN &&
         std::fread(hi.data(), 4, KT_MLP_IN, fp) == KT_MLP_IN;
    std::fclose(fp);
    if (!ok) {
        std::fprintf(stderr, "%s is not a cost model file\n", path);
        return KT_ERR_IO;
    }
    std::vector<kt_count_config> cands(C);
    for (uint32_t i = 0; i < C; ++i) cands[i] = {raw[4 * i], raw[4 * i + 1], raw[4 * i + 2], raw[4 * i + 3]};
    int rc;
    if ((rc = kt_mlp_create(C, 1, out))) return rc;
    if ((rc = kt_mlp_set_params(*out, params.data(), P))) return rc;
    if ((rc = kt_mlp_set_input_range(*out, lo.data(), hi.data()))) return rc;
    if ((rc = kt_plan_set_candidates(plan, cands.data(), C))) return rc;
    return kt_plan_set_pick(plan, KT_PICK_MODEL, *out);
}

int main(int argc, char** argv) {
    Config c;
    const int prc = parseArgs(argc, argv, c);
    if (prc == 1) return 0;
    if (prc) return prc;
    if (c.processor == CPU) {
        std::fprintf(stderr, "-p cpu: this build runs the peel on the GPU only\n");
        return 2;
    }
    Timer timer(c.verbosity >= 1);

    // Set up
    timer.printAndStart("Setting up\n");
    uint32_t n = 0;
    uint64_t m = 0;
    uint32_t *src = nullptr, *dst = nullptr;
    uint32_t flags = KT_INPUT_SYMMETRIC;
    int rc;
    if (c.synthetic) {
        rc = make_synthetic(c.synthetic, &n, &m, &src, &dst);
        flags = KT_INPUT_PAIRS;
    } else if (kt_is_edges_bin(c.graphFileName)) {
        rc = kt_read_edges_bin(c.graphFileName, &n, &m, &src, &dst, &flags);
    } else {
        rc = kt_read_tsv(c.graphFileName, &n, &m, &src, &dst);
    }
    if (rc) return report(1, "reading the graph");
    if (c.binCacheName && (rc = kt_write_edges_bin(c.binCacheName, n, m, src, dst, flags)))
        return report(1, "writing the binary edge cache");
    // the device first: large inputs are then checked and numbered on it
    if ((rc = kt_init(0))) return report(1, "kt_init");
    kt_graph* graph = nullptr;
    if ((rc = kt_graph_create(n, m, src, dst, flags, &graph))) return report(1, "kt_graph_create");
    uint64_t m_und = 0, m_in = 0;
    kt_graph_info(graph, nullptr, &m_und, &m_in, nullptr);
    c.numNodes = n;
    c.numEdges = (unsigned)(2 * m_und);
    timer.stopAndPrint("    Set up time");
    if (c.verbosity >= 1) printConfig(c);

    // Compute: the flow and -v 1 labels of kernel.cu:9-39, kernel_d.cu:11-86
    // and the variant wrappers (kernel_ddc.cu:216-256 and twins), each label
    // printed with the time of the plan stage that does that work
    Timer total(c.verbosity >= 1);
    total.printAndStart("Computing\n");
    kt_variant v;
    kt_variant_default(&v);
    v.directedness = c.directedness;
    v.format = c.inputFormat;
    v.recount = c.recount;
    v.num_edge_remove_iter = c.numEdgeRemoveIter;
    v.block_size = c.blockSize;
    v.tiles_per_dim = c.numTiles;
    v.num_parallel_tiles = c.numParallelTiles;
    v.count_kernel = (uint32_t)c.countKernel;
    kt_plan* plan = nullptr;
    if ((rc = kt_plan_create(graph, &v, &plan))) return report(1, "kt_plan_create");
    kt_mlp* model = nullptr;
    if (c.modelFileName && (rc = load_model(c.modelFileName, plan, &model))) return report(1, "loading the cost model");
    const bool directed = c.directedness != KT_UNDIRECTED;
    const bool tiled = c.inputFormat == KT_TILED_COOCSR;
    const bool by_degree = c.directedness == KT_DIRECTED_BY_DEGREE;
    double sm[KT_PLAN_STAGES] = {};
    kt_plan_stage_times(plan, sm, KT_PLAN_STAGES);
    if (c.verbosity >= 1) {
        if (by_degree) Timer::print(sm[0], "    Relabel vertices (not optimized)");          // kernel.cu:19
        Timer::print(sm[1], "    Copy COO graph to GPU");                                      // kernel_d.cu:20
        Timer::print(sm[2], "    Allocate intermediate data structures");                      // kernel_d.cu:25
        if (directed) Timer::print(sm[3], "    Convert undirected COO to directed COO");      // kernel_d.cu:33
        // kernel_ddc.cu:223, kernel_ddt.cu:234, kernel_duc.cu:201, kernel_dut.cu:213
        const char* build = directed ? (tiled ? "    Convert directed COO to tiled COOCSR (not optimized)"
                                              : "    Convert directed COO to COOCSR (not optimized)")
                                     : (tiled ? "    Convert undirected COO to tiled COOCSR (not optimized)"
                                              : "    Convert undirected COO to COOCSR (not optimized)");
        Timer::print(sm[4], build);
    }

    std::vector<uint32_t> out_src(2 * m_und + 1), out_dst(2 * m_und + 1), tau;
    uint64_t out_m = 0;
    kt_stats st{};
    Timer t(c.verbosity >= 1);
    for (unsigned i = 0; i < c.numWarmupRuns + c.numTimedRuns; ++i) {
        t.printAndStart("    Performing K-truss\n");
        if (c.decompose) rc = kt_plan_decompose(plan, &st);
        else rc = kt_plan_ktruss(plan, c.k, out_src.data(), out_dst.data(), out_src.size(), &out_m, &st);
        if (rc) return report(1, c.decompose ? "kt_plan_decompose" : "kt_plan_ktruss");
        if (c.verbosity >= 2) print_iterations(plan, c);
        if (c.verbosity == 0 && i >= c.numWarmupRuns) {
            printConfigAsCSV(c);
            std::printf("%f ms\n", st.ms_total);
        }
        if (c.verbosity >= 1) {
            Timer::print(st.ms_total, "        Total K-truss time", GREEN);
            if (c.decompose) std::printf("        k_max = %u, phases = %llu, iterations = %llu\n", st.k_max,
                                         (unsigned long long)st.phases, (unsigned long long)st.iterations);
        }
    }
    if (c.decompose) {
        Timer tt(c.verbosity >= 1);
        tt.start();
        tau.resize(m_in);
        if ((rc = kt_plan_truss_numbers(plan, tau.data(), m_in))) return report(1, "kt_plan_truss_numbers");
        tt.stopAndPrint("    Copy truss numbers from GPU");
        // the k-truss for verification: entries with truss number >= k, both directions
        out_m = 0;
        for (uint64_t e = 0; e < m_in; ++e) {
            if (tau[e] < c.k) continue;
            out_src[out_m] = src[e];
            out_dst[out_m] = dst[e];
            ++out_m;
            if (flags == KT_INPUT_PAIRS) {
                out_src[out_m] = dst[e];
                out_dst[out_m] = src[e];
                ++out_m;
            }
        }
    }
    // the results are on the host: the plan's device arrays go now, in the
    // reference's order of frees
    double fm[3] = {0, 0, 0};
    kt_plan_destroy_timed(plan, fm);
    plan = nullptr;
    if (c.verbosity >= 1) {
        print_conversion_stages(sm, fm, c);
    }
    total.stopAndPrint("    Total time");

    // Write result
    if (c.outFileName) {
        if (c.verbosity >= 1) std::printf("Writing result to file: %s\n", c.outFileName);
        if (c.decompose) {
            if ((rc = kt_write_truss_tsv(c.outFileName, m_in, src, dst, tau.data())))
                return report(1, "kt_write_truss_tsv");
        } else if ((rc = kt_write_tsv(c.outFileName, out_m, out_src.data(), out_dst.data()))) {
            return report(1, "kt_write_tsv");
        }
    }

    // Verify result (verify.cu:12-40)
    int status = 0;
    if (std::strcmp(c.compareFileName, "none") != 0) {
        timer.printAndStart("Verifying result\n");
        uint32_t rn = 0;
        uint64_t rm = 0;
        uint32_t *rs = nullptr, *rd = nullptr;
        if ((rc = kt_read_tsv(c.compareFileName, &rn, &rm, &rs, &rd))) return report(1, "reading the compare file");
        const int diff = kt_verify_edges(out_m, out_src.data(), out_dst.data(), rm, rs, rd);
        kt_free(rs);
        kt_free(rd);
        if (diff != 0) {
            std::printf("    FAILED\n");
            status = 1;
        } else if (c.verbosity >= 1) {
            std::printf("    PASSED\n");
        }
        timer.stopAndPrint("    Verification time");
    }

    // Free memory
    timer.printAndStart("Cleaning up\n");
    if (model) kt_mlp_destroy(model);
    kt_graph_destroy(graph);
    kt_free(src);
    kt_free(dst);
    timer.stopAndPrint("    Clean up time");
    return status;
}
