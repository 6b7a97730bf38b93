// config.h -- command-line configuration of the ktruss tool.
//
// Same option letters, meanings and defaults as the reference's Config /
// parseArgs (config.h:163-224 there); additions: -D (full decomposition),
// -s (synthetic input instead of -g), -c none (skip verification), -B (write
// the input as a binary edge cache; -g reads such a cache by its magic), -K
// (support-pass kernel, KT_COUNT_*; default auto, the reference's count), -M
// (the learned cost model's per-phase pick among its candidates).
// Unlike the reference (which prints and exits with status 0 on a bad
// option), a bad option exits with status 2.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <getopt.h>

#include "ktruss.h"

namespace cli {

enum Processor { CPU, GPU };

struct Config {
    const char* graphFileName = "data/loc-gowalla_edges_adj.tsv";
    unsigned k = 3;
    const char* compareFileName = "data/loc-gowalla_edges_adj.tsv.3.reference.tsv";
    const char* outFileName = nullptr;
    unsigned numWarmupRuns = 0;
    unsigned numTimedRuns = 1;
    Processor processor = GPU;
    int directedness = KT_DIRECTED_BY_DEGREE;
    int inputFormat = KT_COOCSR;
    unsigned numEdgeRemoveIter = 0;
    int recount = KT_RECOUNT_ALL;
    unsigned blockSize = 512;
    unsigned numTiles = 8;
    unsigned numParallelTiles = 8;
    unsigned verbosity = 1;
    unsigned numNodes = 0;
    unsigned numEdges = 0;
    // additions
    bool decompose = false;
    const char* synthetic = nullptr;  // rmat:SCALE:EF:SEED | er:N:DEG:SEED | cl:N:GAMMA:DEG:SEED
    const char* binCacheName = nullptr;  // -B: binary edge cache written from the input
    int countKernel = KT_COUNT_AUTO;     // -K
    const char* modelFileName = nullptr;  // -M: cost model (picker/count_picker.ktm)
};

inline void usage() {
    std::fprintf(stderr,
        "\nUsage:  ./ktruss [options]\n"
        "\nOptions:\n"
        "    -g <graphFileName>       input graph, one \"dst src w\" triple per line, 1-based, or a\n"
        "                             binary edge cache written by -B (default =\n"
        "                             data/loc-gowalla_edges_adj.tsv)\n"
        "    -B <binFileName>         also write the input as a binary edge cache\n"
        "    -k <k>                   k (default = 3)\n"
        "    -c <compareFileName>     file the result is verified against; \"none\" skips the check\n"
        "                             (default = data/loc-gowalla_edges_adj.tsv.3.reference.tsv)\n"
        "    -o <outFileName>         write the result to this file (default: not written)\n"
        "    -w <numWarmupRuns>       warmup runs (default = 0)\n"
        "    -x <numTimedRuns>        timed runs (default = 1)\n"
        "    -p <processor>           gpu (default); cpu is not served by this build\n"
        "    -d <directedness>        undirected | directed-by-index | directed-by-degree (default)\n"
        "    -f <inputFormat>         coocsr (default) | tiled-coocsr\n"
        "    -e <numEdgeRemoveIter>   iterations after which deleted edges are compacted away (default = 0)\n"
        "    -r <recount>             all (default) | affected\n"
        "    -b <blockSize>           threads per block of the counting kernel (default = 512)\n"
        "    -t <numTiles>            tiles per dimension for tiled-coocsr (default = 8)\n"
        "    -y <numParallelTiles>    x-tile classes counted in parallel for tiled-coocsr (default = 8)\n"
        "    -v <verbosity>           0: k-truss time as CSV, 1: phase times (default), 2: per-iteration\n"
        "    -D                       full decomposition: truss number of every edge (-o writes\n"
        "                             \"dst src truss\" lines; the k-truss verified is the\n"
        "                             entries with truss number >= k)\n"
        "    -K <countKernel>         support passes: auto (default) | row | search | merge |\n"
        "                             delta (maintained supports; directed-by-degree or\n"
        "                             directed-by-index, coocsr, recount all)\n"
        "    -M <modelFile>           per-phase pick of the support-pass kernel by the learned\n"
        "                             cost model (picker/count_picker.ktm)\n"
        "    -s <spec>                synthetic input instead of -g: rmat:SCALE:EF:SEED,\n"
        "                             er:N:AVGDEG:SEED or cl:N:GAMMA:AVGDEG:SEED\n"
        "    -h                       help\n\n");
}

inline bool parse_enum(const char* s, const char* const* names, int count, int* out) {
    for (int i = 0; i < count; ++i)
        if (std::strcmp(s, names[i]) == 0) {
            *out = i;
            return true;
        }
    return false;
}

static const char* const kDirectedness[] = {"undirected", "directed-by-index", "directed-by-degree"};
static const char* const kFormats[] = {"coocsr", "tiled-coocsr"};
static const char* const kRecounts[] = {"all", "affected"};
// indexed by KT_COUNT_* (0..4)
static const char* const kCountKernels[] = {"auto", "row", "search", "merge", "delta"};
static const char* const kProcessors[] = {"cpu", "gpu"};

inline int parseArgs(int argc, char** argv, Config& c) {
    int opt;
    while ((opt = getopt(argc, argv, "g:k:o:w:x:c:p:d:f:e:r:b:t:y:v:hDs:B:K:M:")) >= 0) {
        int e = 0;
        switch (opt) {
            case 'g': c.graphFileName = optarg; break;
            case 'k': c.k = (unsigned)std::atoi(optarg); break;
            case 'c': c.compareFileName = optarg; break;
            case 'o': c.outFileName = optarg; break;
            case 'w': c.numWarmupRuns = (unsigned)std::atoi(optarg); break;
            case 'x': c.numTimedRuns = (unsigned)std::atoi(optarg); break;
            case 'p':
                if (!parse_enum(optarg, kProcessors, 2, &e)) {
                    std::fprintf(stderr, "Unrecognized -p option: %s\n", optarg);
                    return 2;
                }
                c.processor = (Processor)e;
                break;
            case 'd':
                if (!parse_enum(optarg, kDirectedness, 3, &c.directedness)) {
                    std::fprintf(stderr, "Unrecognized -d option: %s\n", optarg);
                    return 2;
                }
                break;
            case 'f':
                if (!parse_enum(optarg, kFormats, 2, &c.inputFormat)) {
                    std::fprintf(stderr, "Unrecognized -f option: %s\n", optarg);
                    return 2;
                }
                break;
            case 'e': c.numEdgeRemoveIter = (unsigned)std::atoi(optarg); break;
            case 'r':
                if (!parse_enum(optarg, kRecounts, 2, &c.recount)) {
                    std::fprintf(stderr, "Unrecognized -r option: %s\n", optarg);
                    return 2;
                }
                break;
            case 'b': c.blockSize = (unsigned)std::atoi(optarg); break;
            case 't': c.numTiles = (unsigned)std::atoi(optarg); break;
            case 'y': c.numParallelTiles = (unsigned)std::atoi(optarg); break;
            case 'v': c.verbosity = (unsigned)std::atoi(optarg); break;
            case 'D': c.decompose = true; break;
            case 's': c.synthetic = optarg; break;
            case 'B': c.binCacheName = optarg; break;
            case 'M': c.modelFileName = optarg; break;
            case 'K':
                if (!parse_enum(optarg, kCountKernels, 5, &c.countKernel)) {
                    std::fprintf(stderr, "Unrecognized -K option: %s\n", optarg);
                    return 2;
                }
                break;
            case 'h': usage(); return 1;
            default:
                std::fprintf(stderr, "\nUnrecognized option!\n");
                usage();
                return 2;
        }
    }
    return 0;
}

inline void printConfig(const Config& c) {
    std::printf("    Graph: %s\n", c.synthetic ? c.synthetic : c.graphFileName);
    std::printf("        # vertices = %u\n", c.numNodes);
    std::printf("        # edges = %u\n", c.numEdges);
    if (c.decompose) std::printf("    k: all (decomposition)\n");
    else std::printf("    k: %u\n", c.k);
    std::printf("    Processor: %s\n", kProcessors[c.processor]);
    if (c.processor == GPU) std::printf("        Block size = %u\n", c.blockSize);
    std::printf("    Directedness: %s\n", kDirectedness[c.directedness]);
    std::printf("    Input format: %s\n", kFormats[c.inputFormat]);
    if (c.inputFormat == KT_TILED_COOCSR) {
        std::printf("        Number of tiles = %u\n", c.numTiles);
        if (c.processor == GPU)
            std::printf("        Number of tiles processed in parallel = %u\n", c.numParallelTiles);
    }
    std::printf("    Number of initial iterations to remove edges: %u\n", c.numEdgeRemoveIter);
    std::printf("    Recount strategy: %s\n", kRecounts[c.recount]);
    if (c.countKernel != KT_COUNT_AUTO) std::printf("    Count kernel: %s\n", kCountKernels[c.countKernel]);
    if (c.modelFileName) std::printf("    Count kernel per phase: cost model %s\n", c.modelFileName);
}

inline void printConfigAsCSV(const Config& c) {
    std::printf("%s, %u, %u, %u, %s, %s, %s, %u, %s, ", c.synthetic ? c.synthetic : c.graphFileName,
                c.numNodes, c.numEdges, c.k, kProcessors[c.processor], kDirectedness[c.directedness],
                kFormats[c.inputFormat], c.numEdgeRemoveIter, kRecounts[c.recount]);
    if (c.processor == GPU) std::printf("%u, ", c.blockSize);
    else std::printf("N/A, ");
    if (c.inputFormat == KT_TILED_COOCSR) {
        std::printf("%u, ", c.numTiles);
        if (c.processor == GPU) std::printf("%u, ", c.numParallelTiles);
        else std::printf("N/A, ");
    } else {
        std::printf("N/A, N/A, ");
    }
}

}  // namespace cli
