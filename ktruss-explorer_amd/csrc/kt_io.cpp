// kt_io.cpp -- graph files: the reference's TSV edge list (createCOOFromFile,
// graph.cu:37-79; writeCOOGraphToFile, graph.cu:96-104), the same layout
// carrying truss numbers in its third column, and a binary edge cache.
//
// Text is parsed and formatted by all host threads: the file is split at
// whitespace into chunks, each chunk's tokens are counted, a prefix sum
// places every token (token t is column t % 3 of line t / 3), and a second
// pass parses in place.  Errors are the ones a front-to-back reader meets
// first: the lowest chunk's token error, then a truncated triple, then the
// lowest line with a self-loop or a 0 id.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kt_host.h"
#include "kt_internal.h"

namespace kt {

namespace {

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }

// Read-only view of a whole file (mapped; empty files map nothing).
struct FileView {
    const char* p = nullptr;
    size_t size = 0;
    int fd = -1;
    ~FileView() {
        if (p) munmap((void*)p, size);
        if (fd >= 0) close(fd);
    }
    int open_file(const char* path) {
        fd = ::open(path, O_RDONLY);
        if (fd < 0) return fail(KT_ERR_IO, std::string("cannot open ") + path);
        struct stat st;
        if (fstat(fd, &st) != 0) return fail(KT_ERR_IO, std::string("cannot stat ") + path);
        size = (size_t)st.st_size;
        if (size == 0) return KT_OK;
        void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (m == MAP_FAILED) {
            size = 0;
            return fail(KT_ERR_IO, std::string("cannot map ") + path);
        }
        madvise(m, size, MADV_SEQUENTIAL);
        p = (const char*)m;
        return KT_OK;
    }
};

enum : int { kTokOk = 0, kTokNonNumeric = 1, kTokRange = 2 };

// Parses the tokens of [b, e); out(t, v) receives token t (0-based within the
// chunk).  Returns the chunk's token count, stopping at its first error.
template <class F>
uint64_t scan_tokens(const char* b, const char* e, int& err, F&& out) {
    uint64_t t = 0;
    const char* p = b;
    while (p < e) {
        while (p < e && is_space(*p)) ++p;
        if (p >= e) break;
        if (*p < '0' || *p > '9') {
            err = kTokNonNumeric;
            return t;
        }
        uint64_t v = 0;
        while (p < e && *p >= '0' && *p <= '9') {
            v = v * 10 + (uint64_t)(*p - '0');
            if (v > 0xFFFFFFFFull) {
                err = kTokRange;
                return t;
            }
            ++p;
        }
        out(t, (uint32_t)v);
        ++t;
    }
    return t;
}

int parse_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst,
              uint32_t** third) {
    if (!path || !n || !m || !src || !dst) return fail(KT_ERR_ARG, "NULL argument");
    FileView f;
    if (int rc = f.open_file(path)) return rc;
    const char* buf = f.p;
    const size_t sz = f.size;
    // chunk bounds on whitespace (or the end)
    const size_t C = sz < (1u << 20) ? 1 : (size_t)num_threads() * 4;
    std::vector<size_t> bound(C + 1, sz);
    bound[0] = 0;
    for (size_t c = 1; c < C; ++c) {
        size_t q = std::max(bound[c - 1], sz / C * c);
        while (q < sz && !is_space(buf[q])) ++q;
        bound[c] = q;
    }
    std::vector<uint64_t> ntok(C + 1, 0);
    std::vector<int> cerr(C, kTokOk);
    parallel_for(C, 1, [&](size_t c0, size_t c1) {
        for (size_t c = c0; c < c1; ++c)
            ntok[c + 1] = scan_tokens(buf + bound[c], buf + bound[c + 1], cerr[c], [](uint64_t, uint32_t) {});
    });
    for (size_t c = 0; c < C; ++c) {
        if (cerr[c] == kTokNonNumeric) return fail(KT_ERR_IO, std::string("non-numeric token in ") + path);
        if (cerr[c] == kTokRange) return fail(KT_ERR_IO, "id out of range");
    }
    for (size_t c = 0; c < C; ++c) ntok[c + 1] += ntok[c];
    if (ntok[C] % 3) return fail(KT_ERR_IO, std::string("truncated triple in ") + path);
    const uint64_t cnt = ntok[C] / 3;
    const size_t bytes = (cnt ? cnt : 1) * sizeof(uint32_t);
    uint32_t* s = (uint32_t*)std::malloc(bytes);
    uint32_t* d = (uint32_t*)std::malloc(bytes);
    uint32_t* w = third ? (uint32_t*)std::malloc(bytes) : nullptr;
    auto drop = [&] {
        std::free(s);
        std::free(d);
        std::free(w);
    };
    if (!s || !d || (third && !w)) {
        drop();
        return fail(KT_ERR_NOMEM, "tsv arrays");
    }
    // "dst src w": column 0 is the destination, column 1 the source
    parallel_for(C, 1, [&](size_t c0, size_t c1) {
        for (size_t c = c0; c < c1; ++c) {
            const uint64_t base = ntok[c];
            int e = kTokOk;
            scan_tokens(buf + bound[c], buf + bound[c + 1], e, [&](uint64_t t, uint32_t v) {
                const uint64_t g = base + t, line = g / 3;
                const unsigned col = (unsigned)(g % 3);
                if (col == 0) d[line] = v;
                else if (col == 1) s[line] = v;
                else if (w) w[line] = v;
            });
        }
    });
    // per line: self-loop, then 1-based ids; the lowest bad line wins
    const size_t T = std::max<size_t>(1, std::min<size_t>(num_threads(), (cnt + 65535) / 65536));
    std::vector<uint64_t> bad_at(T, ~0ull);
    std::vector<int> bad_code(T, 0);
    std::vector<uint32_t> tmax(T, 0);
    parallel_for(T, 1, [&](size_t t0, size_t t1) {
        for (size_t t = t0; t < t1; ++t) {
            const uint64_t b = cnt * t / T, e = cnt * (t + 1) / T;
            uint32_t mx = 0;
            for (uint64_t i = b; i < e; ++i) {
                const uint32_t a = s[i], c = d[i];
                if (a == c || a == 0 || c == 0) {
                    bad_at[t] = i;
                    bad_code[t] = a == c ? KT_ERR_SELF_LOOP : KT_ERR_RANGE;
                    break;
                }
                mx = std::max(mx, std::max(a, c));
                s[i] = a - 1;
                d[i] = c - 1;
            }
            tmax[t] = mx;
        }
    });
    for (size_t t = 0; t < T; ++t) {
        if (!bad_code[t]) continue;
        drop();
        return fail(bad_code[t], bad_code[t] == KT_ERR_SELF_LOOP ? "Edges from a vertex to itself are not allowed!"
                                                                 : "vertex ids are 1-based");
    }
    *n = *std::max_element(tmax.begin(), tmax.end());
    *m = cnt;
    *src = s;
    *dst = d;
    if (third) *third = w;
    return KT_OK;
}

inline char* put_u32(char* p, uint32_t v) {
    char t[10];
    int k = 0;
    do {
        t[k++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    while (k) *p++ = t[--k];
    return p;
}

// Lines "dst+1 \t src+1 \t w" (w = 1 without a third column), formatted by
// all threads in slices of kSlice lines and written in order.
int write_lines(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst, const uint32_t* third) {
    if (!path || (m && (!src || !dst))) return fail(KT_ERR_ARG, "NULL argument");
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(KT_ERR_IO, std::string("cannot open ") + path);
    constexpr uint64_t kSlice = 1 << 18;
    constexpr size_t kLineMax = 3 * 10 + 3;
    const size_t T = (size_t)std::max<uint64_t>(1, std::min<uint64_t>(num_threads(), (m + kSlice - 1) / kSlice));
    std::vector<std::vector<char>> out(T, std::vector<char>((size_t)std::min<uint64_t>(kSlice, m) * kLineMax));
    std::vector<size_t> len(T, 0);
    bool ok = true;
    for (uint64_t base = 0; base < m && ok; base += kSlice * T) {
        const size_t S = (size_t)std::min<uint64_t>(T, (m - base + kSlice - 1) / kSlice);
        parallel_for(S, 1, [&](size_t s0, size_t s1) {
            for (size_t sl = s0; sl < s1; ++sl) {
                const uint64_t b = base + sl * kSlice, e = std::min(m, b + kSlice);
                char* p = out[sl].data();
                for (uint64_t i = b; i < e; ++i) {
                    p = put_u32(p, dst[i] + 1);
                    *p++ = '\t';
                    p = put_u32(p, src[i] + 1);
                    *p++ = '\t';
                    p = put_u32(p, third ? third[i] : 1u);
                    *p++ = '\n';
                }
                len[sl] = (size_t)(p - out[sl].data());
            }
        });
        for (size_t sl = 0; sl < S && ok; ++sl) ok = std::fwrite(out[sl].data(), 1, len[sl], fp) == len[sl];
    }
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return fail(KT_ERR_IO, std::string("short write of ") + path);
    return KT_OK;
}

// ---------------------------------------------------------------------------
// Binary edge cache, little-endian:
//   bytes 0-7   "KTEDGES1"
//   bytes 8-11  n          (u32)
//   bytes 12-15 flags      (u32, KT_INPUT_SYMMETRIC or KT_INPUT_PAIRS)
//   bytes 16-23 m          (u64)
//   bytes 24-31 checksum   (u64, edge_checksum)
//   then src[m], dst[m]    (u32 each)
// The file size must be exactly 32 + 8 m.

constexpr char kBinMagic[8] = {'K', 'T', 'E', 'D', 'G', 'E', 'S', '1'};
constexpr size_t kBinHeader = 32;

inline uint64_t mix64(uint64_t z) {  // SplitMix64 finaliser (Steele et al., 2014)
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Sum over entries of mix64 of (entry index, src, dst): any changed, moved or
// swapped value changes a term; summed per thread, so it parallelises.
uint64_t edge_checksum(uint64_t m, const uint32_t* src, const uint32_t* dst) {
    const size_t T = std::max<size_t>(1, std::min<size_t>(num_threads(), (m + 65535) / 65536));
    std::vector<uint64_t> part(T, 0);
    parallel_for(T, 1, [&](size_t t0, size_t t1) {
        for (size_t t = t0; t < t1; ++t) {
            uint64_t h = 0;
            for (uint64_t i = m * t / T, e = m * (t + 1) / T; i < e; ++i)
                h += mix64(mix64(i + 0x9E3779B97F4A7C15ull) ^ (((uint64_t)src[i] << 32) | dst[i]));
            part[t] = h;
        }
    });
    uint64_t h = m;
    for (uint64_t p : part) h += p;
    return h;
}

void put_le(unsigned char* p, uint64_t v, int bytes) {
    for (int i = 0; i < bytes; ++i) p[i] = (unsigned char)(v >> (8 * i));
}
uint64_t get_le(const unsigned char* p, int bytes) {
    uint64_t v = 0;
    for (int i = 0; i < bytes; ++i) v |= (uint64_t)p[i] << (8 * i);
    return v;
}

bool host_little_endian() {
    const uint32_t one = 1;
    unsigned char b;
    std::memcpy(&b, &one, 1);
    return b == 1;
}

}  // namespace

int read_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst) {
    return parse_tsv(path, n, m, src, dst, nullptr);
}

int read_truss_tsv(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst, uint32_t** tau) {
    if (!tau) return fail(KT_ERR_ARG, "NULL argument");
    return parse_tsv(path, n, m, src, dst, tau);
}

int write_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst) {
    return write_lines(path, m, src, dst, nullptr);
}

int write_truss_tsv(const char* path, uint64_t m, const uint32_t* src, const uint32_t* dst, const uint32_t* tau) {
    if (m && !tau) return fail(KT_ERR_ARG, "NULL argument");
    return write_lines(path, m, src, dst, tau);
}

int write_edges_bin(const char* path, uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                    uint32_t flags) {
    if (!path || (m && (!src || !dst))) return fail(KT_ERR_ARG, "NULL argument");
    if (flags != KT_INPUT_SYMMETRIC && flags != KT_INPUT_PAIRS) return fail(KT_ERR_ARG, "unknown input flags");
    if (!host_little_endian()) return fail(KT_ERR_ARG, "binary edge files are little-endian");
    unsigned char h[kBinHeader];
    std::memcpy(h, kBinMagic, 8);
    put_le(h + 8, n, 4);
    put_le(h + 12, flags, 4);
    put_le(h + 16, m, 8);
    put_le(h + 24, edge_checksum(m, src, dst), 8);
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fail(KT_ERR_IO, std::string("cannot open ") + path);
    bool ok = std::fwrite(h, 1, kBinHeader, fp) == kBinHeader;
    if (m) {
        ok = ok && std::fwrite(src, 4, m, fp) == m;
        ok = ok && std::fwrite(dst, 4, m, fp) == m;
    }
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return fail(KT_ERR_IO, std::string("short write of ") + path);
    return KT_OK;
}

int read_edges_bin(const char* path, uint32_t* n, uint64_t* m, uint32_t** src, uint32_t** dst,
                   uint32_t* flags) {
    if (!path || !n || !m || !src || !dst) return fail(KT_ERR_ARG, "NULL argument");
    if (!host_little_endian()) return fail(KT_ERR_ARG, "binary edge files are little-endian");
    FileView f;
    if (int rc = f.open_file(path)) return rc;
    if (f.size < kBinHeader || std::memcmp(f.p, kBinMagic, 8) != 0)
        return fail(KT_ERR_IO, std::string("not a binary edge file: ") + path);
    const unsigned char* h = (const unsigned char*)f.p;
    const uint32_t hn = (uint32_t)get_le(h + 8, 4), hf = (uint32_t)get_le(h + 12, 4);
    const uint64_t hm = get_le(h + 16, 8), sum = get_le(h + 24, 8);
    if (hf != KT_INPUT_SYMMETRIC && hf != KT_INPUT_PAIRS)
        return fail(KT_ERR_IO, std::string("unknown input flags in ") + path);
    if (hm > (f.size - kBinHeader) / 8 || kBinHeader + 8 * hm != f.size)
        return fail(KT_ERR_IO, std::string("size does not match the entry count in ") + path);
    const size_t bytes = (hm ? hm : 1) * sizeof(uint32_t);
    uint32_t* s = (uint32_t*)std::malloc(bytes);
    uint32_t* d = (uint32_t*)std::malloc(bytes);
    if (!s || !d) {
        std::free(s);
        std::free(d);
        return fail(KT_ERR_NOMEM, "edge arrays");
    }
    // slices copied by all threads (the mapped pages fault in in parallel)
    parallel_for(hm, 1 << 20, [&](size_t b, size_t e) {
        std::memcpy(s + b, f.p + kBinHeader + b * 4, (e - b) * 4);
        std::memcpy(d + b, f.p + kBinHeader + (hm + b) * 4, (e - b) * 4);
    });
    if (edge_checksum(hm, s, d) != sum) {
        std::free(s);
        std::free(d);
        return fail(KT_ERR_IO, std::string("checksum mismatch in ") + path);
    }
    *n = hn;
    *m = hm;
    *src = s;
    *dst = d;
    if (flags) *flags = hf;
    return KT_OK;
}

int is_edges_bin(const char* path) {
    if (!path) return 0;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return 0;
    char b[8] = {};
    const bool yes = std::fread(b, 1, 8, fp) == 8 && std::memcmp(b, kBinMagic, 8) == 0;
    std::fclose(fp);
    return yes ? 1 : 0;
}

}  // namespace kt
