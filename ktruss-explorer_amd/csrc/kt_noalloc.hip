// kt_noalloc.hip -- one k of the directed COO+CSR peel on a layout the caller
// holds in device memory, allocating nothing.
//
// kt_plan_ktruss runs the peel on a plan, whose layout and peel state
// (Info, info.h:10-19) were allocated when the plan was made.  Here the
// caller's own device arrays are the layout -- the COOCSRGraph of
// kernel_ddc.cu:116 (srcIdx, dstIdx, srcPtr) -- and the peel state lives in a
// fixed device array, `scratch`, of kNoAllocScratchBytes: the counters, the
// count kernel's statistics lines and one support per entry.  The kernels
// are the plan's own (support count, mark); they receive the array's device
// address, found once per device with hipGetSymbolAddress.
//
// Every call on one device shares that device's array and leaves its
// contents behind; each call clears what must start at zero.  Calls that
// overlap on one device (two host threads, or two streams) race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "kt_internal.h"

namespace kt {
namespace {

constexpr uint32_t kBlockNa = 256;
constexpr uint32_t kMaxDevicesNa = 64;

// Layout of the scratch array: counters, layout-check words, the count
// kernel's statistics lines, supports.
constexpr size_t kOffCtr = 0;
constexpr size_t kOffChk = 256;
constexpr size_t kOffStripes = 512;
constexpr size_t kOffSup = kOffStripes + kStripeBytes;
constexpr size_t kNoAllocScratchBytes = kOffSup + (size_t)KT_NOALLOC_MAX_ENTRIES * sizeof(uint32_t);
static_assert(sizeof(DevCounters) <= kOffChk - kOffCtr, "counters overflow their slot");
static_assert(kOffSup % 256 == 0, "supports start on a 256-byte boundary");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Layout-check words: [0] lowest offending entry, [1] lowest offending
// vertex (0xFFFFFFFF: none), [2] longest row.
enum { kChkEntry = 0, kChkVertex = 1, kChkMaxRow = 2, kChkWords = 4 };

// The layout the count kernel assumes: ptr[0] = 0, ptr[n] = E, rows
// non-decreasing; every entry e in the row of src[e], src[e] < dst[e] < n,
// destinations strictly ascending within a row.  Every read is in bounds
// whatever the arrays hold.
__global__ void k_check_layout(uint32_t n, uint32_t E, const uint32_t* __restrict__ ptr,
                               const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                               uint32_t* __restrict__ chk) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t longest = 0;
    for (uint32_t v = t0; v <= n; v += stride) {
        const uint32_t p = ptr[v];
        bool bad = (v == 0 && p != 0) || (v == n && p != E);
        if (v < n) {
            const uint32_t q = ptr[v + 1];
            bad = bad || q < p;
            if (q >= p && q - p > longest) longest = q - p;
        }
        if (bad) atomicMin(chk + kChkVertex, v);
    }
    if (longest) atomicMax(chk + kChkMaxRow, longest);
    for (uint32_t e = t0; e < E; e += stride) {
        const uint32_t u = src[e], x = dst[e];
        bool bad = u >= n || x >= n || x <= u;
        if (!bad) {
            const uint32_t b = ptr[u], f = ptr[u + 1];
            bad = e < b || e >= f || (e > b && dst[e - 1] >= x);
        }
        if (bad) atomicMin(chk + kChkEntry, e);
    }
}

std::mutex g_mu;
void* g_scratch[kMaxDevicesNa] = {};  // scratch's device address per device (null: not looked up)

// The device address of the current device's scratch array, looked up on the
// first call on that device (with it current).
int scratch_of_current(unsigned char** out, int* dev_out) {
    int dev = 0;
    KT_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= (int)kMaxDevicesNa) return fail(KT_ERR_NO_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        hipDeviceProp_t prop;
        KT_HIP_TRY(hipGetDeviceProperties(&prop, dev));
        if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
            return fail(KT_ERR_NO_DEVICE, std::string("current device is ") + prop.gcnArchName + ", not gfx950");
        void* p = nullptr;
        KT_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(KT_ERR_HIP, "hipGetSymbolAddress gave no address for the scratch array");
        g_scratch[dev] = p;
    }
    *out = (unsigned char*)g_scratch[dev];
    *dev_out = dev;
    return KT_OK;
}

uint32_t pow2_at_least(uint32_t x) {
    uint32_t p = 1;
    while (p < x && p < (1u << 31)) p <<= 1;
    return p;
}

}  // namespace

int ktruss_noalloc(uint32_t n, uint32_t E, const uint32_t* ptr, const uint32_t* src, uint32_t* dst,
                   uint32_t k, hipStream_t s, uint64_t* out_live) {
    if (k < 2) return fail(KT_ERR_ARG, "k must be at least 2");
    if (E > KT_NOALLOC_MAX_ENTRIES)
        return fail(KT_ERR_CAPACITY, "kt_ktruss_noalloc takes at most " + std::to_string(KT_NOALLOC_MAX_ENTRIES) +
                                         " entries (its fixed scratch holds one support per entry); got " +
                                         std::to_string(E));
    if (n >= KT_DEAD) return fail(KT_ERR_RANGE, "n must be below 2^31");
    if (!ptr || (E && (!src || !dst))) return fail(KT_ERR_ARG, "null layout array");
    unsigned char* base = nullptr;
    int dev = 0;
    int rc = scratch_of_current(&base, &dev);
    if (rc) return rc;
    DevCounters* ctr = (DevCounters*)(base + kOffCtr);
    uint32_t* chk = (uint32_t*)(base + kOffChk);
    unsigned long long* stripes = (unsigned long long*)(base + kOffStripes);
    uint32_t* sup = (uint32_t*)(base + kOffSup);

    // the previous call's contents stay behind: clear counters and statistics
    KT_HIP_TRY(dev::fill_u32((uint32_t*)ctr, 0, sizeof(DevCounters) / 4, s));
    KT_HIP_TRY(dev::fill_u32(chk, 0xFFFFFFFFu, 2, s));
    KT_HIP_TRY(dev::fill_u32(chk + kChkMaxRow, 0, kChkWords - kChkMaxRow, s));
    KT_HIP_TRY(dev::fill_u32((uint32_t*)stripes, 0, kStripeBytes / 4, s));
    const uint32_t work = std::max<uint32_t>(n + 1, E);
    const uint32_t grid = std::min<uint32_t>(1024, std::max<uint32_t>(1, (work + kBlockNa - 1) / kBlockNa));
    hipLaunchKernelGGL(k_check_layout, dim3(grid), dim3(kBlockNa), 0, s, n, E, ptr, src, dst, chk);
    KT_HIP_TRY(hipGetLastError());
    uint32_t h_chk[kChkWords];
    KT_HIP_TRY(hipMemcpyAsync(h_chk, chk, sizeof(h_chk), hipMemcpyDeviceToHost, s));
    KT_HIP_TRY(hipStreamSynchronize(s));
    if (h_chk[kChkVertex] != 0xFFFFFFFFu)
        return fail(KT_ERR_ARG, "row pointers: ptr[0] must be 0, ptr[n] = E and rows non-decreasing; first bad vertex " +
                                    std::to_string(h_chk[kChkVertex]));
    if (h_chk[kChkEntry] != 0xFFFFFFFFu)
        return fail(KT_ERR_ARG, "entry " + std::to_string(h_chk[kChkEntry]) +
                                    " breaks the oriented layout (src < dst < n, in the row of src, "
                                    "destinations strictly ascending per row)");

    // the count's LDS slice per wave, as a plan resolves it
    CountArgs a{};
    a.ptr = ptr;
    a.src = src;
    a.dst = dst;
    a.sup = sup;
    a.wl = nullptr;
    a.nitems = E;
    a.chunk = 128;
    a.cap_lds = std::min<uint32_t>(4096, std::max<uint32_t>(64, pow2_at_least(h_chk[kChkMaxRow])));
    const uint32_t block = 512;
    while (a.cap_lds > 64 && dev::count_lds_bytes(a.cap_lds, block, true, false, KT_COUNT_ROW) > 65536)
        a.cap_lds >>= 1;
    a.cap = k - 2;
    a.tile_size = 1;
    a.ytiles = 1;
    a.query_factor = 8;
    a.prefilter = 0;
    a.ctr = ctr;
    a.stripes = stripes;

    // the peel loop of kernel_ddc.cu:124-204: count, mark, stop when nothing
    // was deleted (or after the first pass at k = 3, kernel_ddc.cu:183-186)
    uint64_t live = E;
    while (live > 0) {
        KT_HIP_TRY(dev::fill_u32(&ctr->ndel, 0, 1, s));
        KT_HIP_TRY(dev::zero_live_supports(dst, sup, E, s));
        KT_HIP_TRY(dev::count_supports(a, true, false, KT_COUNT_ROW, block, s));
        KT_HIP_TRY(dev::mark_deleted(dst, src, nullptr, sup, nullptr, nullptr, nullptr, nullptr, E, k, false,
                                     ctr, s));
        unsigned int ndel = 0;
        KT_HIP_TRY(hipMemcpyAsync(&ndel, &ctr->ndel, sizeof(ndel), hipMemcpyDeviceToHost, s));
        KT_HIP_TRY(hipStreamSynchronize(s));
        if (ndel > live) return fail(KT_ERR_STATE, "mark pass deleted more entries than were live");
        live -= ndel;
        if (!ndel || k == 3) break;
    }
    if (out_live) *out_live = live;
    return KT_OK;
}

}  // namespace kt
