// bhx_noalloc.hip — the key sort of bhx_sort_keys on device arrays, with no
// allocation: for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The sort's working memory — the second key buffer, the carried values and
// their second buffer, a copy of the caller's indexes, the eight digit
// histograms, the tile counters, the look-back status of every tile and pass,
// and a status block — lives in one file-scope device array, sized for
// BHX_NOALLOC_MAX_KEYS keys.  Each GPU the library runs on holds its own copy
// of the array until the process ends; its device address is taken once per
// device with hipGetSymbolAddress (the array's name in host code is only a
// host-side stand-in).  Calls on one device share the array: they are
// serialised by a host mutex and each returns only when its sort is done.
#include <mutex>

#include "bhx_internal.hpp"

namespace bhx {
namespace {

constexpr size_t kMaxKeys = BHX_NOALLOC_MAX_KEYS;
constexpr size_t kPasses = 8, kBins = 256, kTileKeys = kSortTileKeys;
constexpr size_t kMaxTiles = (kMaxKeys + kTileKeys - 1) / kTileKeys;

// byte offsets of the regions (each a multiple of 256)
constexpr size_t kOffKeysAlt = 0;                                   // u64 x max
constexpr size_t kOffIdxCopy = kOffKeysAlt + 8 * kMaxKeys;          // u64 x max
constexpr size_t kOffVals = kOffIdxCopy + 8 * kMaxKeys;             // u32 x max
constexpr size_t kOffValsAlt = kOffVals + 4 * kMaxKeys;             // u32 x max
constexpr size_t kOffHist = kOffValsAlt + 4 * kMaxKeys;             // u32 x 8 x 256
constexpr size_t kOffTileCtr = kOffHist + 4 * kPasses * kBins;      // u32 x 8 (256 B)
constexpr size_t kOffCounters = kOffTileCtr + 256;                  // Counters (256 B)
constexpr size_t kOffStatus = kOffCounters + 256;                   // u32 x 8 x tiles x 256
constexpr size_t kScratchBytes = kOffStatus + 4 * kPasses * kMaxTiles * kBins;
static_assert(sizeof(Counters) <= 256, "status block fits its region");
static_assert(kOffHist % 256 == 0 && kOffStatus % 256 == 0, "regions stay 256-byte aligned");

static __device__ __align__(256) uint64_t scratch[kScratchBytes / 8];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};   // device address of `scratch` on each device (null: not taken yet)

constexpr int kBlock = 256;

unsigned grid_for(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    if (g == 0) g = 1;
    if (g > 65535u * 16u) g = 65535u * 16u;
    return (unsigned)g;
}

__global__ void k_clear_status(Counters* __restrict__ ctr) {
    if (threadIdx.x == 0) ctr->sort_fail = 0u;
}

__global__ __launch_bounds__(kBlock) void k_copy_idx(const uint64_t* __restrict__ a, size_t n,
                                                     uint64_t* __restrict__ b) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) b[i] = a[i];
}

// idx[i] = old[perm[i]]: the indexes follow their keys (ReorderBodiesByIndexes'
// new[i] = old[indexes[i]], Body.hpp:257-273)
__global__ __launch_bounds__(kBlock) void k_gather_idx(const uint64_t* __restrict__ old,
                                                       const uint32_t* __restrict__ perm, size_t n,
                                                       uint64_t* __restrict__ idx) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        idx[i] = old[perm[i]];
}

// The device address of this device's `scratch` (taken once per device).
int scratch_of(int device, unsigned char** out) {
    if (!g_scratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BHX_E_HIP;
        g_scratch[device] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch[device]);
    return BHX_OK;
}

}  // namespace
}  // namespace bhx

extern "C" int bhx_sort_keys_noalloc(int device, void* stream, uint64_t* keys, uint64_t* idx, size_t n) {
    using namespace bhx;
    if (!keys || !idx) return BHX_E_NULL;
    if (n > kMaxKeys) return BHX_E_CAPACITY;
    if (device < 0 || device >= kMaxDevices) return BHX_E_ARG;
    if (n <= 1) return BHX_OK;
    std::lock_guard<std::mutex> lock(g_mutex);
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return BHX_E_HIP;
    if (prev != device && hipSetDevice(device) != hipSuccess) return BHX_E_HIP;
    int rc = BHX_OK;
    unsigned char* base = nullptr;
    if ((rc = scratch_of(device, &base)) == BHX_OK) {
        hipStream_t s = static_cast<hipStream_t>(stream);
        uint64_t* keys_alt = reinterpret_cast<uint64_t*>(base + kOffKeysAlt);
        uint64_t* idx_copy = reinterpret_cast<uint64_t*>(base + kOffIdxCopy);
        uint32_t* vals = reinterpret_cast<uint32_t*>(base + kOffVals);
        uint32_t* vals_alt = reinterpret_cast<uint32_t*>(base + kOffValsAlt);
        uint32_t* hist = reinterpret_cast<uint32_t*>(base + kOffHist);
        uint32_t* tile_ctr = reinterpret_cast<uint32_t*>(base + kOffTileCtr);
        Counters* ctr = reinterpret_cast<Counters*>(base + kOffCounters);
        uint32_t* status = reinterpret_cast<uint32_t*>(base + kOffStatus);
        // the previous call's contents stay in the array: clear the failure flag
        // (the sort clears its histograms, tile counters and status itself)
        hipLaunchKernelGGL(k_clear_status, dim3(1), dim3(64), 0, s, ctr);
        hipError_t e = launch_iota(vals, n, s);
        if (e == hipSuccess)
            e = radix_sort_pairs(keys, vals, keys_alt, vals_alt, n, hist, status, tile_ctr, ctr, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_copy_idx, dim3(grid_for(n)), dim3(kBlock), 0, s, idx, n, idx_copy);
            hipLaunchKernelGGL(k_gather_idx, dim3(grid_for(n)), dim3(kBlock), 0, s, idx_copy, vals, n, idx);
            e = hipGetLastError();
        }
        uint32_t fail_flag = 0;
        if (e == hipSuccess)
            e = hipMemcpyAsync(&fail_flag, &ctr->sort_fail, sizeof fail_flag, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = BHX_E_HIP;
        else if (fail_flag) rc = BHX_E_SORT;
    }
    if (prev != device) (void)hipSetDevice(prev);
    return rc;
}
