// bhx_sort.hip — stable LSD radix sort of (uint64 key, uint32 value) pairs.
//
// Replaces ThreePassRadixSortMortonKeys / RadixSortPass (MortonKeys.hpp:
// 373-470).  The reference runs nine stable byte passes at overlapping bit
// offsets, which amounts to a stable full sort by the 64-bit key; the same
// order comes out of eight stable passes over disjoint bytes.
//
// A pass whose digit is the same for every key (its histogram has one
// occupied bin) is a plain copy.  Each other pass is a single "onesweep"
// kernel (Adinets & Merrill 2022): a
// workgroup ranks a tile of 8192 keys in LDS, publishes its per-digit counts,
// and finds its global offsets by decoupled look-back (Merrill & Garland
// 2016) over the tiles before it.  Tile indices come from an atomic counter
// in launch order, so every tile a workgroup waits on is already running or
// done.  The spin on a predecessor's status is bounded; running out of spins
// raises an error flag instead of hanging.
#include "bhx_internal.hpp"

namespace bhx {
namespace {

constexpr int kBlock = 256;
constexpr int kPassBlock = 512;              // threads of a pass workgroup (8 waves)
constexpr int kItems = 16;
constexpr int kTile = kPassBlock * kItems;    // 8192 keys per tile
static_assert(kTile == (int)kSortTileKeys, "the tile the scratch is sized for");
constexpr int kPasses = 8;
constexpr int kBins = 256;
constexpr uint32_t kFlagA = 1u << 30;         // tile aggregate available
constexpr uint32_t kFlagP = 2u << 30;         // inclusive prefix available
constexpr uint32_t kCountMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;
constexpr int kLookWindow = 4;               // look-back statuses read at once

__device__ __forceinline__ void st_status(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_status(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one value per thread over a B-thread block.
template <int B = kBlock>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wsum, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < B / 64; ++w) {
        const uint32_t ws = s_wsum[w];
        if (w < wave) wbase += ws;
        tot += ws;
    }
    if (total) *total = tot;
    __syncthreads();
    return wbase + x - v;
}

__global__ void k_sort_init(uint32_t* __restrict__ hist, uint32_t* __restrict__ tile_ctr,
                            uint32_t* __restrict__ status, size_t nstatus) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nstatus; i += stride) status[i] = 0;
    if (blockIdx.x == 0) {
        for (int j = threadIdx.x; j < kPasses * kBins; j += blockDim.x) hist[j] = 0;
        if (threadIdx.x < kPasses) tile_ctr[threadIdx.x] = 0;
    }
}

// Digit histograms of all eight passes in one read of the keys.  With
// `ndev` the key count is read from device memory (at most `n`).
__global__ __launch_bounds__(kBlock) void k_sort_hist(const uint64_t* __restrict__ keys, size_t n,
                                                      uint32_t* __restrict__ hist, const uint32_t* __restrict__ ndev) {
    if (ndev) n = *ndev;
    __shared__ uint32_t h[kPasses * kBins];
    for (int j = threadIdx.x; j < kPasses * kBins; j += kBlock) h[j] = 0;
    __syncthreads();
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const uint64_t k = keys[i];
#pragma unroll
        for (int p = 0; p < kPasses; ++p) atomicAdd(&h[p * kBins + ((k >> (8 * p)) & 0xFF)], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kPasses * kBins; j += kBlock)
        if (h[j]) atomicAdd(&hist[j], h[j]);
}

// Exclusive scan of each pass's histogram: start of every digit in the output.
__global__ __launch_bounds__(kBlock) void k_sort_scan(uint32_t* __restrict__ hist) {
    __shared__ uint32_t s_wsum[kBlock / 64];
    for (int p = 0; p < kPasses; ++p) {
        const uint32_t v = hist[p * kBins + threadIdx.x];
        const uint32_t e = block_exclusive_scan(v, s_wsum, nullptr);
        hist[p * kBins + threadIdx.x] = e;
    }
}

__global__ __launch_bounds__(kPassBlock) void k_sort_pass(const uint64_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint64_t* __restrict__ kout,
                                                      uint32_t* __restrict__ vout, size_t n, int pass,
                                                      const uint32_t* __restrict__ goff,
                                                      uint32_t* __restrict__ status,
                                                      uint32_t* __restrict__ tile_ctr,
                                                      Counters* __restrict__ ctr,
                                                      const uint32_t* __restrict__ ndev) {
    if (ndev) n = *ndev;                       // the grid covers the largest count
    constexpr int kWaves = kPassBlock / 64;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wcnt[kWaves][kBins];
    __shared__ uint32_t s_tstart[kBins];
    __shared__ uint32_t s_gbase[kBins];
    __shared__ uint32_t s_wsum[kWaves];
    __shared__ uint64_t s_keys[kTile];
    __shared__ uint32_t s_vals[kTile];

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const bool dig = tid < kBins;              // threads 0 .. 255 also stand for one digit each
    // a digit every key shares (one occupied bin: its count is n) leaves the
    // order as it is; the pass is then a copy (keys in a small range, or
    // systems whose high key bytes agree)
    {
        bool one = false;
        if (dig) {
            const uint32_t g = goff[pass * kBins + tid];
            const uint32_t gn = tid + 1 < kBins ? goff[pass * kBins + tid + 1] : (uint32_t)n;
            one = gn - g == (uint32_t)n;
        }
        if (__syncthreads_count(one) != 0) {
            for (size_t i = blockIdx.x * (size_t)kPassBlock + tid; i < n; i += (size_t)gridDim.x * kPassBlock) {
                kout[i] = kin[i];
                vout[i] = vin[i];
            }
            return;
        }
    }
    if (tid == 0) s_tile = atomicAdd(tile_ctr + pass, 1u);
    for (int j = tid; j < kWaves * kBins; j += kPassBlock) (&s_wcnt[0][0])[j] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t base = (size_t)tile * kTile;
    // tiles past the device-side count hold no keys: every tile below them
    // took its index earlier and looks back only at lower tiles
    if (base >= n) return;
    const int shift = 8 * pass;

    // Warp-striped load: item j of a lane sits at base + wave*1024 + j*64 + lane,
    // so ranking items j = 0..15 in turn, lanes in order, follows index order.
    uint64_t key[kItems];
    uint32_t val[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const size_t idx = base + (size_t)wave * (64 * kItems) + (size_t)j * 64 + lane;
        if (idx < n) {
            key[j] = kin[idx];
            val[j] = vin[idx];
        } else {
            key[j] = ~0ull;
            val[j] = 0;
        }
    }
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const size_t idx = base + (size_t)wave * (64 * kItems) + (size_t)j * 64 + lane;
        const bool valid = idx < n;
        const uint32_t d = (uint32_t)(key[j] >> shift) & 0xFF;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t before = s_wcnt[wave][d];
        rank[j] = before + (uint32_t)__popcll(peers & lt_mask);
        if (valid && lane == __ffsll((unsigned long long)peers) - 1) s_wcnt[wave][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();

    // Per digit (thread = digit): counts of the waves, tile total.
    uint32_t cw[kWaves];
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        cw[w] = dig ? s_wcnt[w][tid] : 0u;
        tot += cw[w];
    }
    uint32_t* st = status + (size_t)tile * kBins;
    if (dig) st_status(st + tid, (tile == 0 ? kFlagP : kFlagA) | tot);
    const uint32_t tstart = block_exclusive_scan<kPassBlock>(tot, s_wsum, nullptr);

    uint32_t excl = 0;
    if (dig && tile > 0) {
        // kLookWindow predecessors' statuses in flight at once, read nearest first
        int64_t t = (int64_t)tile - 1;
        uint32_t spins = 0;
        for (;;) {
            uint32_t v[kLookWindow];
#pragma unroll
            for (int k = 0; k < kLookWindow; ++k)
                v[k] = t - k >= 0 ? ld_status(status + (size_t)(t - k) * kBins + tid) : kFlagP;
            bool done = false, stall = false;
#pragma unroll
            for (int k = 0; k < kLookWindow; ++k) {
                if (done || stall) continue;
                const uint32_t flag = v[k] & ~kCountMask;
                if (flag == 0) {          // not published yet: wait on this one
                    stall = true;
                    t -= k;
                    continue;
                }
                excl += v[k] & kCountMask;
                if (flag == kFlagP || t - k == 0) done = true;
            }
            if (done) break;
            if (stall) {
                if (++spins > kSpinLimit) {
                    ctr->sort_fail = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            t -= kLookWindow;
        }
        st_status(st + tid, kFlagP | (excl + tot));
    }
    if (dig) {
        s_tstart[tid] = tstart;
        s_gbase[tid] = goff[pass * kBins + tid] + excl;
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            s_wcnt[w][tid] = run;
            run += cw[w];
        }
    }
    __syncthreads();

    // Tile-local sorted order in LDS, then coalesced writes per digit run.
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const size_t idx = base + (size_t)wave * (64 * kItems) + (size_t)j * 64 + lane;
        if (idx < n) {
            const uint32_t d = (uint32_t)(key[j] >> shift) & 0xFF;
            const uint32_t q = s_tstart[d] + s_wcnt[wave][d] + rank[j];
            s_keys[q] = key[j];
            s_vals[q] = val[j];
        }
    }
    __syncthreads();
    const uint32_t nvalid = (uint32_t)((n - base) < (size_t)kTile ? (n - base) : (size_t)kTile);
    for (uint32_t q = tid; q < nvalid; q += kPassBlock) {
        const uint64_t k = s_keys[q];
        const uint32_t d = (uint32_t)(k >> shift) & 0xFF;
        const uint32_t out = s_gbase[d] + (q - s_tstart[d]);
        kout[out] = k;
        vout[out] = s_vals[q];
    }
}

}  // namespace

size_t sort_tiles_for(size_t n) { return (n + kTile - 1) / kTile; }

hipError_t radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, size_t n,
                            uint32_t* hist, uint32_t* status, uint32_t* tile_ctr, Counters* ctr, hipStream_t s,
                            int passes, const uint32_t* ndev) {
    if (passes < 1 || passes > kPasses) return hipErrorInvalidValue;
    // an even pass count leaves the result in keys/vals, an odd one in the alt buffers
    if (n == 0 || (n == 1 && !ndev && !(passes & 1))) return hipSuccess;
    const size_t tiles = sort_tiles_for(n);
    const size_t nstatus = (size_t)passes * tiles * kBins;
    unsigned ig = (unsigned)((nstatus + kBlock - 1) / kBlock);
    if (ig > 4096) ig = 4096;
    hipLaunchKernelGGL(k_sort_init, dim3(ig), dim3(kBlock), 0, s, hist, tile_ctr, status, nstatus);
    unsigned hg = (unsigned)((n + kBlock * 16 - 1) / (kBlock * 16));
    if (hg > 1024) hg = 1024;
    if (hg == 0) hg = 1;
    hipLaunchKernelGGL(k_sort_hist, dim3(hg), dim3(kBlock), 0, s, keys, n, hist, ndev);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(kBlock), 0, s, hist);
    for (int p = 0; p < passes; ++p) {
        const bool even = (p & 1) == 0;
        hipLaunchKernelGGL(k_sort_pass, dim3((unsigned)tiles), dim3(kPassBlock), 0, s, even ? keys : keys_alt,
                           even ? vals : vals_alt, even ? keys_alt : keys, even ? vals_alt : vals, n, p, hist,
                           status + (size_t)p * tiles * kBins, tile_ctr, ctr, ndev);
    }
    return hipGetLastError();
}

namespace {

// Second-stage keys of the segmented sort: the system of each sorted element
// (values are body positions) and the element's place in the key order.
__global__ __launch_bounds__(kBlock) void k_sys_keys(const uint32_t* __restrict__ vals,
                                                     const uint32_t* __restrict__ sysid, size_t n,
                                                     uint64_t* __restrict__ skey, uint32_t* __restrict__ sidx) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        skey[i] = sysid[vals[i]];
        sidx[i] = (uint32_t)i;
    }
}

__global__ __launch_bounds__(kBlock) void k_sys_gather(const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ vals,
                                                       const uint32_t* __restrict__ sidx, size_t n,
                                                       uint64_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const uint32_t j = sidx[i];
        kout[i] = keys[j];
        vout[i] = vals[j];
    }
}

__global__ __launch_bounds__(kBlock) void k_copy_pairs(const uint64_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin, size_t n,
                                                       uint64_t* __restrict__ kout, uint32_t* __restrict__ vout) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        kout[i] = kin[i];
        vout[i] = vin[i];
    }
}

unsigned grid_n(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    if (g == 0) g = 1;
    if (g > 65535u * 16u) g = 65535u * 16u;
    return (unsigned)g;
}

}  // namespace

// Several systems: a stable sort by key, then a stable sort by system (LSD
// order), gives every system's range sorted by key with ties in their prior
// order — the single-system order, system by system.
hipError_t sort_pairs(bhx_ctx* c, uint64_t* keys, uint32_t* vals, hipStream_t s, const SortScratch* sc) {
    SortScratch m;
    if (!sc) {
        m.keys_alt = c->keys_alt; m.vals_alt = c->vals_alt;
        m.hist = c->sort_hist; m.status = c->sort_status; m.tile_ctr = c->sort_tile_ctr;
        m.skey = c->skey; m.skey_alt = c->skey_alt; m.sidx = c->sidx; m.sidx_alt = c->sidx_alt;
        sc = &m;
    }
    const size_t n = c->n;
    hipError_t e = radix_sort_pairs(keys, vals, sc->keys_alt, sc->vals_alt, n, sc->hist, sc->status, sc->tile_ctr,
                                    c->ctr, s);
    if (e != hipSuccess || c->nsys <= 1 || n <= 1) return e;
    hipLaunchKernelGGL(k_sys_keys, dim3(grid_n(n)), dim3(kBlock), 0, s, vals, c->sysid, n, sc->skey, sc->sidx);
    e = radix_sort_pairs(sc->skey, sc->sidx, sc->skey_alt, sc->sidx_alt, n, sc->hist, sc->status, sc->tile_ctr, c->ctr,
                         s, 2);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sys_gather, dim3(grid_n(n)), dim3(kBlock), 0, s, keys, vals, sc->sidx, n, sc->keys_alt,
                       sc->vals_alt);
    hipLaunchKernelGGL(k_copy_pairs, dim3(grid_n(n)), dim3(kBlock), 0, s, sc->keys_alt, sc->vals_alt, n, keys, vals);
    return hipGetLastError();
}

}  // namespace bhx
