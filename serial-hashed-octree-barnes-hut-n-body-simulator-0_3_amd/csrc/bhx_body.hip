// bhx_body.hip — per-body kernels whose FP64 results must match the reference
// bit for bit: bounds, Morton keys, tree path keys, gathers and the leapfrog.
// Compiled with -ffp-contract=off: every operation rounds as written.
#include <climits>

#include "bhx_internal.hpp"

namespace bhx {
namespace {

constexpr int kBlock = 256;
// Chunk of the per-system sequential mean (doubles per axis and buffer; each
// row is padded by 16 doubles so that the chain's read-ahead never needs a
// bounds test): 2 x 3 x (1024 + 16) doubles = 48.75 KiB of LDS, which finds
// room beside other waves.
constexpr int kMeanChunkShared = 1024;
// The one-system mean in parallel: tiles of 1024 bodies, at most 64 pieces per
// tile and axis.
constexpr int kSumTile = 1024;
constexpr int kSumPieces = 64;
constexpr int32_t kCodeSingle = INT32_MIN;   // an element that stands alone
constexpr int16_t kPieceSingle = -32768;     // SumPiece::e of one plain addition
constexpr int16_t kPieceTile = -32767;       // SumPiece::e of a tile added element by element

inline unsigned grid_for(size_t n, int per_block = kBlock, unsigned cap = 65535u * 16u) {
    size_t g = (n + per_block - 1) / per_block;
    if (g == 0) g = 1;
    return (unsigned)(g < cap ? g : cap);
}

// ---------------------------------------------------------------- bounds

// OctantBounds(Body*, N) centre — OctantUtils.cpp:31-38 with Vec3D::operator/=
// (MathUtils.cpp:81-90): a left-to-right FP64 sum in array order, then a
// multiply by fl(1/N).  The sum is one dependent chain per axis; waves 1..3
// stream chunks of positions into LDS (double buffered) while lanes 0..2 of
// wave 0 run the x, y and z chains over the staged chunk.  One workgroup per
// system; the chains of different systems run side by side.  gate (one
// system): runs only when gate[3] is set (see k_sum_offsets).
template <int kMeanChunk>
__global__ __launch_bounds__(kBlock) void k_bounds_mean(const double* __restrict__ px,
                                                        const double* __restrict__ py,
                                                        const double* __restrict__ pz, SysView sv,
                                                        double* __restrict__ bounds4,
                                                        const uint32_t* __restrict__ gate) {
    if (gate && gate[3] == 0u) return;
    constexpr int kMeanRow = kMeanChunk + 16;
    __shared__ double buf[2][3][kMeanRow];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const uint32_t sys = blockIdx.x;
    const size_t first = sv.off[sys];
    const size_t n = sv.off[sys + 1] - first;
    double* out4 = bounds4 + 4 * (size_t)sys;
    const size_t nchunks = (n + kMeanChunk - 1) / kMeanChunk;
    const double* const src[3] = {px + first, py + first, pz + first};
    const bool pairs = (first & 1) == 0;   // 16-byte aligned double pairs

    double s = 0.0;
    for (size_t ch = 0; ch <= nchunks; ++ch) {
        const int b = (int)(ch & 1);
        if (wave != 0) {
            // waves 1..3 stage chunk ch (axis = wave - 1) while wave 0 sums chunk ch-1
            if (ch < nchunks) {
                const int a = wave - 1;
                const size_t base = ch * kMeanChunk;
                const double* p = src[a] + base;
                double* d = buf[b][a];
                if (pairs && base + kMeanChunk <= n) {
                    for (int j = 2 * lane; j < kMeanChunk; j += 128) {
                        const double2 v = *reinterpret_cast<const double2*>(p + j);
                        d[j] = v.x;
                        d[j + 1] = v.y;
                    }
                } else {
                    for (int j = lane; j < kMeanChunk; j += 64) d[j] = (base + j < n) ? p[j] : 0.0;
                }
            }
        } else if (ch > 0 && lane < 3) {
            // the chain's wave issues ahead of other waves on its SIMD (walk waves
            // of the same frame run beside it)
            if (ch == 1) __builtin_amdgcn_s_setprio(3);
            // the dependent chain: register double buffering keeps the LDS reads
            // of the next 16 values in flight while the current 16 are added
            const size_t base = (ch - 1) * kMeanChunk;
            const int cnt = (int)((n - base) < (size_t)kMeanChunk ? (n - base) : (size_t)kMeanChunk);
            const double* v = buf[b ^ 1][lane];
            int j = 0;
            if (cnt == kMeanChunk) {
                // the last read-ahead lands in the row's padding and is unused;
                // scheduling barriers keep each read-ahead ahead of the adds
                // that precede its use (the scheduler would otherwise sink it)
                double t0[16], t1[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) t0[u] = v[u];
                for (j = 0; j < kMeanChunk; j += 32) {
#pragma unroll
                    for (int u = 0; u < 16; ++u) t1[u] = v[j + 16 + u];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 16; ++u) s = s + t0[u];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 16; ++u) t0[u] = v[j + 32 + u];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int u = 0; u < 16; ++u) s = s + t1[u];
                    __builtin_amdgcn_sched_barrier(0);
                }
            } else {
                for (; j < cnt; ++j) s = s + v[j];
            }
        }
        __syncthreads();
    }
    if (wave == 0 && lane < 3) {
        const double dn = (double)n;
        if (dn != 0.0) {
            const double inv = 1.0 / dn;
            s = s * inv;
        }
        out4[lane] = s;
    }
    if (tid == 0) out4[3] = 0.0;
}

// ---- the one-system mean: the same sequential sum, computed in parallel
//
// The sum s_k = fl(s_{k-1} + x_k) (s_0 = 0) is one dependent chain.  Within a
// binade the rounding does not depend on where s is: let u = 2^(e-52) be the
// spacing of the doubles in [2^e, 2^(e+1)), s_{k-1} a multiple of u, and
// r_k = x_k/u rounded to the nearest integer, with x_k/u not halfway between
// two integers.  Then t = s_{k-1} + u*r_k is within u/2 of the exact sum, and
// when 2^e + u <= |t| <= 2^(e+1) - u no other double is that close, so
// s_k = t.  A run of additions whose partial sums stay in one binade is
// therefore s_end = s_start + u * (r_a + ... + r_b): an integer sum, which
// tiles compute in parallel.  The chain that remains walks the runs in order
// and, per run, checks the condition exactly from s_start and the least and
// greatest prefix of the r's; a run that fails the check (or that holds a
// halfway case, a non-finite value or a too-large increment) is added one
// element at a time.  The result equals the sequential sum bit for bit,
// whatever the estimate below; the estimate only decides where runs break.
//
// Estimated partial sums (any order of additions) choose each element's
// binade; elements whose estimate is near a binade edge stand alone.

__device__ __forceinline__ double readlane_d(double v, uint32_t j) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, (int)j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), (int)j);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Exclusive prefix over the 256 threads of a block (all threads call it);
// `total` gets the block's sum.  s_w: 4 slots of shared memory.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* s_w, T& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const T t = __shfl_up(inc, o, 64);
        if (lane >= o) inc = inc + t;
    }
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    T off = 0;
    total = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        const T x = s_w[w];
        if (w < wave) off = off + x;
        total = total + x;
    }
    __syncthreads();
    return off + (inc - v);
}

// Tile sums of the three axes (estimate: any order).
__global__ __launch_bounds__(kBlock) void k_sum_tiles(const double* __restrict__ px, const double* __restrict__ py,
                                                      const double* __restrict__ pz, size_t n,
                                                      double* __restrict__ tsum, uint32_t* __restrict__ gate) {
    if (blockIdx.x == 0 && threadIdx.x < 4) gate[threadIdx.x] = 0u;
    const size_t base = (size_t)blockIdx.x * kSumTile;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int i = threadIdx.x; i < kSumTile; i += kBlock) {
        const size_t k = base + i;
        if (k < n) {
            a0 = a0 + px[k];
            a1 = a1 + py[k];
            a2 = a2 + pz[k];
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        a0 = a0 + __shfl_down(a0, o, 64);
        a1 = a1 + __shfl_down(a1, o, 64);
        a2 = a2 + __shfl_down(a2, o, 64);
    }
    __shared__ double s_w[3][kBlock / 64];
    if (lane == 0) {
        s_w[0][wave] = a0;
        s_w[1][wave] = a1;
        s_w[2][wave] = a2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const double* w = s_w[threadIdx.x];
        tsum[3 * (size_t)blockIdx.x + threadIdx.x] = (w[0] + w[1]) + (w[2] + w[3]);
    }
}

// Exclusive prefix of the tile sums, per axis (one block).
__global__ __launch_bounds__(kBlock) void k_sum_scan(const double* __restrict__ tsum, uint32_t ntiles,
                                                     double* __restrict__ tpre) {
    __shared__ double s_w[kBlock / 64];
    for (int a = 0; a < 3; ++a) {
        double carry = 0.0;
        for (uint32_t t0 = 0; t0 < ntiles; t0 += kBlock) {
            const uint32_t t = t0 + threadIdx.x;
            const double v = t < ntiles ? tsum[3 * (size_t)t + a] : 0.0;
            double total;
            const double ex = block_excl_scan(v, s_w, total);
            if (t < ntiles) tpre[3 * (size_t)t + a] = carry + ex;
            carry = carry + total;
        }
    }
}

// Pieces of tile blockIdx.x, axis blockIdx.y: thread j holds elements 4j..4j+3.
__global__ __launch_bounds__(kBlock) void k_sum_pieces(const double* __restrict__ px, const double* __restrict__ py,
                                                       const double* __restrict__ pz, size_t n,
                                                       const double* __restrict__ tpre, uint32_t* __restrict__ cnt,
                                                       SumPiece* __restrict__ pieces, uint32_t* __restrict__ gate) {
    const uint32_t t = blockIdx.x, a = blockIdx.y;
    const double* __restrict__ x = a == 0 ? px : (a == 1 ? py : pz);
    const size_t base = (size_t)t * kSumTile;
    const int count = (int)((n - base) < (size_t)kSumTile ? (n - base) : (size_t)kSumTile);
    const int tid = threadIdx.x, lane = tid & 63;
    __shared__ int32_t s_code[kSumTile];
    __shared__ double s_wd[kBlock / 64];
    __shared__ int s_wi[kBlock / 64];
    __shared__ int s_pstart[kSumPieces];
    __shared__ int32_t s_pcode[kSumPieces];
    __shared__ double s_pbase[kSumPieces], s_plast[kSumPieces], s_pval[kSumPieces];
    __shared__ long long s_pmin[kSumPieces], s_pmax[kSumPieces];

    double v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * tid + i;
        v[i] = k < count ? x[base + k] : 0.0;
    }
    // estimated partial sums -> binade, increment and standing of each element
    double loc[4], run = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        run = run + v[i];
        loc[i] = run;
    }
    double tot;
    const double p0 = tpre[3 * (size_t)t + a] + block_excl_scan(run, s_wd, tot);
    int32_t code[4];
    double r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double ap = fabs(p0 + loc[i]);
        bool single = !(ap >= 0x1p-960 && ap < 0x1p1000);
        int e = 0;
        double rv = 0.0;
        if (!single) {
            int ex;
            const double fm = frexp(ap, &ex);
            e = ex - 1;
            if (fm < 0.5 + 0x1p-25 || fm > 1.0 - 0x1p-24) single = true;
            const double q = ldexp(v[i], 52 - e);
            rv = rint(q);
            if (!(fabs(rv) < 0x1p42) || fabs(q - rv) == 0.5) single = true;
        }
        code[i] = single ? kCodeSingle : e;
        r[i] = single ? 0.0 : rv;
        const int k = 4 * tid + i;
        if (k < count) s_code[k] = code[i];
    }
    if (tid < kSumPieces) {
        s_pmin[tid] = LLONG_MAX;
        s_pmax[tid] = LLONG_MIN;
    }
    __syncthreads();
    // piece boundaries: a change of binade, and around every single element
    bool st[4], last[4];
    int nst = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * tid + i;
        st[i] = false;
        last[i] = false;
        if (k < count) {
            const int32_t prev = k > 0 ? s_code[k - 1] : 0;
            st[i] = k == 0 || code[i] == kCodeSingle || prev == kCodeSingle || code[i] != prev;
            const int32_t next = k + 1 < count ? s_code[k + 1] : 0;
            last[i] = k + 1 == count || next == kCodeSingle || code[i] == kCodeSingle || next != code[i];
        }
        nst += st[i] ? 1 : 0;
    }
    int np;
    const int pex = block_excl_scan(nst, s_wi, np);
    if (np > kSumPieces) {   // (np is the same in every thread): one piece, element by element
        if (tid == 0) {
            SumPiece q;
            q.sum = q.lo = q.hi = 0.0;
            q.e = kPieceTile;
            q.len = (uint16_t)count;
            q.start = (uint32_t)base;
            pieces[(3 * (size_t)t + a) * kSumPieces] = q;
            cnt[3 * (size_t)t + a] = 1u;
            atomicAdd(&gate[a], 1u);
        }
        return;
    }
    // exact prefix of the increments over the tile (integers below 2^52)
    double rl[4], rr = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        rr = rr + r[i];
        rl[i] = rr;
    }
    double rtot;
    const double rex = block_excl_scan(rr, s_wd, rtot);
    int pid = pex - 1;
    // least and greatest prefix per run piece: this thread's share first
    int cur = -1;
    long long mn = LLONG_MAX, mx = LLONG_MIN;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = 4 * tid + i;
        if (k >= count) continue;
        const double R = rex + rl[i];
        if (st[i]) {
            ++pid;
            s_pstart[pid] = k;
            s_pcode[pid] = code[i];
            s_pbase[pid] = R - r[i];
            s_pval[pid] = v[i];
        }
        if (last[i]) s_plast[pid] = R;
        if (code[i] == kCodeSingle) continue;
        if (pid != cur) {
            if (cur >= 0) {
                atomicMin(&s_pmin[cur], mn);
                atomicMax(&s_pmax[cur], mx);
            }
            cur = pid;
            mn = LLONG_MAX;
            mx = LLONG_MIN;
        }
        const long long Ri = (long long)R;
        mn = Ri < mn ? Ri : mn;
        mx = Ri > mx ? Ri : mx;
    }
    // a wave whose run elements all lie in one piece reduces before its one atomic
    const int c0 = __shfl(cur, 0, 64);
    const bool same = (cur == c0 || cur < 0) && c0 >= 0;
    if (__all(same)) {
        for (int o = 32; o > 0; o >>= 1) {
            const long long omn = __shfl_xor(mn, o, 64), omx = __shfl_xor(mx, o, 64);
            mn = omn < mn ? omn : mn;
            mx = omx > mx ? omx : mx;
        }
        if (lane == 0) {
            atomicMin(&s_pmin[c0], mn);
            atomicMax(&s_pmax[c0], mx);
        }
    } else if (cur >= 0) {
        atomicMin(&s_pmin[cur], mn);
        atomicMax(&s_pmax[cur], mx);
    }
    __syncthreads();
    if (tid < np) {
        SumPiece q;
        const int start = s_pstart[tid];
        const int end = tid + 1 < np ? s_pstart[tid + 1] : count;
        const int32_t code = s_pcode[tid];
        q.e = code == kCodeSingle ? kPieceSingle : (int16_t)code;
        q.len = (uint16_t)(end - start);
        q.start = (uint32_t)(base + start);
        if (code == kCodeSingle) {
            q.sum = s_pval[tid];
            q.lo = 0.0;
            q.hi = 0.0;
        } else {
            const double b = s_pbase[tid];
            q.sum = s_plast[tid] - b;
            q.lo = (double)s_pmin[tid] - b;
            q.hi = (double)s_pmax[tid] - b;
        }
        pieces[(3 * (size_t)t + a) * kSumPieces + tid] = q;
    }
    if (tid == 0) cnt[3 * (size_t)t + a] = (uint32_t)np;
}

// Exclusive prefix of the piece counts per axis (one block); entry ntiles
// holds the axis's total.  When more than three quarters of some axis's tiles
// are added element by element (unsorted input: partial sums near zero), the
// plain chain (k_bounds_mean) is faster than the walk over pieces; gate[3]
// selects it.  Either gives the same bits.
__global__ __launch_bounds__(kBlock) void k_sum_offsets(const uint32_t* __restrict__ cnt, uint32_t ntiles,
                                                        uint32_t* __restrict__ off, uint32_t* __restrict__ gate) {
    if (threadIdx.x == 0) {
        uint32_t mx = gate[0] > gate[1] ? gate[0] : gate[1];
        mx = mx > gate[2] ? mx : gate[2];
        gate[3] = 4ull * mx > 3ull * ntiles ? 1u : 0u;
    }
    __shared__ uint32_t s_w[kBlock / 64];
    for (int a = 0; a < 3; ++a) {
        uint32_t carry = 0;
        for (uint32_t t0 = 0; t0 < ntiles; t0 += kBlock) {
            const uint32_t t = t0 + threadIdx.x;
            const uint32_t v = t < ntiles ? cnt[3 * (size_t)t + a] : 0u;
            uint32_t total;
            const uint32_t ex = block_excl_scan(v, s_w, total);
            if (t < ntiles) off[(size_t)a * (ntiles + 1) + t] = carry + ex;
            carry += total;
        }
        if (threadIdx.x == 0) off[(size_t)a * (ntiles + 1) + ntiles] = carry;
    }
}

// The pieces of every tile, one contiguous stream per axis.
__global__ __launch_bounds__(64) void k_sum_compact(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off,
                                                    uint32_t ntiles, const SumPiece* __restrict__ pieces,
                                                    SumPiece* __restrict__ out, size_t axis_stride) {
    const uint32_t t = blockIdx.x, a = blockIdx.y;
    const uint32_t c = cnt[3 * (size_t)t + a];
    if (threadIdx.x < c)
        out[(size_t)a * axis_stride + off[(size_t)a * (ntiles + 1) + t] + threadIdx.x] =
            pieces[(3 * (size_t)t + a) * kSumPieces + threadIdx.x];
}

// s + buf[0] + buf[1] + ... + buf[len-1], left to right, every lane alike.
// Register double buffering keeps the LDS reads of the next 16 values in
// flight while the current 16 are added; buf holds 32 doubles of padding for
// the last read-ahead.
__device__ __forceinline__ double add_from_lds(const double* buf, int len, double s) {
    int j = 0;
    const double2* b2 = reinterpret_cast<const double2*>(buf);   // 16-byte reads
    if (len >= 32) {
        double t0[16], t1[16];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double2 w = b2[u];
            t0[2 * u] = w.x;
            t0[2 * u + 1] = w.y;
        }
        for (; j + 32 <= len; j += 32) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const double2 w = b2[(j + 16) / 2 + u];
                t1[2 * u] = w.x;
                t1[2 * u + 1] = w.y;
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < 16; ++u) s = s + t0[u];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const double2 w = b2[(j + 32) / 2 + u];
                t0[2 * u] = w.x;
                t0[2 * u + 1] = w.y;
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < 16; ++u) s = s + t1[u];
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    for (; j < len; ++j) s = s + buf[j];
    return s;
}

// Elements x[0 .. len) of one chunk (1 <= len <= 1024), 16 per lane (lane +
// 64 k); entries past len repeat x[len-1] (never added).  The loads take no
// branch and nothing reads them before they are written to LDS, so all 16
// stay in flight while the previous chunk is added.
__device__ __forceinline__ void load_chunk(const double* __restrict__ x, int len, int lane, double (&v)[16]) {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int i = lane + 64 * k;
        v[k] = x[i < len ? i : len - 1];
    }
}

// One chunk into a 1024-entry LDS buffer.
__device__ __forceinline__ void chunk_to_lds(const double (&v)[16], int lane, double* buf) {
#pragma unroll
    for (int k = 0; k < 16; ++k) buf[lane + 64 * k] = v[k];
    // the wave's LDS writes complete before any lane reads them back
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// s + x[0] + x[1] + ... + x[len-1], left to right, every lane alike (len >= 1).
// Chunks of up to 1024 elements pass through two LDS buffers: the next chunk's
// loads go out before the current chunk's additions and are written to the
// other buffer after them.
__device__ __forceinline__ double add_range(const double* __restrict__ x, int len, double s, int lane,
                                            double* buf0, double* buf1) {
    int c0 = 0, cl = len < kSumTile ? len : kSumTile;
    double v[16];
    load_chunk(x, cl, lane, v);
    chunk_to_lds(v, lane, buf0);
    double* cur = buf0;
    double* nxt = buf1;
    for (;;) {
        const int n0 = c0 + cl;
        const int nl = (len - n0) < kSumTile ? (len - n0) : kSumTile;
        if (nl > 0) load_chunk(x + n0, nl, lane, v);
        s = add_from_lds(cur, cl, s);
        if (nl <= 0) break;
        chunk_to_lds(v, lane, nxt);
        double* t = cur;
        cur = nxt;
        nxt = t;
        c0 = n0;
        cl = nl;
    }
    return s;
}

// The chain over the pieces: workgroup a (one wave) sums axis a, so the three
// chains issue from different compute units.  Lane j holds piece g0 + j
// of the axis's stream; the next 64 are loaded while these are applied.
// Consecutive whole-tile pieces are consecutive tiles, added as one range.
__global__ __launch_bounds__(64) void k_sum_walk(const double* __restrict__ px, const double* __restrict__ py,
                                                  const double* __restrict__ pz, size_t n, uint32_t ntiles,
                                                  const uint32_t* __restrict__ off,
                                                  const SumPiece* __restrict__ stream, size_t axis_stride,
                                                  double* __restrict__ bounds4, const uint32_t* __restrict__ gate) {
    if (gate[3] != 0u) return;   // the plain chain runs instead
    const int a = (int)blockIdx.x;
    const int lane = threadIdx.x;
    const double* __restrict__ x = a == 0 ? px : (a == 1 ? py : pz);
    __shared__ __align__(16) double s_buf[2][kSumTile + 32];
    // the chain issues ahead of other kernels' waves on its SIMD
    __builtin_amdgcn_s_setprio(3);
    const uint32_t np = (uint32_t)__builtin_amdgcn_readfirstlane((int)off[(size_t)a * (ntiles + 1) + ntiles]);
    const SumPiece* __restrict__ ps = stream + (size_t)a * axis_stride;
    SumPiece q{}, nq{};
    if ((uint32_t)lane < np) q = ps[lane];
    if (64u + lane < np) nq = ps[64 + lane];
    double s = 0.0;
    for (uint32_t g0 = 0; g0 < np; g0 += 64) {
        const uint32_t m = (np - g0) < 64u ? (np - g0) : 64u;
        const unsigned long long tiles = __ballot((uint32_t)lane < m && q.e == kPieceTile);
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)(((uint32_t)(uint16_t)q.e) |
                                                                       ((uint32_t)q.len << 16)), (int)j);
            const int e = (int16_t)(w & 0xFFFFu);
            if (e == kPieceSingle) {
                s = s + readlane_d(q.sum, j);
                continue;
            }
            const uint32_t start = (uint32_t)__builtin_amdgcn_readlane((int)q.start, (int)j);
            if (e == kPieceTile) {
                // whole-tile pieces j .. j+k-1 (this block): one contiguous range
                const unsigned long long rest = ~(tiles >> j);
                uint32_t k = rest ? (uint32_t)__builtin_ctzll(rest) : 64u - j;
                k = k < m - j ? k : m - j;
                const uint32_t jl = j + k - 1;
                const uint32_t wl = (uint32_t)__builtin_amdgcn_readlane((int)q.len, (int)jl) & 0xFFFFu;
                const uint32_t sl = (uint32_t)__builtin_amdgcn_readlane((int)q.start, (int)jl);
                s = add_range(x + start, (int)(sl + wl - start), s, lane, s_buf[0], s_buf[1]);
                j = jl;
                continue;
            }
            // a run: one exact addition when its partial sums stay in the binade
            const double sum = readlane_d(q.sum, j), lo = readlane_d(q.lo, j), hi = readlane_d(q.hi, j);
            const double S = s * ldexp(1.0, 52 - e);   // exact: a power-of-two scaling
            const double A = S + lo, B = S + hi;
            const bool ok = S == rint(S) && ((A >= 0x1p52 + 1.0 && B <= 0x1p53 - 1.0) ||
                                             (B <= -(0x1p52 + 1.0) && A >= -(0x1p53 - 1.0)));
            if (ok)
                s = (S + sum) * ldexp(1.0, e - 52);
            else   // element by element
                s = add_range(x + start, (int)(w >> 16), s, lane, s_buf[0], s_buf[1]);
        }
        q = nq;
        if (g0 + 128u + lane < np) nq = ps[g0 + 128u + lane];
    }
    if (lane == 0) {
        const double dn = (double)n;
        if (dn != 0.0) {
            const double inv = 1.0 / dn;
            s = s * inv;
        }
        bounds4[a] = s;
        if (a == 0) bounds4[3] = 0.0;
    }
}

// OctantBounds size — OctantUtils.cpp:47-58: max over bodies and axes of
// fabs(p) - c, floored at 0.  A max is exact in any order, so a parallel
// reduction reproduces it; non-negative doubles order like their bit patterns.
// blockIdx.y is the system, blockIdx.x strides over its bodies.
__global__ __launch_bounds__(kBlock) void k_bounds_size(const double* __restrict__ px,
                                                        const double* __restrict__ py,
                                                        const double* __restrict__ pz, SysView sv,
                                                        double* __restrict__ bounds4) {
    const uint32_t sys = blockIdx.y;
    const size_t first = sv.off[sys], end = sv.off[sys + 1];
    double* out4 = bounds4 + 4 * (size_t)sys;
    if (first + (size_t)blockIdx.x * kBlock >= end) return;
    const double cx = out4[0], cy = out4[1], cz = out4[2];
    double s = 0.0;
    for (size_t i = first + blockIdx.x * (size_t)kBlock + threadIdx.x; i < end; i += (size_t)gridDim.x * kBlock) {
        double t;
        t = fabs(px[i]) - cx; if (t > s) s = t;
        t = fabs(py[i]) - cy; if (t > s) s = t;
        t = fabs(pz[i]) - cz; if (t > s) s = t;
    }
    __shared__ double red[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(s, off, 64);
        if (o > s) s = o;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w)
            if (red[w] > s) s = red[w];
        atomicMax(reinterpret_cast<unsigned long long*>(out4 + 3),
                  (unsigned long long)__double_as_longlong(s));
    }
}

// ---------------------------------------------------------------- keys

// (spatialKey)(xScaled) in MortonKeys.hpp:162-164 with x86-64 SSE2 semantics:
// truncate to int64 (0x8000000000000000 when out of range or NaN), then
// reinterpret as uint64.
__device__ __forceinline__ uint64_t cvt_u64_x86(double u) {
    if (!(u >= -9223372036854775808.0 && u < 9223372036854775808.0)) return 0x8000000000000000ull;
    return (uint64_t)(int64_t)u;
}

// interleaveBits' per-axis spread (MortonKeys.hpp:100-116) with the magic-bit
// masks of MortonKeys.hpp:51-58.
__device__ __forceinline__ uint64_t spread21(uint64_t t) {
    t = (t | (t << 32)) & 0x001f00000000ffffull;
    t = (t | (t << 16)) & 0x001f0000ff0000ffull;
    t = (t | (t << 8)) & 0x100f00f00f00f00full;
    t = (t | (t << 4)) & 0x10c30c30c30c30c3ull;
    t = (t | (t << 2)) & 0x1249249249249249ull;
    return t;
}

// computeMortonKey — MortonKeys.hpp:137-181; the sort payload starts as the
// identity (RadixSortBodies' indexes[i] = i, Body.hpp:285-289).
__global__ __launch_bounds__(kBlock) void k_morton(const double* __restrict__ px,
                                                   const double* __restrict__ py,
                                                   const double* __restrict__ pz, size_t n,
                                                   const uint32_t* __restrict__ sysid,
                                                   const double* __restrict__ bounds4,
                                                   uint64_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals,
                                                   uint32_t* __restrict__ nonfinite) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const double size = bounds4[4 * (size_t)(sysid ? sysid[i] : 0u) + 3];
        const double frac = 2097152.0 / (2.0 * size);
        const double x = px[i], y = py[i], z = pz[i];
        if (nonfinite && !(isfinite(x) && isfinite(y) && isfinite(z))) *nonfinite = 1u;
        double xs = x + size;
        double ys = y + size;
        double zs = z + size;
        xs = xs * frac;
        ys = ys * frac;
        zs = zs * frac;
        const uint64_t X = spread21(cvt_u64_x86(xs));
        const uint64_t Y = spread21(cvt_u64_x86(ys));
        const uint64_t Z = spread21(cvt_u64_x86(zs));
        keys[i] = (1ull << 63) | (X << 2) | (Y << 1) | Z;
        if (vals) vals[i] = (uint32_t)i;
    }
}

// Root edge: the bounds size (reference geometry) or twice it (enclosing
// cube); the factor is a power of two, so both recurrences are exact.
__device__ __forceinline__ uint64_t path_key(double x, double y, double z, const double* __restrict__ b4,
                                             double tscale) {
    double cx = b4[0], cy = b4[1], cz = b4[2], s = b4[3] * tscale;
    uint64_t k = 1;
#pragma unroll
    for (int L = 1; L <= kMaxLevel; ++L) {
        const unsigned ox = x > cx, oy = y > cy, oz = z > cz;
        k = (k << 3) | (ox << 2) | (oy << 1) | oz;
        s = s * 0.5;
        cx = cx + ((ox ? 1.0 : -1.0) * 0.5) * s;
        cy = cy + ((oy ? 1.0 : -1.0) * 0.5) * s;
        cz = cz + ((oz ? 1.0 : -1.0) * 0.5) * s;
    }
    return k;
}

// The descent a body makes through insertBody (HashedOctree.cpp:474-505):
// octant by strict > against the cell centre (DetermineOctant,
// OctantUtils.hpp:64-70), child key parent<<3|o (GetChildKey :97-101), child
// edge size*0.5 and centre c + (dir*0.5)*edge (HOTNode child constructor,
// HashedNode.cpp:50-56).  Written as one 64-bit path key over 21 levels.
__global__ __launch_bounds__(kBlock) void k_treekeys(const double* __restrict__ px,
                                                     const double* __restrict__ py,
                                                     const double* __restrict__ pz, size_t n,
                                                     const uint32_t* __restrict__ sysid,
                                                     const double* __restrict__ bounds4, double tscale,
                                                     uint64_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        keys[i] = path_key(px[i], py[i], pz[i], bounds4 + 4 * (size_t)(sysid ? sysid[i] : 0u), tscale);
        vals[i] = (uint32_t)i;
    }
}

__global__ void k_iota(uint32_t* __restrict__ v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        v[i] = (uint32_t)i;
}

// ---------------------------------------------------------------- gathers / leapfrog

struct GatherArgs {
    const double* src[16];
    double* dst[16];
    int narr;
};

// ReorderBodiesByIndexes — Body.hpp:257-273: new[i] = old[perm[i]].
__global__ __launch_bounds__(kBlock) void k_gather(GatherArgs g, const uint32_t* __restrict__ perm, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const uint32_t j = perm[i];
        for (int a = 0; a < g.narr; ++a) g.dst[a][i] = g.src[a][j];
    }
}

struct BodyArrays {
    const double *px, *py, *pz, *vx, *vy, *vz, *m, *ax, *ay, *az;
    const double *cpx, *cpy, *cpz, *cvx, *cvy, *cvz;
    double *opx, *opy, *opz, *ovx, *ovy, *ovz, *om;
    double *ocpx, *ocpy, *ocpz, *ocvx, *ocvy, *ocvz;
};

// Kick then drift — ComputePositionAtHalfTimeStep, Body.hpp:306-313:
// v = v + (a*dt)*0.5; x = x + v*dt.  The compensated form keeps the rounding
// error of each accumulation (Kahan) and feeds it into the next one.
__device__ __forceinline__ void kick_drift_one(double& x, double& v, double a, double dt) {
    v = v + (a * dt) * 0.5;
    x = x + v * dt;
}

// One body's kick then drift from its state at index j (acceleration at ja);
// the frame's x_{f+1} and v_{f+1/2}.
struct Drifted {
    double x, y, z, vx, vy, vz, cpx, cpy, cpz, cvx, cvy, cvz;
};
template <bool kComp>
__device__ __forceinline__ Drifted kick_drift_body(const BodyArrays& b, size_t j, size_t ja, double dt) {
    Drifted d;
    d.x = b.px[j]; d.y = b.py[j]; d.z = b.pz[j];
    d.vx = b.vx[j]; d.vy = b.vy[j]; d.vz = b.vz[j];
    const double ax = b.ax[ja], ay = b.ay[ja], az = b.az[ja];
    if (!kComp) {
        kick_drift_one(d.x, d.vx, ax, dt);
        kick_drift_one(d.y, d.vy, ay, dt);
        kick_drift_one(d.z, d.vz, az, dt);
        d.cpx = d.cpy = d.cpz = d.cvx = d.cvy = d.cvz = 0.0;
    } else {
        d.cvx = b.cvx[j]; d.cvy = b.cvy[j]; d.cvz = b.cvz[j];
        d.cpx = b.cpx[j]; d.cpy = b.cpy[j]; d.cpz = b.cpz[j];
        kahan_add(d.vx, d.cvx, (ax * dt) * 0.5);
        kahan_add(d.vy, d.cvy, (ay * dt) * 0.5);
        kahan_add(d.vz, d.cvz, (az * dt) * 0.5);
        kahan_add(d.x, d.cpx, d.vx * dt);
        kahan_add(d.y, d.cpy, d.vy * dt);
        kahan_add(d.z, d.cpz, d.vz * dt);
    }
    return d;
}

// Fused RadixSortBodies gather (Body.hpp:296) + kick/drift (Body.hpp:306-313)
// for the reference frame order: bodies are reordered by `perm` and moved in
// one pass.  With permute_acc = 0 the acceleration slot i is not reordered
// (the as-run behaviour, ofApp.cpp:176,248).  `inv` (may be null) receives
// the inverse permutation: inv[perm[i]] = i.
template <bool kComp>
__global__ __launch_bounds__(kBlock) void k_permute_kick_drift(BodyArrays b, const uint32_t* __restrict__ perm,
                                                               size_t n, double dt, int permute_acc,
                                                               uint32_t* __restrict__ inv) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const size_t j = perm ? perm[i] : i;
        const size_t ja = permute_acc ? j : i;
        const Drifted d = kick_drift_body<kComp>(b, j, ja, dt);
        if (kComp) {
            b.ocvx[i] = d.cvx; b.ocvy[i] = d.cvy; b.ocvz[i] = d.cvz;
            b.ocpx[i] = d.cpx; b.ocpy[i] = d.cpy; b.ocpz[i] = d.cpz;
        }
        b.opx[i] = d.x; b.opy[i] = d.y; b.opz[i] = d.z;
        b.ovx[i] = d.vx; b.ovy[i] = d.vy; b.ovz[i] = d.vz;
        b.om[i] = b.m[j];
        if (inv) inv[j] = (uint32_t)i;
    }
}

// The tree branch of a frame whose accelerations move with their bodies: the
// path key of every body's drifted position, in the order before the body
// sort (kick_drift_body gives bitwise the positions k_permute_kick_drift
// writes), with the identity payload.
template <bool kComp>
__global__ __launch_bounds__(kBlock) void k_drift_treekeys(BodyArrays b, size_t n, double dt,
                                                           const uint32_t* __restrict__ sysid,
                                                           const double* __restrict__ bounds4, double tscale,
                                                           uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    for (size_t j = blockIdx.x * (size_t)kBlock + threadIdx.x; j < n; j += (size_t)gridDim.x * kBlock) {
        const Drifted d = kick_drift_body<kComp>(b, j, j, dt);
        keys[j] = path_key(d.x, d.y, d.z, bounds4 + 4 * (size_t)(sysid ? sysid[j] : 0u), tscale);
        vals[j] = (uint32_t)j;
    }
}

// Tree order in the new array order from the tree branch's sort: slot t holds
// inv[told[t]].  The branch sorted stably from the order before the body
// sort, so bodies of equal path key sit in that order; the frame's tree order
// has them in the new array order, so each run of equal keys (rare: one
// finest cell) is put in ascending new index by the run's first slot.  Then
// the tree-order copies of positions and masses.
__global__ __launch_bounds__(kBlock) void k_tree_order(const uint64_t* __restrict__ tkeys,
                                                       const uint32_t* __restrict__ told,
                                                       const uint32_t* __restrict__ inv, size_t n,
                                                       const double* __restrict__ px, const double* __restrict__ py,
                                                       const double* __restrict__ pz, const double* __restrict__ m,
                                                       uint32_t* __restrict__ tperm, double* __restrict__ tx,
                                                       double* __restrict__ ty, double* __restrict__ tz,
                                                       double* __restrict__ tm) {
    for (size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x; t < n; t += (size_t)gridDim.x * kBlock) {
        const uint64_t k = tkeys[t];
        if (t > 0 && tkeys[t - 1] == k) continue;   // inside a run: its first slot places it
        size_t e = t + 1;
        while (e < n && tkeys[e] == k) ++e;
        if (e > t + 1) {
            // insertion sort of the run by new index (in place in tperm)
            for (size_t u = t; u < e; ++u) {
                const uint32_t v = inv[told[u]];
                size_t w = u;
                while (w > t && tperm[w - 1] > v) {
                    tperm[w] = tperm[w - 1];
                    --w;
                }
                tperm[w] = v;
            }
        } else {
            tperm[t] = inv[told[t]];
        }
        for (size_t u = t; u < e; ++u) {
            const uint32_t j = tperm[u];
            tx[u] = px[j];
            ty[u] = py[j];
            tz[u] = pz[j];
            tm[u] = m[j];
        }
    }
}

// Second kick — ComputeVelocityAndPosition, Body.hpp:317-324: v = v + (a*dt)*0.5.
template <bool kComp>
__global__ __launch_bounds__(kBlock) void k_kick(double* __restrict__ vx, double* __restrict__ vy,
                                                 double* __restrict__ vz, const double* __restrict__ ax,
                                                 const double* __restrict__ ay, const double* __restrict__ az,
                                                 double* __restrict__ cvx, double* __restrict__ cvy,
                                                 double* __restrict__ cvz, size_t n, double dt,
                                                 uint32_t* __restrict__ nonfinite) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        double v0 = vx[i], v1 = vy[i], v2 = vz[i];
        if (!kComp) {
            half_kick(v0, v1, v2, nullptr, ax[i], ay[i], az[i], dt);
        } else {
            double c[3] = {cvx[i], cvy[i], cvz[i]};
            half_kick(v0, v1, v2, c, ax[i], ay[i], az[i], dt);
            cvx[i] = c[0]; cvy[i] = c[1]; cvz[i] = c[2];
        }
        vx[i] = v0; vy[i] = v1; vz[i] = v2;
        if (nonfinite && !(isfinite(v0) && isfinite(v1) && isfinite(v2))) *nonfinite = 1u;
    }
}

// Tree-order copies of positions and masses for the walk.
__global__ __launch_bounds__(kBlock) void k_tree_gather(const double* __restrict__ px, const double* __restrict__ py,
                                                        const double* __restrict__ pz, const double* __restrict__ m,
                                                        const uint32_t* __restrict__ perm, size_t n,
                                                        double* __restrict__ tx, double* __restrict__ ty,
                                                        double* __restrict__ tz, double* __restrict__ tm) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const uint32_t j = perm[i];
        tx[i] = px[j];
        ty[i] = py[j];
        tz[i] = pz[j];
        tm[i] = m[j];
    }
}

}  // namespace

size_t sum_tiles_for(size_t n) { return (n + kSumTile - 1) / kSumTile + 1; }

hipError_t launch_bounds(bhx_ctx* c, double* bounds4, hipStream_t s) {
    const SysView sv = c->sys();
    if (sv.nsys == 0 || c->n == 0) return hipSuccess;
    if (sv.nsys == 1) {
        // one system: the chain over runs of additions (several systems run
        // their short chains side by side below)
        const uint32_t nt = (uint32_t)((c->n + kSumTile - 1) / kSumTile);
        if (nt + 1 > c->sum_tiles_cap) return hipErrorInvalidValue;
        double* tsum = c->sum_tile;
        double* tpre = c->sum_tile + 3 * c->sum_tiles_cap;
        uint32_t* gate = c->sum_cnt + 6 * c->sum_tiles_cap;   // 4 words
        hipLaunchKernelGGL(k_sum_tiles, dim3(nt), dim3(kBlock), 0, s, c->px, c->py, c->pz, c->n, tsum, gate);
        hipLaunchKernelGGL(k_sum_scan, dim3(1), dim3(kBlock), 0, s, tsum, nt, tpre);
        uint32_t* cnt = c->sum_cnt;
        uint32_t* off = c->sum_cnt + 3 * c->sum_tiles_cap;
        SumPiece* pieces = c->sum_piece;
        SumPiece* stream = c->sum_piece + 3 * (size_t)kSumPieces * c->sum_tiles_cap;
        const size_t axis_stride = (size_t)kSumPieces * c->sum_tiles_cap;
        hipLaunchKernelGGL(k_sum_pieces, dim3(nt, 3), dim3(kBlock), 0, s, c->px, c->py, c->pz, c->n, tpre, cnt,
                           pieces, gate);
        hipLaunchKernelGGL(k_sum_offsets, dim3(1), dim3(kBlock), 0, s, cnt, nt, off, gate);
        hipLaunchKernelGGL(k_sum_compact, dim3(nt, 3), dim3(64), 0, s, cnt, off, nt, pieces, stream, axis_stride);
        hipLaunchKernelGGL(k_sum_walk, dim3(3), dim3(64), 0, s, c->px, c->py, c->pz, c->n, nt, off, stream,
                           axis_stride, bounds4, gate);
        hipLaunchKernelGGL(k_bounds_mean<kMeanChunkShared>, dim3(1), dim3(kBlock), 0, s, c->px, c->py, c->pz, sv,
                           bounds4, gate);
    } else {
        hipLaunchKernelGGL(k_bounds_mean<kMeanChunkShared>, dim3(sv.nsys), dim3(kBlock), 0, s, c->px, c->py, c->pz,
                           sv, bounds4, (const uint32_t*)nullptr);
    }
    hipLaunchKernelGGL(k_bounds_size, dim3(grid_for(c->max_sys_n, kBlock * 8, 2048), sv.nsys), dim3(kBlock), 0, s,
                       c->px, c->py, c->pz, sv, bounds4);
    return hipGetLastError();
}

hipError_t launch_morton(const double* px, const double* py, const double* pz, size_t n, const uint32_t* sysid,
                         const double* b4, uint64_t* keys, uint32_t* vals, hipStream_t s, uint32_t* nonfinite) {
    hipLaunchKernelGGL(k_morton, dim3(grid_for(n)), dim3(kBlock), 0, s, px, py, pz, n, sysid, b4, keys, vals,
                       nonfinite);
    return hipGetLastError();
}

hipError_t launch_treekeys(const double* px, const double* py, const double* pz, size_t n, const uint32_t* sysid,
                           const double* b4, double tscale, uint64_t* keys, uint32_t* vals, hipStream_t s) {
    hipLaunchKernelGGL(k_treekeys, dim3(grid_for(n)), dim3(kBlock), 0, s, px, py, pz, n, sysid, b4, tscale, keys,
                       vals);
    return hipGetLastError();
}

hipError_t launch_iota(uint32_t* vals, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(kBlock), 0, s, vals, n);
    return hipGetLastError();
}

hipError_t launch_gather(double* const* src, double* const* dst, int narr, const uint32_t* perm, size_t n,
                         hipStream_t s) {
    GatherArgs g{};
    g.narr = narr;
    for (int a = 0; a < narr; ++a) {
        g.src[a] = src[a];
        g.dst[a] = dst[a];
    }
    hipLaunchKernelGGL(k_gather, dim3(grid_for(n)), dim3(kBlock), 0, s, g, perm, n);
    return hipGetLastError();
}

static BodyArrays body_arrays(bhx_ctx* c) {
    BodyArrays b{};
    b.px = c->px; b.py = c->py; b.pz = c->pz;
    b.vx = c->vx; b.vy = c->vy; b.vz = c->vz;
    b.m = c->m; b.ax = c->ax; b.ay = c->ay; b.az = c->az;
    b.cpx = c->cpx; b.cpy = c->cpy; b.cpz = c->cpz;
    b.cvx = c->cvx; b.cvy = c->cvy; b.cvz = c->cvz;
    b.opx = c->alt[0]; b.opy = c->alt[1]; b.opz = c->alt[2];
    b.ovx = c->alt[3]; b.ovy = c->alt[4]; b.ovz = c->alt[5];
    b.om = c->alt[6];
    b.ocpx = c->alt[7]; b.ocpy = c->alt[8]; b.ocpz = c->alt[9];
    b.ocvx = c->alt[10]; b.ocvy = c->alt[11]; b.ocvz = c->alt[12];
    return b;
}

static void swap_alt(bhx_ctx* c, bool comp) {
    double** cur[13] = {&c->px, &c->py, &c->pz, &c->vx, &c->vy, &c->vz, &c->m,
                        &c->cpx, &c->cpy, &c->cpz, &c->cvx, &c->cvy, &c->cvz};
    const int na = comp ? 13 : 7;
    for (int a = 0; a < na; ++a) {
        double* t = *cur[a];
        *cur[a] = c->alt[a];
        c->alt[a] = t;
    }
}

hipError_t launch_permute_kick_drift(bhx_ctx* c, const uint32_t* perm, double dt, hipStream_t s, uint32_t* inv) {
    BodyArrays b = body_arrays(c);
    const bool comp = c->cfg.compensated != 0;
    if (comp)
        hipLaunchKernelGGL(k_permute_kick_drift<true>, dim3(grid_for(c->n)), dim3(kBlock), 0, s, b, perm, c->n, dt,
                           c->cfg.permute_acc, inv);
    else
        hipLaunchKernelGGL(k_permute_kick_drift<false>, dim3(grid_for(c->n)), dim3(kBlock), 0, s, b, perm, c->n, dt,
                           c->cfg.permute_acc, inv);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) swap_alt(c, comp);
    return e;
}

void swap_body_buffers(bhx_ctx* c) { swap_alt(c, c->cfg.compensated != 0); }

hipError_t launch_kick_drift(bhx_ctx* c, double dt, hipStream_t s) {
    return launch_permute_kick_drift(c, nullptr, dt, s);
}

hipError_t launch_kick(bhx_ctx* c, double dt, hipStream_t s, uint32_t* nonfinite) {
    if (c->cfg.compensated)
        hipLaunchKernelGGL(k_kick<true>, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->vx, c->vy, c->vz, c->ax, c->ay,
                           c->az, c->cvx, c->cvy, c->cvz, c->n, dt, nonfinite);
    else
        hipLaunchKernelGGL(k_kick<false>, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->vx, c->vy, c->vz, c->ax, c->ay,
                           c->az, c->cvx, c->cvy, c->cvz, c->n, dt, nonfinite);
    return hipGetLastError();
}

hipError_t launch_drift_treekeys(bhx_ctx* c, double dt, const double* b4, hipStream_t s) {
    const BodyArrays b = body_arrays(c);
    const uint32_t* sysid = c->sys().id;
    if (c->cfg.compensated)
        hipLaunchKernelGGL(k_drift_treekeys<true>, dim3(grid_for(c->n)), dim3(kBlock), 0, s, b, c->n, dt, sysid, b4,
                           c->cfg_tscale(), c->tkeys, c->tperm_old);
    else
        hipLaunchKernelGGL(k_drift_treekeys<false>, dim3(grid_for(c->n)), dim3(kBlock), 0, s, b, c->n, dt, sysid, b4,
                           c->cfg_tscale(), c->tkeys, c->tperm_old);
    return hipGetLastError();
}

hipError_t launch_tree_order(bhx_ctx* c, hipStream_t s) {
    hipLaunchKernelGGL(k_tree_order, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->tkeys, c->tperm_old, c->inv, c->n,
                       c->px, c->py, c->pz, c->m, c->tperm, c->tx, c->ty, c->tz, c->tm);
    return hipGetLastError();
}

hipError_t launch_tree_gather(bhx_ctx* c, hipStream_t s) {
    hipLaunchKernelGGL(k_tree_gather, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->px, c->py, c->pz, c->m, c->tperm,
                       c->n, c->tx, c->ty, c->tz, c->tm);
    return hipGetLastError();
}

}  // namespace bhx
