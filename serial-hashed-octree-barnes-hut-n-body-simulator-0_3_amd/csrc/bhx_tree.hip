// bhx_tree.hip — hashed-octree construction on the GPU.
//
// Replaces buildLinearHashedOctreeInPlace and the insertion routines
// (HashedOctree.hpp:211-244, HashedOctree.cpp:450-711) plus the node hash map
// (LinearHashedOctree::nodes, insertHOTNode / lookUpNode, HashedOctree.cpp:
// 111-115,152-205).
//
// The reference inserts bodies one at a time.  Its result does not depend on
// the insertion order: a cell exists iff it is the root, or its parent holds
// more than `leaf_capacity` bodies (1 in the reference) and it holds at least
// one.  Given every body's 21-level path key sorted ascending, the cells of
// level L are the runs of equal L-prefixes, so the whole node set follows
// from adjacent key comparisons:
//   lcp(i)  = levels shared by keys i-1 and i,
//   D(i)    = deepest level at which the run containing body i holds more
//             than B bodies = max over the windows [j, j+B] containing i of
//             the levels shared by keys j and j+B,
//   body i starts the cells of levels lcp(i)+1 .. min(D(i)+1, 21).
// Numbering the cells by (first body, level) gives a depth-first pre-order in
// which a cell's first child is the next record and its subtree is a
// contiguous id range, closed by `skip` = first cell starting at the cell's
// end.  Moments are then combined level by level from the leaves up.
#include "bhx_internal.hpp"

namespace bhx {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;

inline unsigned grid_for(size_t n, int per_block = kBlock, unsigned cap = 65535u * 16u) {
    size_t g = (n + per_block - 1) / per_block;
    if (g == 0) g = 1;
    return (unsigned)(g < cap ? g : cap);
}

// Levels (below the root) shared by two path keys; 21 when equal.
__device__ __forceinline__ int shared_levels(uint64_t a, uint64_t b) {
    const uint64_t x = a ^ b;
    if (x == 0) return kMaxLevel;
    return (__clzll((long long)x) - 1) / 3;
}

__device__ __forceinline__ uint64_t prefix(uint64_t k, int L) { return k >> (63 - 3 * L); }

// ---------------------------------------------------------------- scan (u32)

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_wsum, uint32_t* total) {
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
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t ws = s_wsum[w];
        if (w < wave) wbase += ws;
        tot += ws;
    }
    if (total) *total = tot;
    __syncthreads();
    return wbase + x - v;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ in, size_t n,
                                                        uint32_t* __restrict__ part) {
    __shared__ uint32_t s_wsum[kBlock / 64];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t sum = 0;
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = base + (size_t)j * kBlock + threadIdx.x;
        if (i < n) sum += in[i];
    }
    uint32_t tot;
    block_excl_scan(sum, s_wsum, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_scan_top(uint32_t* __restrict__ part, size_t nparts,
                                                     uint32_t* __restrict__ total) {
    __shared__ uint32_t s_wsum[kBlock / 64];
    uint32_t carry = 0;
    for (size_t b = 0; b < nparts; b += kBlock) {
        const size_t i = b + threadIdx.x;
        const uint32_t v = i < nparts ? part[i] : 0;
        uint32_t tot;
        const uint32_t e = block_excl_scan(v, s_wsum, &tot);
        if (i < nparts) part[i] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const uint32_t* __restrict__ in, size_t n,
                                                      const uint32_t* __restrict__ part,
                                                      uint32_t* __restrict__ out) {
    __shared__ uint32_t s_wsum[kBlock / 64];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    // blocked arrangement: thread t owns items [t*kScanItems, (t+1)*kScanItems)
    uint32_t v[kScanItems];
    uint32_t sum = 0;
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = base + (size_t)threadIdx.x * kScanItems + j;
        v[j] = i < n ? in[i] : 0;
        sum += v[j];
    }
    uint32_t run = part[blockIdx.x] + block_excl_scan(sum, s_wsum, nullptr);
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = base + (size_t)threadIdx.x * kScanItems + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
}

// ---------------------------------------------------------------- node counting

// Per body: how many nodes start at it — the cells of levels lo..hi (the
// counting half of the insertion loop, HashedOctree.cpp:474-505) plus the
// root when the body opens its system.  Windows and shared prefixes never
// reach across a system boundary.
// Also clears the octant masks of the first `nzero` node slots for the
// build's atomic ORs (the grid covers them: about twice the bodies).
__global__ __launch_bounds__(kBlock) void k_count(const uint64_t* __restrict__ keys, uint32_t n, int B, SysView sv,
                                                  uint32_t* __restrict__ nn, uint32_t* __restrict__ nlo,
                                                  uint32_t* __restrict__ child_byte, size_t nzero) {
    for (size_t z = blockIdx.x * (size_t)kBlock + threadIdx.x; z < nzero; z += (size_t)gridDim.x * kBlock)
        child_byte[z] = 0u;
    __shared__ uint64_t s_key[kBlock + 2 * kMaxLeafCap];
    __shared__ uint32_t s_sys[kBlock + 2 * kMaxLeafCap];
    __shared__ int8_t s_w[kBlock + kMaxLeafCap];
    const int64_t b0 = (int64_t)blockIdx.x * kBlock;
    // keys [b0 - B, b0 + kBlock + B)
    for (int t = threadIdx.x; t < kBlock + 2 * B; t += kBlock) {
        const int64_t g = b0 - B + t;
        const bool in = g >= 0 && g < (int64_t)n;
        s_key[t] = in ? keys[g] : 0;
        s_sys[t] = in ? sys_of(sv, (uint32_t)g) : 0xFFFFFFFFu;
    }
    __syncthreads();
    // window levels W(j) for j in [b0 - B, b0 + kBlock): keys j and j+B
    for (int t = threadIdx.x; t < kBlock + B; t += kBlock) {
        const int64_t j = b0 - B + t;
        int w = -1;
        if (j >= 0 && j + B < (int64_t)n && s_sys[t] == s_sys[t + B]) w = shared_levels(s_key[t], s_key[t + B]);
        s_w[t] = (int8_t)w;
    }
    __syncthreads();
    const int64_t i = b0 + threadIdx.x;
    if (i >= (int64_t)n) return;
    const int li = threadIdx.x + B;   // local index of body i in s_key
    const bool first = i == 0 || s_sys[li - 1] != s_sys[li];
    const int lcp = first ? -1 : shared_levels(s_key[li - 1], s_key[li]);
    int D = -1;
    for (int t = 0; t <= B; ++t) {     // windows starting at i-B .. i
        const int w = s_w[threadIdx.x + t];
        if (w > D) D = w;
    }
    const int lo = (lcp + 1 > 1) ? lcp + 1 : 1;
    const int hi = (D + 1 < kMaxLevel) ? D + 1 : kMaxLevel;
    nn[i] = (hi >= lo ? (uint32_t)(hi - lo + 1) : 0u) + (first ? 1u : 0u);
    nlo[i] = (uint32_t)lo;
}

// End of the level-L run starting at i: first j > i (j <= lim) with a different L-prefix.
__device__ uint32_t run_end(const uint64_t* __restrict__ keys, uint32_t lim, uint32_t i, int L) {
    const uint64_t p = prefix(keys[i], L);
    uint32_t lo = i, step = 1;          // keys[lo] has prefix p
    uint32_t hi;
    for (;;) {
        const uint64_t j = (uint64_t)i + step;
        if (j >= lim) { hi = lim; break; }
        if (prefix(keys[j], L) != p) { hi = (uint32_t)j; break; }
        lo = (uint32_t)j;
        step <<= 1;
    }
    while (hi - lo > 1) {               // keys[lo] in run, keys[hi] (or lim) out
        const uint32_t mid = lo + (hi - lo) / 2;
        if (prefix(keys[mid], L) == p) lo = mid; else hi = mid;
    }
    return hi;
}

// Start of the level-L run containing i: first j in [base, i] with the same L-prefix.
__device__ uint32_t run_start(const uint64_t* __restrict__ keys, uint32_t base, uint32_t i, int L) {
    const uint64_t p = prefix(keys[i], L);
    uint32_t hi = i, step = 1;          // keys[hi] has prefix p
    int64_t lo;
    for (;;) {
        const int64_t j = (int64_t)i - step;
        if (j < (int64_t)base) { lo = (int64_t)base - 1; break; }
        if (prefix(keys[j], L) != p) { lo = j; break; }
        hi = (uint32_t)j;
        step <<= 1;
    }
    while ((int64_t)hi - lo > 1) {      // keys[lo] out (or base-1), keys[hi] in
        const int64_t mid = lo + ((int64_t)hi - lo) / 2;
        if (prefix(keys[mid], L) == p) hi = (uint32_t)mid; else lo = mid;
    }
    return hi;
}

struct TreeArgs {
    uint64_t* key;
    uint32_t *start, *count, *skip, *parent, *child_byte, *level, *first;
    uint32_t* child;   // 8 per node, by octant
    double* center;
    double* com;
    double* mass;
    double* quad;
    float* quadf;
    WalkNode* walk;
    WalkVisit32* walk32;
    WalkPay32* pay32;
    float4* tb4;
    uint32_t* tleaf;
    const double* scale32;   // per system: FP32 walk length and mass factors
    double tscale;   // root edge / bounds size (1: reference geometry, 2: enclosing cube)
};

// Child cell geometry — HOTNode(parent, octant, ...) HashedNode.cpp:50-56.
__device__ __forceinline__ void descend_center(uint64_t key, int L, const double* b4, double tscale, double& cx,
                                               double& cy, double& cz, double& s) {
#pragma clang fp contract(off)
    cx = b4[0]; cy = b4[1]; cz = b4[2]; s = b4[3] * tscale;
    for (int l = 1; l <= L; ++l) {
        const unsigned o = (unsigned)(key >> (3 * (L - l))) & 7u;
        s = s * 0.5;
        cx = cx + (((o & 4) ? 1.0 : -1.0) * 0.5) * s;
        cy = cy + (((o & 2) ? 1.0 : -1.0) * 0.5) * s;
        cz = cz + (((o & 1) ? 1.0 : -1.0) * 0.5) * s;
    }
}

// Writes every node's key, first body, level and centre, the roots whole,
// and the parent links that need no search (createChildNode / HOTNode child
// constructor, HashedOctree.cpp:697-711, HashedNode.cpp:48-61).  Node ids:
// the nodes starting at body i are off[i] + (root first, then levels lo,
// lo+1, ...); k_emit_links finishes each node's run.
__global__ __launch_bounds__(kBlock) void k_emit(const uint64_t* __restrict__ keys, uint32_t n, int B, SysView sv,
                                                 const uint32_t* __restrict__ nn, const uint32_t* __restrict__ nlo,
                                                 const uint32_t* __restrict__ off, const double* __restrict__ bounds4,
                                                 uint32_t cap, TreeArgs t, uint32_t* __restrict__ sys_root,
                                                 Counters* __restrict__ ctr) {
    const uint32_t total = ctr->total_nodes;
    const size_t stride = (size_t)gridDim.x * kBlock;
    for (size_t ii = blockIdx.x * (size_t)kBlock + threadIdx.x; ii < n; ii += stride) {
        const uint32_t i = (uint32_t)ii;
        const uint32_t sys = sys_of(sv, i);
        const uint32_t base = sv.off[sys], lim = sv.off[sys + 1];
        const double* b4 = bounds4 + 4 * (size_t)sys;
        const bool first = i == base;
        uint32_t id = off[i];
        if (first) {
            sys_root[sys] = id;
            if (id < cap) {
                t.key[id] = 1;
                t.start[id] = i;
                t.count[id] = lim - i;
                t.skip[id] = lim < n ? off[lim] : total;
                t.parent[id] = kNone;
                t.level[id] = 0;
                t.center[3 * (size_t)id] = b4[0];
                t.center[3 * (size_t)id + 1] = b4[1];
                t.center[3 * (size_t)id + 2] = b4[2];
            }
            ++id;
        }
        const uint32_t cnt = nn[i] - (first ? 1u : 0u);
        if (cnt == 0) continue;
        const int lo = (int)nlo[i];
        for (int L = lo; L < lo + (int)cnt; ++L, ++id) {
            if (id >= cap) break;
            const uint64_t k = prefix(keys[i], L);
            t.key[id] = k;
            t.start[id] = i;
            t.level[id] = (uint32_t)L;
            if (L > lo || L == 1) {   // the parent is the level above here, or the root
                const uint32_t par = L > lo ? id - 1 : off[base];
                t.parent[id] = par;
                atomicOr(&t.child_byte[par], 1u << (unsigned)(k & 7));
                t.child[8 * (size_t)par + (unsigned)(k & 7)] = id;   // one writer per slot
            }
            double cx, cy, cz, s;
            descend_center(k, L, b4, t.tscale, cx, cy, cz, s);
            t.center[3 * (size_t)id] = cx;
            t.center[3 * (size_t)id + 1] = cy;
            t.center[3 * (size_t)id + 2] = cz;
        }
    }
}

// One thread per node below the roots: the level-L run of its first body
// gives its body count and pre-order skip; the shallowest node a body starts
// (below level 1) finds its parent in the run of the level above.  Per node
// rather than per body, so the long searches of the top levels run side by
// side instead of one after another in the thread of a run's first body.
__global__ __launch_bounds__(kBlock) void k_emit_links(const uint64_t* __restrict__ keys, uint32_t n, int B,
                                                       SysView sv, const uint32_t* __restrict__ nlo,
                                                       const uint32_t* __restrict__ off, uint32_t cap, TreeArgs t,
                                                       Counters* __restrict__ ctr) {
    const uint32_t total = ctr->total_nodes;
    const uint32_t nodes = total < cap ? total : cap;
    for (size_t id = blockIdx.x * (size_t)kBlock + threadIdx.x; id < nodes; id += (size_t)gridDim.x * kBlock) {
        const int L = (int)t.level[id];
        if (L == 0) continue;
        const uint32_t i = t.start[id];
        const uint32_t sys = sys_of(sv, i);
        const uint32_t base = sv.off[sys], lim = sv.off[sys + 1];
        const uint32_t end = run_end(keys, lim, i, L);
        t.count[id] = end - i;
        t.skip[id] = end < n ? off[end] : total;
        if (L == kMaxLevel && (int)(end - i) > B) atomicAdd(&ctr->unsplit, 1u);
        if (L == (int)nlo[i] && L > 1) {
            const uint32_t ps = run_start(keys, base, i, L - 1);
            const uint32_t par = off[ps] + (ps == base ? 1u : 0u) + (uint32_t)(L - 1 - (int)nlo[ps]);
            const uint64_t k = t.key[id];
            t.parent[id] = par;
            atomicOr(&t.child_byte[par], 1u << (unsigned)(k & 7));
            t.child[8 * (size_t)par + (unsigned)(k & 7)] = (uint32_t)id;   // one writer per slot
        }
    }
}

// Total node count against the capacity: the clamped count bounds every
// later kernel of the build, the overflow flags tell the host to grow.
// FP32 walk factors of system k: lengths from the tree's root edge, masses
// from the system's first body (tree order), both powers of two.
__device__ __forceinline__ void scale32_of(uint32_t k, const double* __restrict__ bounds4, double tscale,
                                           const SysView& sv, const double* __restrict__ tm,
                                           double* __restrict__ scale) {
    const uint32_t first = sv.off[k];
    scale[2 * (size_t)k] = pow2_inverse_near(bounds4[4 * (size_t)k + 3] * tscale);
    scale[2 * (size_t)k + 1] = sv.off[k + 1] > first ? pow2_inverse_near(fabs(tm[first])) : 1.0;
}

// The node count against the capacity (the clamped count bounds every later
// kernel of the build, the overflow flags tell the host to grow), the
// build's counters and per-level straddler counts cleared, and the FP32 walk
// factors of every system: the build's small jobs in one launch.
__global__ void k_build_setup(uint32_t cap, Counters* __restrict__ ctr, uint32_t* __restrict__ nrest, int nlevels,
                              const double* __restrict__ bounds4, double tscale, SysView sv,
                              const double* __restrict__ tm, double* __restrict__ scale) {
    if (threadIdx.x == 0) {
        const uint32_t total = ctr->total_nodes;
        const bool over = total > cap;
        ctr->node_overflow = over ? 1u : 0u;
        if (over) ctr->overflow_seen = 1u;
        if (total > ctr->max_nodes) ctr->max_nodes = total;
        ctr->nodes_kept = over ? cap : total;
        ctr->unsplit = 0u;
        ctr->hash_fail = 0u;
    }
    for (int l = threadIdx.x; l < nlevels; l += blockDim.x) nrest[l] = 0u;
    for (uint32_t k = threadIdx.x; k < sv.nsys; k += blockDim.x) scale32_of(k, bounds4, tscale, sv, tm, scale);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* __restrict__ p, size_t n, uint64_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------- hash table

// insertHOTNode — HashedOctree.cpp:152-205: open addressing with linear
// probing in HBM; an empty slot holds key 0 (no valid key is 0: every key
// carries the root sentinel bit).  Every node is unique, so an insert takes
// the first empty slot of its probe sequence.
__global__ __launch_bounds__(kBlock) void k_hash_insert(const uint64_t* __restrict__ key,
                                                        const uint32_t* __restrict__ start, SysView sv,
                                                        uint64_t* __restrict__ hkey, uint32_t* __restrict__ hval,
                                                        uint64_t mask, Counters* __restrict__ ctr) {
    const uint32_t total = ctr->nodes_kept;
    for (size_t id = blockIdx.x * (size_t)kBlock + threadIdx.x; id < total; id += (size_t)gridDim.x * kBlock) {
        const uint64_t k = key[id];
        uint64_t h = hash_home(k, sys_of(sv, start[id])) & mask;
        bool done = false;
        for (uint64_t probe = 0; probe <= mask; ++probe) {
            const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(hkey + h), 0ull,
                                                      (unsigned long long)k);
            if (prev == 0ull) {
                hval[h] = (uint32_t)id;
                done = true;
                break;
            }
            h = (h + 1) & mask;
        }
        if (!done) atomicAdd(&ctr->hash_fail, 1u);
    }
}

// lookUpNode — HashedOctree.cpp:111-115, in system `sys`.
__global__ __launch_bounds__(kBlock) void k_hash_lookup(const uint64_t* __restrict__ hkey, const uint32_t* __restrict__ hval,
                                                        const uint32_t* __restrict__ start, SysView sv, uint32_t sys,
                                                        uint64_t mask, const uint64_t* __restrict__ q, size_t n,
                                                        int64_t* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const uint64_t k = q[i];
        int64_t r = -1;
        if (k != 0) {
            uint64_t h = hash_home(k, sys) & mask;
            for (uint64_t probe = 0; probe <= mask; ++probe) {
                const uint64_t s = hkey[h];
                if (s == 0) break;
                const uint32_t v = hval[h];
                if (s == k && v != kHashErased && sys_of(sv, start[v]) == sys) { r = (int64_t)v; break; }
                h = (h + 1) & mask;
            }
        }
        out[i] = r;
    }
}

// ---------------------------------------------------------------- moments

__device__ __forceinline__ void quad_add(double* q, double m, double dx, double dy, double dz) {
    const double d2 = dx * dx + dy * dy + dz * dz;
    q[0] += m * (3.0 * dx * dx - d2);
    q[1] += m * (3.0 * dx * dy);
    q[2] += m * (3.0 * dx * dz);
    q[3] += m * (3.0 * dy * dy - d2);
    q[4] += m * (3.0 * dy * dz);
    q[5] += m * (3.0 * dz * dz - d2);
}

// Payload of the cells of one level, children (level L+1) already done.
//  AS_RUN (what the reference holds after insertion, HashedOctree.cpp:480-504,
//  536-559, 697-711; the moments pass never runs): N = count, barycentre =
//  position of the first body in array order, mass = sum of member masses,
//  quadrupole = 0.
//  INTENDED (computeNodeBaryCenters' aim, HashedOctree.cpp:721-827, written
//  from the physics): mass-weighted centre of mass and the parallel-axis
//  quadrupole sum Q = sum_k [M_k (3 d d^T - |d|^2 I) + Q_k].
//  Children's payloads come through `Child`: the node arrays in global memory,
//  or the slice's LDS copy when the node's whole subtree lies in the slice.
//  The node's own payload is written to the node arrays and returned.
struct Payload {
    double M, cx, cy, cz;
    double q[6];
    uint32_t first;
};

struct ChildFromGlobal {
    static constexpr bool kPreload = true;   // far memory: load all children before summing
    const TreeArgs& t;
    __device__ double mass(uint32_t c) const { return t.mass[c]; }
    __device__ double com(uint32_t c, int k) const { return t.com[3 * (size_t)c + k]; }
    __device__ uint32_t first(uint32_t c) const { return t.first[c]; }
    __device__ double quad(uint32_t c, int k) const { return t.quad[6 * (size_t)c + k]; }
};

struct ChildFromSlice {
    static constexpr bool kPreload = false;  // LDS: child by child keeps registers (and occupancy) low
    const double* m;
    const double (*c)[3];
    const uint32_t* f;
    const double (*q)[6];
    uint32_t b0;
    __device__ double mass(uint32_t i) const { return m[i - b0]; }
    __device__ double com(uint32_t i, int k) const { return c[i - b0][k]; }
    __device__ uint32_t first(uint32_t i) const { return f[i - b0]; }
    __device__ double quad(uint32_t i, int k) const { return q[i - b0][k]; }
};

constexpr int kStage = 1024;     // leaf bodies staged in LDS at once by the slice kernel

// What combining node p reads of its own records, loaded in one batch
// (ahead of the level loop in the slice kernel: none of it depends on the
// children).
struct NodeInfo {
    uint32_t L, st, cnt, cb, skip;
    uint32_t ch[8];          // child id by octant (where cb has the bit)
    double lam, mu;          // the system's FP32 factors
    double b4[4];            // the system's root bounds
};

__device__ __forceinline__ NodeInfo node_info(uint32_t p, const TreeArgs& t, const SysView& sv,
                                              const double* __restrict__ bounds4) {
    NodeInfo I;
    I.L = t.level[p];
    I.st = t.start[p];
    I.cnt = t.count[p];
    I.cb = t.child_byte[p];
    I.skip = t.skip[p];
    const uint4 ca = *reinterpret_cast<const uint4*>(t.child + 8 * (size_t)p);
    const uint4 cb4 = *reinterpret_cast<const uint4*>(t.child + 8 * (size_t)p + 4);
    I.ch[0] = ca.x; I.ch[1] = ca.y; I.ch[2] = ca.z; I.ch[3] = ca.w;
    I.ch[4] = cb4.x; I.ch[5] = cb4.y; I.ch[6] = cb4.z; I.ch[7] = cb4.w;
    const uint32_t sysn = sys_of(sv, I.st);
    I.lam = t.scale32[2 * (size_t)sysn];
    I.mu = t.scale32[2 * (size_t)sysn + 1];
#pragma unroll
    for (int k = 0; k < 4; ++k) I.b4[k] = bounds4[4 * (size_t)sysn + k];
    return I;
}

// The node's records: FP64 payload, the FP64 and FP32 walk records and the
// FP32 payload (the FP32 copies in the system's scaled units, exact powers of
// two).
__device__ __forceinline__ void write_node(uint32_t p, const NodeInfo& I, const TreeArgs& t, bool leaf, double M,
                                           double cx, double cy, double cz, const double* q, uint32_t first) {
    const int L = (int)I.L;
    const uint32_t st = I.st, cnt = I.cnt, cb = I.cb;
    const double lam = I.lam, mu = I.mu;
    t.mass[p] = M;
    t.first[p] = first;
    t.com[3 * (size_t)p] = cx;
    t.com[3 * (size_t)p + 1] = cy;
    t.com[3 * (size_t)p + 2] = cz;
    const double lam2 = lam * lam, qs = mu * lam2;
    for (int k = 0; k < 6; ++k) {
        t.quad[6 * (size_t)p + k] = q[k];
        t.quadf[6 * (size_t)p + k] = (float)(q[k] * qs);
    }
    const double* b4n = I.b4;
    double s = b4n[3] * t.tscale;
    for (int l = 0; l < L; ++l) s = s * 0.5;
    WalkNode w;
    w.cx = cx; w.cy = cy; w.cz = cz;
    w.mass = M;
    w.size2 = s * s;
    w.skip = I.skip;
    w.start = st;
    w.count = cnt;
    w.level_leaf = (uint32_t)L | (leaf ? 256u : 0u);
    w.size2f = (float)(s * s * lam2);
    w.massf = (float)(M * mu);
    t.walk[p] = w;
    WalkVisit32 v32;
    v32.cx = (float)((cx - b4n[0]) * lam);
    v32.cy = (float)((cy - b4n[1]) * lam);
    v32.cz = (float)((cz - b4n[2]) * lam);
    v32.meta = (uint32_t)L | (leaf ? 32u : 0u) | ((cnt < (1u << 26) ? cnt : (1u << 26) - 1) << 6);
    v32.skip = I.skip;
    v32.start = st;
    v32.child_byte = cb;
    v32.size2 = (float)(s * s * lam2);
    t.walk32[p] = v32;
    WalkPay32 p32;
    p32.mass = (float)(M * mu);
    for (int k = 0; k < 6; ++k) p32.q[k] = cnt > 1 ? (float)(q[k] * qs) : 0.0f;
    p32.pad = 0u;
    t.pay32[p] = p32;
}

// Inner nodes only: the leaves are combined by the slice kernel.
template <class Child>
__device__ Payload moments_of(uint32_t p, const NodeInfo& I, int intended, const TreeArgs& t,
                              const double* __restrict__ px, const double* __restrict__ py,
                              const double* __restrict__ pz, const Child& ch) {
    const uint32_t cb = I.cb;
    double M = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    double q[6] = {0, 0, 0, 0, 0, 0};
    uint32_t first = 0xFFFFFFFFu;
    double cx, cy, cz;
    // children in octant order = pre-order: the sums keep their order; the
    // child ids come from the octant table, so the child records are
    // fetched independently rather than along the pre-order skip chain
    const uint32_t* ids0 = I.ch;
    if constexpr (Child::kPreload) {
        // every child's record is loaded before any sum uses it (an absent
        // octant reads the node's own slot, never used), so the loads
        // overlap instead of one round trip per child
        uint32_t ids[8];
        double mc[8], c0[8], c1[8], c2[8];
        uint32_t fc[8];
        double qc[8][6];   // with the other loads: one round trip, not two
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            ids[o] = ((cb >> o) & 1u) ? ids0[o] : p;
            mc[o] = ch.mass(ids[o]);
            c0[o] = ch.com(ids[o], 0);
            c1[o] = ch.com(ids[o], 1);
            c2[o] = ch.com(ids[o], 2);
            fc[o] = ch.first(ids[o]);
            if (intended) {
#pragma unroll
                for (int k = 0; k < 6; ++k) qc[o][k] = ch.quad(ids[o], k);
            }
        }
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            if (!((cb >> o) & 1u)) continue;
            M += mc[o];
            sx += mc[o] * c0[o];
            sy += mc[o] * c1[o];
            sz += mc[o] * c2[o];
            if (fc[o] < first) first = fc[o];
        }
        if (!intended) {
            cx = px[first]; cy = py[first]; cz = pz[first];
        } else {
            cx = sx / M; cy = sy / M; cz = sz / M;
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                if (!((cb >> o) & 1u)) continue;
                quad_add(q, mc[o], c0[o] - cx, c1[o] - cy, c2[o] - cz);
                for (int k = 0; k < 6; ++k) q[k] += qc[o][k];
            }
        }
    } else {
#pragma unroll
        for (int o = 0; o < 8; ++o) {
            if (!((cb >> o) & 1u)) continue;
            const uint32_t c = ids0[o];
            const double mcc = ch.mass(c);
            M += mcc;
            sx += mcc * ch.com(c, 0);
            sy += mcc * ch.com(c, 1);
            sz += mcc * ch.com(c, 2);
            const uint32_t fcc = ch.first(c);
            if (fcc < first) first = fcc;
        }
        if (!intended) {
            cx = px[first]; cy = py[first]; cz = pz[first];
        } else {
            cx = sx / M; cy = sy / M; cz = sz / M;
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                if (!((cb >> o) & 1u)) continue;
                const uint32_t c = ids0[o];
                quad_add(q, ch.mass(c), ch.com(c, 0) - cx, ch.com(c, 1) - cy, ch.com(c, 2) - cz);
                for (int k = 0; k < 6; ++k) q[k] += ch.quad(c, k);
            }
        }
    }
    write_node(p, I, t, false, M, cx, cy, cz, q, first);
    Payload out;
    out.M = M;
    out.cx = cx; out.cy = cy; out.cz = cz;
    for (int k = 0; k < 6; ++k) out.q[k] = q[k];
    out.first = first;
    return out;
}

// Moments in two launches.  Phase 1: workgroup b owns the pre-order slice
// of nodes [256 b, 256 b + 256); a node whose whole subtree [p, skip) lies in
// the slice is combined there: every leaf at once first (leaves read only
// their bodies, so all their loads overlap instead of one level's leaves
// waiting for the level below), then the inner nodes deepest level first,
// with a workgroup barrier between levels (children and parent share the
// compute unit, so no device-wide ordering is needed).  Nodes whose subtree leaves the slice
// are ancestors of the first node after the slice, so a slice has at most
// one per level: they are listed by level (rest[L * slices + i], count
// nrest[L]) for phase 2, where one workgroup combines them level by level.
// Every node is combined once, from its final children in pre-order.
__global__ __launch_bounds__(kBlock) void k_moments_local(const Counters* __restrict__ ctr, SysView sv, int intended,
                                                          TreeArgs t, const double* __restrict__ bounds4,
                                                          const double* __restrict__ px, const double* __restrict__ py,
                                                          const double* __restrict__ pz, const double* __restrict__ tx,
                                                          const double* __restrict__ ty, const double* __restrict__ tz,
                                                          const double* __restrict__ tm,
                                                          const uint32_t* __restrict__ tperm,
                                                          uint32_t* __restrict__ rest, uint32_t* __restrict__ nrest) {
    __shared__ uint32_t s_levels, s_lo, s_hi;
    __shared__ double s_m[kBlock], s_c[kBlock][3], s_q[kBlock][6];
    __shared__ uint32_t s_f[kBlock];
    __shared__ double s_bm[kStage], s_bx[kStage], s_by[kStage], s_bz[kStage];
    __shared__ uint32_t s_bp[kStage];
    if (ctr->node_overflow) return;   // the node table is incomplete: nothing to combine
    const uint32_t total = ctr->total_nodes;
    const uint32_t b0 = blockIdx.x * kBlock;
    if (b0 >= total) return;
    const uint32_t b1 = b0 + kBlock < total ? b0 + kBlock : total;
    const uint32_t p = b0 + threadIdx.x;
    const bool valid = p < b1;
    if (threadIdx.x == 0) {
        s_levels = 0;
        s_lo = 0xFFFFFFFFu;
        s_hi = 0;
    }
    __syncthreads();
    uint32_t L = 0;
    bool local = false, leaf = false;
    NodeInfo info{};
    if (valid) {
        info = node_info(p, t, sv, bounds4);
        L = info.L;
        local = info.skip <= b1;   // always for a leaf: its subtree is itself
        leaf = info.cb == 0u;
        if (local && !leaf) atomicOr(&s_levels, 1u << L);
        else if (!local) rest[L * (size_t)gridDim.x + atomicAdd(&nrest[L], 1u)] = p;
        if (leaf) {
            atomicMin(&s_lo, info.st);
            atomicMax(&s_hi, info.st + info.cnt);
        }
    }
    const ChildFromSlice slice{s_m, s_c, s_f, s_q, b0};
    auto combine = [&]() {
        const Payload r = moments_of(p, info, intended, t, px, py, pz, slice);
        const uint32_t i = threadIdx.x;
        s_m[i] = r.M;
        s_c[i][0] = r.cx; s_c[i][1] = r.cy; s_c[i][2] = r.cz;
        for (int k = 0; k < 6; ++k) s_q[i][k] = r.q[k];
        s_f[i] = r.first;
    };
    __syncthreads();
    // The leaves: their bodies (one tree-order range for the slice) pass
    // through LDS kStage at a time in coalesced loads, and each leaf's thread
    // adds up its own bodies from there in body order: masses, first moments
    // and the least array index, then (the centre of mass known) the
    // quadrupole and the FP32 body copies.
    {
        const uint32_t lo = s_lo, hi = s_hi;
        const uint32_t st = info.st, e = info.st + info.cnt;
        const double lam = info.lam, mu = info.mu;
        double M = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
        double q[6] = {0, 0, 0, 0, 0, 0};
        uint32_t first = 0xFFFFFFFFu;
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1 && leaf) {
                if (!intended || info.cnt == 1) {
                    cx = px[first]; cy = py[first]; cz = pz[first];
                } else {
                    cx = sx / M; cy = sy / M; cz = sz / M;
                }
            }
            for (uint32_t c0 = lo; c0 < hi; c0 += kStage) {
                const uint32_t c1 = hi - c0 < (uint32_t)kStage ? hi : c0 + kStage;
                for (uint32_t j = c0 + threadIdx.x; j < c1; j += kBlock) {
                    s_bm[j - c0] = tm[j];
                    s_bx[j - c0] = tx[j];
                    s_by[j - c0] = ty[j];
                    s_bz[j - c0] = tz[j];
                    if (pass == 0) s_bp[j - c0] = tperm[j];
                }
                __syncthreads();
                if (leaf) {
                    const uint32_t a0 = st > c0 ? st : c0, a1 = e < c1 ? e : c1;
                    for (uint32_t j = a0; j < a1; ++j) {
                        const double mj = s_bm[j - c0], xj = s_bx[j - c0], yj = s_by[j - c0], zj = s_bz[j - c0];
                        if (pass == 0) {
                            M += mj;
                            sx += mj * xj;
                            sy += mj * yj;
                            sz += mj * zj;
                            const uint32_t aj = s_bp[j - c0];
                            if (aj < first) first = aj;
                        } else {
                            const double dx = xj - cx, dy = yj - cy, dz = zj - cz;
                            if (intended && info.cnt > 1) quad_add(q, mj, dx, dy, dz);
                            t.tb4[j] = make_float4((float)(dx * lam), (float)(dy * lam), (float)(dz * lam),
                                                   (float)(mj * mu));
                            t.tleaf[j] = p;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (leaf) {
            write_node(p, info, t, true, M, cx, cy, cz, q, first);
            const uint32_t i = threadIdx.x;
            s_m[i] = M;
            s_c[i][0] = cx; s_c[i][1] = cy; s_c[i][2] = cz;
            for (int k = 0; k < 6; ++k) s_q[i][k] = q[k];
            s_f[i] = first;
        }
    }
    __syncthreads();
    const uint32_t levels = s_levels;
    for (int l = kMaxLevel; l >= 0; --l) {
        if (!((levels >> l) & 1u)) continue;   // block-uniform
        if (local && !leaf && (int)L == l) combine();
        __syncthreads();
    }
}

// Phase 2.  One system: workgroup g (of 64) combines the straddling nodes
// whose octants at levels 1 and 2 (their key's first two digits) are g / 8
// and g % 8, level by level; the last of an octant's eight workgroups to
// finish combines that octant's level-1 node, and the last of those the root
// (each hand-over is one agent-scope fence and one count; the counts, in
// `meet`, return to zero for the next build).  Several systems: workgroup g
// of G combines, level by level, every straddling node of the systems s with
// s mod G = g (systems share no node, so nothing is handed over).
constexpr int kRestSplit = 64;

__device__ __forceinline__ bool meet_last(uint32_t* cnt, uint32_t of) {
    __threadfence();                           // this side's records, visible to every compute unit
    if (atomicAdd(cnt, 1u) + 1u != of) return false;
    __threadfence();                           // the other sides' records
    *cnt = 0u;
    return true;
}

__global__ __launch_bounds__(256) void k_moments_rest(Counters* __restrict__ ctr, SysView sv, int intended,
                                                       TreeArgs t, const double* __restrict__ bounds4,
                                                       const double* __restrict__ px, const double* __restrict__ py,
                                                       const double* __restrict__ pz,
                                                       const uint32_t* __restrict__ rest,
                                                       const uint32_t* __restrict__ nrest, uint32_t slices,
                                                       uint32_t* __restrict__ meet) {
    if (ctr->node_overflow) return;
    const bool split = gridDim.x > 1 && sv.nsys == 1;     // by octants below the root
    const bool bysys = gridDim.x > 1 && sv.nsys > 1;      // by systems
    const uint32_t g = blockIdx.x;
    const int lmin = split ? 2 : 0;
    // a level's node records do not depend on the level below: each thread's
    // first node of the next non-empty level is fetched while this one combines
    int l = kMaxLevel;
    while (l >= lmin && nrest[l] == 0) --l;
    uint32_t pn = 0;
    uint64_t kn = 0;
    NodeInfo next{};
    if (l >= lmin && threadIdx.x < nrest[l]) {
        pn = rest[l * (size_t)slices + threadIdx.x];
        next = node_info(pn, t, sv, bounds4);
        kn = t.key[pn];
    }
    while (l >= lmin) {
        const uint32_t m = nrest[l];
        const uint32_t p0 = pn;
        const NodeInfo info0 = next;
        const uint64_t k0 = kn;
        int ln = l - 1;
        while (ln >= lmin && nrest[ln] == 0) --ln;
        if (ln >= lmin && threadIdx.x < nrest[ln]) {
            pn = rest[ln * (size_t)slices + threadIdx.x];
            next = node_info(pn, t, sv, bounds4);
            kn = t.key[pn];
        }
        for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
            const bool first = i == threadIdx.x;
            const uint32_t p = first ? p0 : rest[l * (size_t)slices + i];
            const uint64_t k = first ? k0 : t.key[p];
            if (split && (uint32_t)((k >> (3 * (l - 2))) & 63u) != g) continue;   // another workgroup's
            const NodeInfo I = first ? info0 : node_info(p, t, sv, bounds4);
            if (bysys && sys_of(sv, I.st) % gridDim.x != g) continue;
            (void)moments_of(p, I, intended, t, px, py, pz, ChildFromGlobal{t});
        }
        __syncthreads();
        l = ln;
    }
    if (split) {
        __builtin_amdgcn_s_waitcnt(0);         // this workgroup's records are written
        __syncthreads();
        if (threadIdx.x == 0 && meet_last(&meet[g / 8], 8u)) {
            // the octant's level-1 node, when it straddles
            for (uint32_t i = 0; i < nrest[1]; ++i) {
                const uint32_t p = rest[1 * (size_t)slices + i];
                if ((uint32_t)(t.key[p] & 7u) == g / 8)
                    (void)moments_of(p, node_info(p, t, sv, bounds4), intended, t, px, py, pz, ChildFromGlobal{t});
            }
            if (meet_last(&meet[8], 8u) && nrest[0] > 0) {
                const uint32_t r = rest[0];
                (void)moments_of(r, node_info(r, t, sv, bounds4), intended, t, px, py, pz, ChildFromGlobal{t});
            }
        }
    }
}

TreeArgs tree_args(bhx_ctx* c) {
    DeviceTree& d = c->tree;
    TreeArgs t;
    t.key = d.key; t.start = d.start; t.count = d.count; t.skip = d.skip; t.parent = d.parent;
    t.child_byte = d.child_byte; t.level = d.level; t.first = d.first; t.center = d.center;
    t.child = d.child;
    t.com = d.com; t.mass = d.mass; t.quad = d.quad; t.quadf = d.quadf; t.walk = d.walk;
    t.walk32 = d.walk32;
    t.pay32 = d.pay32;
    t.scale32 = c->sys_scale;
    t.tb4 = c->tb4;
    t.tleaf = c->tleaf;
    t.tscale = c->cfg_tscale();
    return t;
}

}  // namespace

hipError_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* tmp, uint32_t* total_dev,
                              hipStream_t s) {
    const size_t parts = (n + kScanTile - 1) / kScanTile;
    if (parts == 0) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, total_dev, (size_t)1, 0u);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)parts), dim3(kBlock), 0, s, in, n, tmp);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, s, tmp, parts, total_dev);
    hipLaunchKernelGGL(k_scan_down, dim3((unsigned)parts), dim3(kBlock), 0, s, in, n, tmp, out);
    return hipGetLastError();
}

// Builds the tree of the current bodies (sorted tree keys in c->tkeys) with
// the bounds in slot c->bslot.  With `sync` the node count is read back
// first and *total_out reports it (the caller grows the arrays and calls
// again when it exceeds the capacity); without, nothing waits on the host —
// grids cover the capacity, the kernels stop at the count the device holds,
// and an overflow is left in the counters for the host to find later.
hipError_t build_tree(bhx_ctx* c, hipStream_t s, size_t* total_out, bool sync) {
    const uint32_t n = (uint32_t)c->n;
    const int B = c->cfg.leaf_capacity;
    const double* b4 = c->bounds_slot(c->bslot);
    const SysView sv = c->sys();
    const uint32_t cap = (uint32_t)c->tree.cap;
    c->tree_slot = c->bslot;
    c->tree_tscale = c->cfg_tscale();
    uint32_t* total_dev = &c->ctr->total_nodes;
    uint32_t* nrest = c->tree.lvl;   // per-level counts of slice-straddling nodes
    hipLaunchKernelGGL(k_count, dim3(grid_for(n)), dim3(kBlock), 0, s, c->tkeys, n, B, sv, c->nn, c->nlo,
                       c->tree.child_byte, (size_t)cap);
    hipError_t e = scan_exclusive_u32(c->nn, c->noff, n, c->scan_tmp, total_dev, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_build_setup, dim3(1), dim3(kBlock), 0, s, cap, c->ctr, nrest, kMaxLevel + 1, b4,
                       c->tree_tscale, sv, c->tm, c->sys_scale);
    size_t span = cap;
    if (sync) {
        e = hipMemcpyAsync(c->ctr_host, c->ctr, sizeof(Counters), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        const size_t total = c->ctr_host->total_nodes;
        *total_out = total;
        if (total > cap) return hipSuccess;   // caller grows and retries
        span = total;
    } else {
        *total_out = 0;
    }
    TreeArgs t = tree_args(c);
    hipLaunchKernelGGL(k_emit, dim3(grid_for(n)), dim3(kBlock), 0, s, c->tkeys, n, B, sv, c->nn, c->nlo, c->noff, b4,
                       cap, t, c->sys_root, c->ctr);
    hipLaunchKernelGGL(k_emit_links, dim3(grid_for(span)), dim3(kBlock), 0, s, c->tkeys, n, B, sv, c->nlo, c->noff,
                       cap, t, c->ctr);
    c->hash_valid = false;
    const int intended = c->cfg.mode == BHX_MODE_INTENDED;
    // per-level lists of slice-straddling nodes: (kMaxLevel + 1) x slices
    // entries, within the node capacity (slices = capacity / 256, rounded up)
    const unsigned slices = grid_for(span, kBlock, 0xFFFFFFFFu);
    if ((size_t)(kMaxLevel + 1) * slices > c->tree.cap) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_moments_local, dim3(slices), dim3(kBlock), 0, s, c->ctr, sv, intended, t, b4, c->px, c->py,
                       c->pz, c->tx, c->ty, c->tz, c->tm, c->tperm, c->tree.by_level, nrest);
    const unsigned rest_groups = sv.nsys == 1 ? (unsigned)kRestSplit : (sv.nsys < 256u ? sv.nsys : 256u);
    hipLaunchKernelGGL(k_moments_rest, dim3(rest_groups), dim3(256), 0, s, c->ctr, sv, intended, t,
                       b4, c->px, c->py, c->pz, c->tree.by_level, nrest, (uint32_t)slices, nrest + (kMaxLevel + 1));
    return hipGetLastError();
}

hipError_t ensure_hash(bhx_ctx* c, hipStream_t s) {
    if (c->hash_valid) return hipSuccess;
    const size_t span = c->tree.cap;   // the insert stops at the node count the build left on the device
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(c->tree.hcap)), dim3(kBlock), 0, s, c->tree.hkey, c->tree.hcap,
                       0ull);
    hipLaunchKernelGGL(k_hash_insert, dim3(grid_for(span)), dim3(kBlock), 0, s, c->tree.key, c->tree.start, c->sys(),
                       c->tree.hkey, c->tree.hval, (uint64_t)(c->tree.hcap - 1), c->ctr);
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) c->hash_valid = true;
    return e;
}

hipError_t launch_lookup(bhx_ctx* c, uint32_t sys, const uint64_t* keys, size_t n, int64_t* ids, hipStream_t s) {
    const hipError_t e = ensure_hash(c, s);
    if (e != hipSuccess) return e;
    const DeviceTree& t = c->tree;
    hipLaunchKernelGGL(k_hash_lookup, dim3(grid_for(n)), dim3(kBlock), 0, s, t.hkey, t.hval, t.start, c->sys(), sys,
                       (uint64_t)(t.hcap - 1), keys, n, ids);
    return hipGetLastError();
}

}  // namespace bhx
