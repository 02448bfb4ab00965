// bhx_dropin.hip — device side of the reference's per-body tree API.
//
// ComputeHOTOctreeForce (HashedOctree.hpp:262-272) loops over the bodies
// calling TraverseHOTInteractionList (:336-385) and
// ComputeHOTForceInteractionList (:393-459); code written against those two
// calls, LinearHashedOctree::MAC (HashedOctree.cpp:872-885), and the node
// table members insertHOTNode / deleteNode / pruneEmptyNodes
// (HashedOctree.cpp:152-205,610-660) is served here on the device tree.  The
// batched walks of bhx_walk are the fast path; these kernels reproduce the
// reference's own per-body traversal: one thread per query point runs the
// explicit-stack depth-first search from the root through the node hash
// table, children scanned in octant order, accepted children appended,
// opened ones pushed, last pushed popped first — so every list comes out in
// the reference's order, and the force over a list sums in that order with
// the reference's FP64 operation sequence.
//
// decodeMortonKey (MortonKeys.hpp:204-228) is here too.
#include "bhx_internal.hpp"

namespace bhx {
namespace {

constexpr int kLB = 64;             // threads (query points) per block
// The explicit stack never holds more than 7 pending siblings per level on
// the current path plus the node being expanded: 7 * 21 + 8 < 168.
constexpr int kStackDepth = 168;

struct TreeView {
    const uint64_t* key;
    const uint32_t* count;
    const uint32_t* child_byte;
    const uint32_t* level;
    const uint32_t* start;
    const double* com;
    const double* mass;
    const double* quad;
    const uint64_t* hkey;
    const uint32_t* hval;
    uint64_t mask;
    const uint32_t* tperm;     // tree order -> array index
    const double *px, *py, *pz, *m;   // bodies, array order
    uint32_t nbodies;
    uint32_t nnodes;
    double root_size;          // edge of the root cell
};

__device__ __forceinline__ int64_t hlookup(const TreeView& t, uint64_t k) {
    uint64_t h = hash_home(k, 0) & t.mask;
    for (uint64_t probe = 0; probe <= t.mask; ++probe) {
        const uint64_t s = t.hkey[h];
        if (s == 0) return -1;
        if (s == k) {
            const uint32_t v = t.hval[h];
            return v == kHashErased ? -1 : (int64_t)v;
        }
        h = (h + 1) & t.mask;
    }
    return -1;
}

// nodeBounds.size of a level-L cell: the root edge halved L times (exact),
// the recurrence of HashedNode.cpp:50-56.
__device__ __forceinline__ double cell_size(const TreeView& t, uint32_t lvl) {
    double s = t.root_size;
    for (uint32_t l = 0; l < lvl; ++l) s = s * 0.5;
    return s;
}

// LinearHashedOctree::MAC — HashedOctree.cpp:872-885.
__device__ __forceinline__ bool mac_ref(double size, const double* com, double px, double py, double pz,
                                        double theta) {
    const double dx = com[0] - px;
    const double dy = com[1] - py;
    const double dz = com[2] - pz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double s2 = size * size;
    return (s2 / d2) < (theta * theta);
}

// One term of ComputeHOTForceInteractionList (HashedOctree.hpp:404-455).
__device__ __forceinline__ void ref_term(double& ax, double& ay, double& az, const double* c, double m, double px,
                                         double py, double pz, double eps2, const double* q) {
    const double dx = c[0] - px, dy = c[1] - py, dz = c[2] - pz;
    double D2 = dx * dx + dy * dy + dz * dz;
    D2 += eps2;
    double D1 = 1.0 / sqrt(D2);
    D2 = 1.0 / D2;
    D1 = D1 * D2;
    ax += m * dx * D1;
    ay += m * dy * D1;
    az += m * dz * D1;
    if (q) {
        double qx = q[0] * dx + q[1] * dy + q[2] * dz;
        const double qy = q[1] * dx + q[3] * dy + q[4] * dz;
        const double qz = q[2] * dx + q[4] * dy + q[5] * dz;
        D1 *= D2;
        ax -= qx * D1;
        ay -= qy * D1;
        az -= qz * D1;
        qx = dx * qx + dy * qy + dz * qz;
        qx *= 2.5;
        D1 *= D2;
        ax += qx * dx * D1;
        ay += qx * dy * D1;
        az += qx * dz * D1;
    }
}

// TraverseHOTInteractionList for nq points.  Count pass (kFill false): the
// list length of each point into counts.  Fill pass: the entries from
// offsets[q] on: a node's index (>= 0), or -(1 + i) for body i of a leaf
// holding several bodies (each body except one at the point itself).
template <bool kFill>
__global__ __launch_bounds__(kLB) void k_ref_lists(TreeView t, const double* __restrict__ pts, size_t nq,
                                                    double theta, const uint64_t* __restrict__ offsets,
                                                    uint64_t* __restrict__ counts, int64_t* __restrict__ entries,
                                                    uint32_t* __restrict__ err) {
    __shared__ uint32_t s_stack[kStackDepth][kLB];
    const uint32_t lane = threadIdx.x;
    const size_t q = blockIdx.x * (size_t)kLB + lane;
    if (q >= nq) return;
    const double px = pts[3 * q], py = pts[3 * q + 1], pz = pts[3 * q + 2];
    const uint64_t base = kFill ? offsets[q] : 0;
    uint64_t nl = 0;
    auto emit = [&](int64_t e) {
        if (kFill) entries[base + nl] = e;
        ++nl;
    };
    const int64_t root = hlookup(t, 1ull);
    if (root >= 0) {
        int sp = 0;
        s_stack[sp++][lane] = (uint32_t)root;
        while (sp > 0) {
            const uint32_t id = s_stack[--sp][lane];
            const uint32_t cb = t.child_byte[id];
            if (cb != 0u) {
                const uint64_t k = t.key[id];
                for (uint32_t o = 0; o < 8; ++o) {
                    if (!((cb >> o) & 1u)) continue;
                    const int64_t cid = hlookup(t, (k << 3) | o);
                    if (cid < 0) continue;               // a deleted child
                    if (mac_ref(cell_size(t, t.level[cid]), t.com + 3 * cid, px, py, pz, theta)) {
                        emit(cid);
                    } else if (sp < kStackDepth) {
                        s_stack[sp++][lane] = (uint32_t)cid;
                    } else {
                        atomicOr(err, 1u);               // cannot happen on trees of <= 21 levels
                    }
                }
            } else {
                const uint32_t cnt = t.count[id];
                if (cnt <= 1u) {
                    const double* c = t.com + 3 * (size_t)id;
                    if (c[0] != px || c[1] != py || c[2] != pz) emit((int64_t)id);
                } else {
                    const uint32_t st = t.start[id];
                    for (uint32_t j = st; j < st + cnt && j < t.nbodies; ++j) {
                        const uint32_t i = t.tperm[j];
                        if (t.px[i] != px || t.py[i] != py || t.pz[i] != pz) emit(-1 - (int64_t)i);
                    }
                }
            }
        }
    }
    if (!kFill) counts[q] = nl;
}

// ComputeHOTForceInteractionList for nq points: acc (3 per point, in and out)
// += the terms of entries[offsets[q], offsets[q+1]) in list order.
__global__ __launch_bounds__(kLB) void k_force_lists(TreeView t, const double* __restrict__ pts, size_t nq,
                                                     const uint64_t* __restrict__ offsets,
                                                     const int64_t* __restrict__ entries, double eps2, int quad,
                                                     double* __restrict__ acc, uint32_t* __restrict__ err) {
    const size_t q = blockIdx.x * (size_t)kLB + threadIdx.x;
    if (q >= nq) return;
    const double px = pts[3 * q], py = pts[3 * q + 1], pz = pts[3 * q + 2];
    double ax = acc[3 * q], ay = acc[3 * q + 1], az = acc[3 * q + 2];
    for (uint64_t e = offsets[q]; e < offsets[q + 1]; ++e) {
        const int64_t v = entries[e];
        if (v >= 0) {
            if ((uint64_t)v >= t.nnodes || t.key[v] == 0) {
                atomicOr(err, 1u);
                continue;
            }
            const double* qd = (quad && t.count[v] > 1u) ? t.quad + 6 * (size_t)v : nullptr;
            ref_term(ax, ay, az, t.com + 3 * (size_t)v, t.mass[v], px, py, pz, eps2, qd);
        } else {
            const uint64_t i = (uint64_t)(-1 - v);
            if (i >= t.nbodies) {
                atomicOr(err, 1u);
                continue;
            }
            const double c[3] = {t.px[i], t.py[i], t.pz[i]};
            ref_term(ax, ay, az, c, t.m[i], px, py, pz, eps2, nullptr);
        }
    }
    acc[3 * q] = ax;
    acc[3 * q + 1] = ay;
    acc[3 * q + 2] = az;
}

__global__ __launch_bounds__(kLB) void k_mac(TreeView t, const int64_t* __restrict__ ids,
                                             const double* __restrict__ pts, size_t n, double theta,
                                             uint8_t* __restrict__ out, uint32_t* __restrict__ err) {
    const size_t q = blockIdx.x * (size_t)kLB + threadIdx.x;
    if (q >= n) return;
    const int64_t id = ids[q];
    if (id < 0 || (uint64_t)id >= t.nnodes || t.key[id] == 0) {
        atomicOr(err, 1u);
        out[q] = 0;
        return;
    }
    out[q] = mac_ref(cell_size(t, t.level[id]), t.com + 3 * (size_t)id, pts[3 * q], pts[3 * q + 1], pts[3 * q + 2],
                     theta) ? 1 : 0;
}

// insertHOTNode: the record for `key` at node index `id` (an existing node of
// that key, or the next free index, then entered into the hash table).
struct NodeRec {
    uint64_t key;
    double center[3], com[3], mass, quad[6];
    uint32_t count, first, level, child_byte, start;
};

__global__ void k_node_put(NodeRec r, uint32_t id, int insert, uint64_t* __restrict__ key,
                           uint32_t* __restrict__ count, uint32_t* __restrict__ first, uint32_t* __restrict__ level,
                           uint32_t* __restrict__ child_byte, uint32_t* __restrict__ start,
                           uint32_t* __restrict__ skip, double* __restrict__ center, double* __restrict__ com,
                           double* __restrict__ mass, double* __restrict__ quad, uint64_t* __restrict__ hkey,
                           uint32_t* __restrict__ hval, uint64_t mask, uint32_t* __restrict__ err) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    key[id] = r.key;
    count[id] = r.count;
    first[id] = r.first;
    level[id] = r.level;
    child_byte[id] = r.child_byte;
    start[id] = r.start;
    if (insert) skip[id] = id + 1;    // a new record: no subtree in the pre-order layout
    for (int a = 0; a < 3; ++a) {
        center[3 * (size_t)id + a] = r.center[a];
        com[3 * (size_t)id + a] = r.com[a];
    }
    mass[id] = r.mass;
    for (int k = 0; k < 6; ++k) quad[6 * (size_t)id + k] = r.quad[k];
    if (!insert) return;
    uint64_t h = hash_home(r.key, 0) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        const uint64_t s = hkey[h];
        if (s == r.key || s == 0) {   // the key's erased slot, or the first empty one
            hkey[h] = r.key;
            hval[h] = id;
            return;
        }
        h = (h + 1) & mask;
    }
    atomicOr(err, 1u);                // table full
}

// Erase node `k` from the table: its slot keeps the key with the erased value
// and its record's key becomes 0 (a deleted record).
__device__ __forceinline__ bool erase_key(uint64_t k, uint64_t* __restrict__ key, uint64_t* __restrict__ hkey,
                                          uint32_t* __restrict__ hval, uint64_t mask) {
    uint64_t h = hash_home(k, 0) & mask;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
        const uint64_t s = hkey[h];
        if (s == 0) return false;
        if (s == k) {
            const uint32_t v = hval[h];
            if (v == kHashErased) return false;
            hval[h] = kHashErased;
            key[v] = 0;
            return true;
        }
        h = (h + 1) & mask;
    }
    return false;
}

__global__ void k_node_erase(uint64_t k, uint64_t* __restrict__ key, uint64_t* __restrict__ hkey,
                             uint32_t* __restrict__ hval, uint64_t mask, uint32_t* __restrict__ done) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *done = erase_key(k, key, hkey, hval, mask) ? 1u : 0u;
}

// pruneEmptyNodes(node) — HashedOctree.cpp:610-639: every node at or below
// `root_key` with N == 0 and no children is erased.  (The reference's
// recursion never clears a parent's child bits, so a parent is erased only
// when its own N and childByte are 0: one pass decides every node.)
__global__ __launch_bounds__(256) void k_prune(uint64_t root_key, uint32_t root_level, uint32_t nnodes,
                                               uint64_t* __restrict__ key, const uint32_t* __restrict__ count,
                                               const uint32_t* __restrict__ child_byte,
                                               const uint32_t* __restrict__ level, uint64_t* __restrict__ hkey,
                                               uint32_t* __restrict__ hval, uint64_t mask,
                                               uint32_t* __restrict__ removed) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nnodes; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t k = key[i];
        if (k == 0 || level[i] < root_level) continue;
        if ((k >> (3 * (level[i] - root_level))) != root_key) continue;
        if (count[i] == 0u && child_byte[i] == 0u && erase_key(k, key, hkey, hval, mask)) atomicAdd(removed, 1u);
    }
}

// decodeMortonKey — MortonKeys.hpp:204-228: the three 21-bit coordinates of
// a key, each scaled back as x / 2^21 * 2 * size - size.
__global__ __launch_bounds__(256) void k_decode(const uint64_t* __restrict__ keys, size_t n, double size,
                                                double* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[i];
        uint64_t c[3] = {0, 0, 0};
        for (int b = 0; b < 21; ++b) {
            c[0] |= ((k >> (3 * b + 2)) & 1ull) << b;
            c[1] |= ((k >> (3 * b + 1)) & 1ull) << b;
            c[2] |= ((k >> (3 * b)) & 1ull) << b;
        }
        for (int a = 0; a < 3; ++a) {
            double v = (double)c[a];
            v = v / 2097152.0 * 2.0 * size - size;
            out[3 * i + a] = v;
        }
    }
}

TreeView view_of(bhx_ctx* c) {
    const DeviceTree& t = c->tree;
    TreeView v{};
    v.key = t.key;
    v.count = t.count;
    v.child_byte = t.child_byte;
    v.level = t.level;
    v.start = t.start;
    v.com = t.com;
    v.mass = t.mass;
    v.quad = t.quad;
    v.hkey = t.hkey;
    v.hval = t.hval;
    v.mask = (uint64_t)(t.hcap - 1);
    v.tperm = c->tperm;
    v.px = c->px;
    v.py = c->py;
    v.pz = c->pz;
    v.m = c->m;
    v.nbodies = (uint32_t)c->tree_n;
    v.nnodes = (uint32_t)c->n_nodes;
    v.root_size = c->tree_bounds_h.size() >= 4 ? c->tree_bounds_h[3] * c->tree_tscale : 0.0;
    return v;
}

unsigned blocks_for(size_t n, unsigned per) {
    size_t g = (n + per - 1) / per;
    return (unsigned)(g == 0 ? 1 : g);
}

}  // namespace

hipError_t launch_ref_lists(bhx_ctx* c, const double* pts, size_t nq, double theta, const uint64_t* offsets,
                            uint64_t* counts, int64_t* entries, uint32_t* err, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    if (const hipError_t he = ensure_hash(c, s); he != hipSuccess) return he;
    const TreeView v = view_of(c);
    if (entries)
        hipLaunchKernelGGL(k_ref_lists<true>, dim3(blocks_for(nq, kLB)), dim3(kLB), 0, s, v, pts, nq, theta, offsets,
                           counts, entries, err);
    else
        hipLaunchKernelGGL(k_ref_lists<false>, dim3(blocks_for(nq, kLB)), dim3(kLB), 0, s, v, pts, nq, theta,
                           offsets, counts, entries, err);
    return hipGetLastError();
}

hipError_t launch_force_lists(bhx_ctx* c, const double* pts, size_t nq, const uint64_t* offsets,
                              const int64_t* entries, double* acc, uint32_t* err, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    if (const hipError_t he = ensure_hash(c, s); he != hipSuccess) return he;
    hipLaunchKernelGGL(k_force_lists, dim3(blocks_for(nq, kLB)), dim3(kLB), 0, s, view_of(c), pts, nq, offsets,
                       entries, c->cfg.eps * c->cfg.eps, c->cfg.quadrupole, acc, err);
    return hipGetLastError();
}

hipError_t launch_mac(bhx_ctx* c, const int64_t* ids, const double* pts, size_t n, double theta, uint8_t* out,
                      uint32_t* err, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (const hipError_t he = ensure_hash(c, s); he != hipSuccess) return he;
    hipLaunchKernelGGL(k_mac, dim3(blocks_for(n, kLB)), dim3(kLB), 0, s, view_of(c), ids, pts, n, theta, out, err);
    return hipGetLastError();
}

hipError_t launch_node_put(bhx_ctx* c, const bhx_node& rec, uint32_t id, bool insert, uint32_t* err,
                           hipStream_t s) {
    if (const hipError_t he = ensure_hash(c, s); he != hipSuccess) return he;
    DeviceTree& t = c->tree;
    NodeRec r{};
    r.key = rec.key;
    for (int a = 0; a < 3; ++a) {
        r.center[a] = rec.center[a];
        r.com[a] = rec.com[a];
    }
    r.mass = rec.mass;
    for (int k = 0; k < 6; ++k) r.quad[k] = rec.quad[k];
    r.count = (uint32_t)rec.n;
    r.first = (uint32_t)rec.first;
    r.level = (uint32_t)rec.level;
    r.child_byte = rec.child_byte & 0xFFu;
    r.start = (uint32_t)rec.body_start;
    hipLaunchKernelGGL(k_node_put, dim3(1), dim3(64), 0, s, r, id, insert ? 1 : 0, t.key, t.count, t.first, t.level,
                       t.child_byte, t.start, t.skip, t.center, t.com, t.mass, t.quad, t.hkey, t.hval,
                       (uint64_t)(t.hcap - 1), err);
    return hipGetLastError();
}

hipError_t launch_node_erase(bhx_ctx* c, uint64_t key, uint32_t* done, hipStream_t s) {
    if (const hipError_t he = ensure_hash(c, s); he != hipSuccess) return he;
    DeviceTree& t = c->tree;
    hipLaunchKernelGGL(k_node_erase, dim3(1), dim3(64), 0, s, key, t.key, t.hkey, t.hval, (uint64_t)(t.hcap - 1), done);
    return hipGetLastError();
}

hipError_t launch_prune(bhx_ctx* c, uint64_t root_key, uint32_t root_level, uint32_t* removed, hipStream_t s) {
    if (const hipError_t he = ensure_hash(c, s); he != hipSuccess) return he;
    DeviceTree& t = c->tree;
    if (c->n_nodes == 0) return hipSuccess;
    unsigned g = blocks_for(c->n_nodes, 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_prune, dim3(g), dim3(256), 0, s, root_key, root_level, (uint32_t)c->n_nodes, t.key, t.count,
                       t.child_byte, t.level, t.hkey, t.hval, (uint64_t)(t.hcap - 1), removed);
    return hipGetLastError();
}

hipError_t launch_decode(const uint64_t* keys, size_t n, double size, double* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    unsigned g = blocks_for(n, 256);
    if (g > 65535u * 16u) g = 65535u * 16u;
    hipLaunchKernelGGL(k_decode, dim3(g), dim3(256), 0, s, keys, n, size, out);
    return hipGetLastError();
}

}  // namespace bhx
