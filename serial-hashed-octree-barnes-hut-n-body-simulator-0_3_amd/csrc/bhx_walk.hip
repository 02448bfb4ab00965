// bhx_walk.hip — the theta-criterion tree walk and gravity.
//
// Replaces ComputeHOTOctreeForce (HashedOctree.hpp:262-272),
// TraverseHOTInteractionList (:336-385), LinearHashedOctree::MAC
// (HashedOctree.cpp:872-885) and ComputeHOTForceInteractionList
// (HashedOctree.hpp:393-459).
//
// One wave walks the tree for 64 consecutive bodies in tree order (spatially
// close).  The traversal is wave-uniform and stackless over the pre-order
// node array: every node is fetched once per wave through scalar loads, each
// lane evaluates the reference's per-body opening test for it, and the wave
// descends into a node when any lane opens it.  A per-lane bit per level
// records which lanes opened the current ancestor at that level, so each
// lane interacts with exactly the nodes of its own reference interaction
// list (the same set, not the same order).
#include <algorithm>
#include <mutex>

#include "bhx_internal.hpp"

namespace bhx {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

typedef const __attribute__((address_space(4))) WalkNode* ConstNodes;
typedef const __attribute__((address_space(4))) double* ConstD;
typedef const __attribute__((address_space(4))) float* ConstF;
typedef const __attribute__((address_space(4))) unsigned long long* ConstU64;

// One 64-byte node record as eight 8-byte scalar loads issued together (the
// backend merges them into one burst), before the first field is used.
__device__ __forceinline__ WalkNode load_node(ConstU64 raw, uint32_t p) {
    const size_t b = 8 * (size_t)p;
    const unsigned long long w0 = raw[b], w1 = raw[b + 1], w2 = raw[b + 2], w3 = raw[b + 3];
    const unsigned long long w4 = raw[b + 4], w5 = raw[b + 5], w6 = raw[b + 6], w7 = raw[b + 7];
    WalkNode nd;
    nd.cx = __longlong_as_double((long long)w0);
    nd.cy = __longlong_as_double((long long)w1);
    nd.cz = __longlong_as_double((long long)w2);
    nd.mass = __longlong_as_double((long long)w3);
    nd.size2 = __longlong_as_double((long long)w4);
    nd.skip = (uint32_t)w5;
    nd.start = (uint32_t)(w5 >> 32);
    nd.count = (uint32_t)w6;
    nd.level_leaf = (uint32_t)(w6 >> 32);
    nd.size2f = __uint_as_float((uint32_t)w7);
    nd.massf = __uint_as_float((uint32_t)(w7 >> 32));
    return nd;
}

struct WalkArgs {
    const WalkNode* nodes;
    const WalkVisit32* nodes32;   // FP32 visit records
    const WalkPay32* pay32;       // FP32 payloads (mass, quadrupole)
    const double* scale32;        // per system: FP32 length and mass factors (lam, mu)
    const double* com;       // FP64 centres of mass, 3 per node (the FP32 walk's entries)
    const float4* tb4;       // tree order: offset from the leaf's centre of mass, mass
    const uint32_t* tleaf;   // tree order: node id of the body's leaf
    uint32_t groups;         // walk groups (grid is rounded up to a multiple of 8)
    const uint4* gtab;       // per group: first body, end body (tree order), system
    const uint32_t* sys_root;  // root node of each system
    const uint32_t* abort_flag;  // non-zero: the node table is incomplete, walk nothing
    const double* b4;        // root bounds (cx, cy, cz, size) per system
    const double* quad;      // 6 per node (FP64)
    const float* quadf;      // 6 per node (FP32)
    const double *tx, *ty, *tz, *tm;
    const uint32_t* tperm;
    uint32_t n;
    uint32_t quad_on;
    double theta2, eps2;
    float theta2f, eps2f;
    double tscale;           // root edge / bounds size of the walked tree
    double *ax, *ay, *az;
    double* pot;             // potential per body (array order), potential walks only
    uint32_t* counts;
    uint32_t* qcounts;       // per-body reference walk: the interactions that carry the quadrupole term
    // adjoint walk: weights w = dL/da per body (array order) in, dL/dx of each
    // body's own position out (array order, 3 per body), and per list entry
    // (dL/dX, dL/dM) summed over the targets: 4 per node, 4 per body (tree order)
    const double *wx, *wy, *wz;
    double* gself;
    double* gnode;
    double* gbody;
    // persistent walks: per-XCD counters of the groups taken (8, zero at launch)
    uint32_t* wq;
    // FP32 adjoint: per group a region of list_cap entries (cell ids from the
    // front, tree-order body indices from the back) and the two counts; lists
    // that do not fit are evaluated by the traversal itself
    uint32_t* lists;
    uint2* list_n;
    uint32_t list_cap;
    // and the evaluation kernel's entry sums: one record of `part_stride`
    // doubles per stored entry at its place in the groups' concatenated lists
    // (list_off), with a sort key (cell id, or node_cap + tree-order body
    // index) and the place itself
    const uint32_t* list_off;
    double* part;
    uint64_t* pkey;
    uint32_t* pval;
    uint32_t node_cap;
    const uint32_t* child;   // octant table: 8 child ids per node
    const uint32_t* child_byte;
    // decision digest (FP64 group walk): [0] sum, [1] collect when non-zero, [2] frame index
    unsigned long long* digest;
    // the frame's closing half kick, applied by the group walks as each body's
    // acceleration is written (kick: 0 none, 1 plain, 2 compensated)
    uint32_t kick;
    double kdt;
    double *kvx, *kvy, *kvz;      // velocities (array order)
    double *kcx, *kcy, *kcz;      // their compensation terms
    uint32_t* knonfinite;         // set when a kicked velocity is not finite

};

// One term of ComputeHOTForceInteractionList (HashedOctree.hpp:411-454) in
// the reference's FP64 operation order: D2 = |d|^2 + eps^2,
// D1 = (1/sqrt(D2)) * (1/D2), a += (m*d)*D1, then the quadrupole correction
// for cells holding more than one body.
__device__ __forceinline__ void term64(double& ax, double& ay, double& az, double dx, double dy, double dz,
                                       double d2, double m, double eps2, const double* q) {
#pragma clang fp contract(off)
    double D2 = d2;
    D2 += eps2;
    double D1 = 1.0 / sqrt(D2);
    D2 = 1.0 / D2;
    D1 = D1 * D2;
    ax += m * dx * D1;
    ay += m * dy * D1;
    az += m * dz * D1;
    if (q) {
        double qx = q[0] * dx + q[1] * dy + q[2] * dz;
        double qy = q[1] * dx + q[3] * dy + q[4] * dz;
        double qz = q[2] * dx + q[4] * dy + q[5] * dz;
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

template <bool kQuad>
__global__ __launch_bounds__(kBlock) void k_walk64(WalkArgs a) {
#pragma clang fp contract(off)
    const uint32_t group = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (group >= a.groups || *a.abort_flag) return;
    const uint4 gr = a.gtab[group];
    const uint32_t i = gr.x + lane;
    const bool valid = i < gr.y;
    const uint32_t iv = valid ? i : gr.x;
    const uint32_t root = a.sys_root[gr.z];
    const double x = a.tx[iv], y = a.ty[iv], z = a.tz[iv];
    ConstNodes nodes = (ConstNodes)a.nodes;
    ConstD tx = (ConstD)a.tx;
    ConstD ty = (ConstD)a.ty;
    ConstD tz = (ConstD)a.tz;
    ConstD tm = (ConstD)a.tm;
    double ax = 0.0, ay = 0.0, az = 0.0;
    uint32_t cnt = 0, qcnt = 0;
    uint32_t openbits = 1;                 // bit L: this lane opened the current level-L ancestor
    const bool root_leaf = (nodes[root].level_leaf & 256u) != 0;
    uint32_t p = root_leaf ? root : root + 1;
    const uint32_t end = nodes[root].skip;
    while (p < end) {
        p = __builtin_amdgcn_readfirstlane(p);
        const double cx = nodes[p].cx, cy = nodes[p].cy, cz = nodes[p].cz;
        const double m = nodes[p].mass, s2 = nodes[p].size2;
        const uint32_t ll = nodes[p].level_leaf, nskip = nodes[p].skip;
        if (nskip <= p) break;             // a subtree always ends after its root
        const uint32_t L = ll & 255u;
        const bool leaf = (ll & 256u) != 0;
        const uint32_t ncount = nodes[p].count;
        const bool active = valid && (L == 0 || ((openbits >> (L - 1)) & 1u));
        const double dx = cx - x, dy = cy - y, dz = cz - z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (leaf && ncount == 1) {
            // a one-body leaf enters the list whether it passes (as a node) or
            // is popped (as a leaf), unless it is the body itself
            if (active && (cx != x || cy != y || cz != z)) {
                term64(ax, ay, az, dx, dy, dz, d2, m, a.eps2, nullptr);
                ++cnt;
            }
            p = nskip;
            continue;
        }
        const bool pass = L != 0 && s2 / d2 < a.theta2;
        const bool use = active && pass;
        if (__ballot(use)) {
            if (use) {
                const double* q = nullptr;
                if (kQuad && ncount > 1) q = a.quad + 6 * (size_t)p;
                term64(ax, ay, az, dx, dy, dz, d2, m, a.eps2, q);
                ++cnt;
                if (q) ++qcnt;
            }
        }
        const bool open = active && !pass;
        const uint64_t om = __ballot(open);
        if (leaf) {
            if (om) {
                const uint32_t st = nodes[p].start;
                for (uint32_t j = st; j < st + ncount; ++j) {
                    const double bx = tx[j], by = ty[j], bz = tz[j], bm = tm[j];
                    if (open && (bx != x || by != y || bz != z)) {
                        const double ex = bx - x, ey = by - y, ez = bz - z;
                        term64(ax, ay, az, ex, ey, ez, ex * ex + ey * ey + ez * ez, bm, a.eps2, nullptr);
                        ++cnt;
                    }
                }
            }
            p = nskip;
        } else if (om) {
            openbits = (openbits & ~(1u << L)) | ((open ? 1u : 0u) << L);
            p = p + 1;
        } else {
            p = nskip;
        }
    }
    if (valid) {
        const uint32_t o = a.tperm[i];
        a.ax[o] = ax;
        a.ay[o] = ay;
        a.az[o] = az;
        if (a.counts) a.counts[o] = cnt;
        if (a.qcounts) a.qcounts[o] = qcnt;
    }
}

// FP32 walk in group-relative coordinates: positions are taken relative to
// the wave's first body (an FP64 subtraction, then a rounding to FP32), so
// FP32 carries only the small separations.  Opening test size^2 < theta^2 d^2,
// softened monopole with a hardware reciprocal square root, FP32 sums.
template <bool kQuad>
__global__ __launch_bounds__(kBlock) void k_walk32(WalkArgs a) {
    const uint32_t group = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (group >= a.groups || *a.abort_flag) return;
    const uint4 gr = a.gtab[group];
    const uint32_t i = gr.x + lane;
    const bool valid = i < gr.y;
    const uint32_t iv = valid ? i : gr.x;
    const uint32_t root = a.sys_root[gr.z];
    ConstNodes nodes = (ConstNodes)a.nodes;
    ConstD tx = (ConstD)a.tx;
    ConstD ty = (ConstD)a.ty;
    ConstD tz = (ConstD)a.tz;
    ConstD tm = (ConstD)a.tm;
    ConstF qf = (ConstF)a.quadf;
    const uint32_t g0 = gr.x;
    const double ox = tx[g0], oy = ty[g0], oz = tz[g0];
    const double x = a.tx[iv], y = a.ty[iv], z = a.tz[iv];
    const double lam = a.scale32[2 * (size_t)gr.z], mu = a.scale32[2 * (size_t)gr.z + 1];
    const float xr = (float)((x - ox) * lam), yr = (float)((y - oy) * lam), zr = (float)((z - oz) * lam);
    const float eps2 = (float)(a.eps2 * lam * lam), th2 = a.theta2f;
    float ax = 0.f, ay = 0.f, az = 0.f;
    uint32_t cnt = 0;
    uint32_t openbits = 1;
    const bool root_leaf = (nodes[root].level_leaf & 256u) != 0;
    uint32_t p = root_leaf ? root : root + 1;
    const uint32_t end = nodes[root].skip;
    while (p < end) {
        p = __builtin_amdgcn_readfirstlane(p);
        const double cx = nodes[p].cx, cy = nodes[p].cy, cz = nodes[p].cz;
        const uint32_t ll = nodes[p].level_leaf, nskip = nodes[p].skip;
        if (nskip <= p) break;             // a subtree always ends after its root
        const uint32_t L = ll & 255u;
        const bool leaf = (ll & 256u) != 0;
        const uint32_t ncount = nodes[p].count;
        const float mf = nodes[p].massf, s2 = nodes[p].size2f;
        const bool active = valid && (L == 0 || ((openbits >> (L - 1)) & 1u));
        const float dx = (float)((cx - ox) * lam) - xr, dy = (float)((cy - oy) * lam) - yr,
                    dz = (float)((cz - oz) * lam) - zr;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (leaf && ncount == 1) {
            if (active && (cx != x || cy != y || cz != z)) {
                const float inv = __builtin_amdgcn_rsqf(d2 + eps2);
                const float f = mf * (inv * inv * inv);
                ax += f * dx; ay += f * dy; az += f * dz;
                ++cnt;
            }
            p = nskip;
            continue;
        }
        const bool pass = L != 0 && s2 < th2 * d2;
        const bool use = active && pass;
        if (__ballot(use)) {
            if (use) {
                const float inv = __builtin_amdgcn_rsqf(d2 + eps2);
                const float inv2 = inv * inv;
                float inv3 = inv * inv2;
                const float f = mf * inv3;
                ax += f * dx; ay += f * dy; az += f * dz;
                if (kQuad && ncount > 1) {
                    const float* q = (const float*)0;
                    (void)q;
                    const size_t qb = 6 * (size_t)p;
                    const float q0 = qf[qb], q1 = qf[qb + 1], q2 = qf[qb + 2], q3 = qf[qb + 3], q4 = qf[qb + 4],
                                q5 = qf[qb + 5];
                    const float qx = q0 * dx + q1 * dy + q2 * dz;
                    const float qy = q1 * dx + q3 * dy + q4 * dz;
                    const float qz = q2 * dx + q4 * dy + q5 * dz;
                    const float inv5 = inv3 * inv2;
                    ax -= qx * inv5; ay -= qy * inv5; az -= qz * inv5;
                    const float rqr = 2.5f * (dx * qx + dy * qy + dz * qz) * (inv5 * inv2);
                    ax += rqr * dx; ay += rqr * dy; az += rqr * dz;
                }
                ++cnt;
            }
        }
        const bool open = active && !pass;
        const uint64_t om = __ballot(open);
        if (leaf) {
            if (om) {
                const uint32_t st = nodes[p].start;
                for (uint32_t j = st; j < st + ncount; ++j) {
                    const double bx = tx[j], by = ty[j], bz = tz[j];
                    const float bm = (float)(tm[j] * mu);
                    if (open && (bx != x || by != y || bz != z)) {
                        const float ex = (float)((bx - ox) * lam) - xr, ey = (float)((by - oy) * lam) - yr,
                                    ez = (float)((bz - oz) * lam) - zr;
                        const float inv = __builtin_amdgcn_rsqf(ex * ex + ey * ey + ez * ez + eps2);
                        const float f = bm * (inv * inv * inv);
                        ax += f * ex; ay += f * ey; az += f * ez;
                        ++cnt;
                    }
                }
            }
            p = nskip;
        } else if (om) {
            openbits = (openbits & ~(1u << L)) | ((open ? 1u : 0u) << L);
            p = p + 1;
        } else {
            p = nskip;
        }
    }
    if (valid) {
        const uint32_t o = a.tperm[i];
        const double back = lam * lam / mu;   // scaled units to physical (exact)
        a.ax[o] = (double)ax * back;
        a.ay[o] = (double)ay * back;
        a.az[o] = (double)az * back;
        if (a.counts) a.counts[o] = cnt;
    }
}

// ---------------------------------------------------------------- group walks: the shared terms
//
// The group walks (k_walk_lp32, k_walk_lp64) build ONE interaction list for
// a group of 64 consecutive tree-order bodies that every lane evaluates: a
// cell enters the shared list only when every body accepts it under the
// reference opening test (size^2/d^2 < theta^2, d to the cell's centre of
// mass), and is opened as soon as one body opens it.  The shared list
// therefore refines each body's own reference list (a cell is replaced by
// its children wherever any neighbour needs them), so no body is
// approximated more coarsely than the reference would, and all 64 lanes do
// useful work on every entry.  Below: the list terms and their adjoints.
// FP32 works in offsets from the group's first body (only the offsets are
// rounded); FP64 keeps absolute coordinates and the reference's unfused
// opening test, and forms the term's 1/sqrt(D2) and 1/D2 from a
// Newton-refined reciprocal square root (the per-body walk k_walk64 keeps
// the reference's rounded divisions).

template <typename T>
__device__ __forceinline__ T rsqrt_t(T r2);
template <>
__device__ __forceinline__ float rsqrt_t<float>(float r2) { return __builtin_amdgcn_rsqf(r2); }
// FP64: the hardware estimate refined by two Newton steps y += y (1 - x y^2) / 2
// (each squares the relative error; the result is within ~1 ulp of 1/sqrt(x)).
// Scaling x by 4^k scales every step by 2^-k exactly.  x = 0 gives a NaN or
// infinity, which the zero-separation guard discards.
template <>
__device__ __forceinline__ double rsqrt_t<double>(double r2) {
    double y = __builtin_amdgcn_rsq(r2);
    double e = fma(-r2 * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-r2 * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// 1/r2 from 1/sqrt(r2): squared (within ~2 ulp)
template <typename T>
__device__ __forceinline__ T inv_sq_t(T inv, T r2);
template <>
__device__ __forceinline__ float inv_sq_t<float>(float inv, float) { return inv * inv; }
template <>
__device__ __forceinline__ double inv_sq_t<double>(double inv, double) { return inv * inv; }

// The reference opening test size^2/d^2 < theta^2: exact FP64 form, and the
// FP32 form size^2 < theta^2 d^2 (equal away from rounding at the boundary).
__device__ __forceinline__ bool mac_t(double s2, double d2, double th2, float, float) {
#pragma clang fp contract(off)
    return s2 / d2 < th2;
}
__device__ __forceinline__ bool mac_t(float, float d2, float, float s2f, float th2f) { return s2f < th2f * d2; }

// |d|^2 as the reference forms it: unfused in FP64, fused multiply-adds in FP32
__device__ __forceinline__ double dist2(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}
__device__ __forceinline__ float dist2(float dx, float dy, float dz) { return dx * dx + dy * dy + dz * dz; }

// kGuard: without softening a zero separation (the body itself) must add
// nothing; with softening d = 0 already contributes 0.
// kPot: also the softened potential phi -= M/r + (d.Q.d) / (2 r^5) (a = -grad
// phi for the force above); an entry at zero separation (the body itself)
// adds no potential.
// r^2 = |d|^2 + eps^2 of a list term: three fused multiply-adds, or, where the
// zero-separation test needs |d|^2 itself (no softening, or the potential),
// |d|^2 as the opening test forms it plus eps^2.
template <typename T, bool kNeedD2>
__device__ __forceinline__ void softened_r2(T dx, T dy, T dz, T eps2, T& d2, T& r2) {
    if constexpr (kNeedD2) {
        d2 = dist2(dx, dy, dz);
        r2 = d2 + eps2;
    } else {
        r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
        d2 = r2;   // not read
    }
}

template <typename T, bool kQuad, bool kGuard, bool kPot = false>
__device__ __forceinline__ void eval_nodes(const T (*nd)[4], const T (*q)[6], int nn, T xr, T yr, T zr, T eps2,
                                           T& ax, T& ay, T& az, T& phi) {
    for (int k = 0; k < nn; ++k) {
        const T dx = nd[k][0] - xr, dy = nd[k][1] - yr, dz = nd[k][2] - zr;
        T d2, r2;
        softened_r2<T, kGuard || kPot>(dx, dy, dz, eps2, d2, r2);
        T inv = rsqrt_t<T>(r2);
        T inv2 = inv_sq_t<T>(inv, r2);
        if (kGuard) {   // both factors: in FP64 1/r2 comes from r2, which is 0 here
            inv = d2 > (T)0 ? inv : (T)0;
            inv2 = d2 > (T)0 ? inv2 : (T)0;
        }
        const T inv3 = inv * inv2;
        const T f = nd[k][3] * inv3;
        ax += f * dx;
        ay += f * dy;
        az += f * dz;
        T pt = kPot ? nd[k][3] * inv : (T)0;
        if (kQuad) {
            const T q0 = q[k][0], q1 = q[k][1], q2 = q[k][2], q3 = q[k][3], q4 = q[k][4], q5 = q[k][5];
            const T qx = q0 * dx + q1 * dy + q2 * dz;
            const T qy = q1 * dx + q3 * dy + q4 * dz;
            const T qz = q2 * dx + q4 * dy + q5 * dz;
            const T inv5 = inv3 * inv2;
            const T dqd = dx * qx + dy * qy + dz * qz;
            const T rq = (T)2.5 * dqd * (inv5 * inv2);
            ax += rq * dx - qx * inv5;
            ay += rq * dy - qy * inv5;
            az += rq * dz - qz * inv5;
            if (kPot) pt += (T)0.5 * dqd * inv5;
        }
        if (kPot) phi -= d2 > (T)0 ? pt : (T)0;
    }
}

template <typename T, bool kGuard, bool kPot = false>
__device__ __forceinline__ void eval_bodies(const T (*bd)[4], int nb, T xr, T yr, T zr, T eps2, T& ax, T& ay,
                                            T& az, T& phi) {
    for (int k = 0; k < nb; ++k) {
        const T dx = bd[k][0] - xr, dy = bd[k][1] - yr, dz = bd[k][2] - zr;
        T d2, r2;
        softened_r2<T, kGuard || kPot>(dx, dy, dz, eps2, d2, r2);
        T inv = rsqrt_t<T>(r2);
        T inv2 = inv_sq_t<T>(inv, r2);
        if (kGuard) {   // both factors: in FP64 1/r2 comes from r2, which is 0 here
            inv = d2 > (T)0 ? inv : (T)0;
            inv2 = d2 > (T)0 ? inv2 : (T)0;
        }
        const T inv3 = inv * inv2;
        const T f = bd[k][3] * inv3;
        ax += f * dx;
        ay += f * dy;
        az += f * dz;
        if (kPot) phi -= d2 > (T)0 ? bd[k][3] * inv : (T)0;
    }
}

// ---- adjoint of the list evaluation (frozen lists).  For an entry of mass
// M at X and a target at x with weight w = dL/da, d = X - x, r^2 = |d|^2 +
// eps^2, the term M d r^-3 has Jacobian M (r^-3 I - 3 r^-5 d d^T) in d, so
//   dL/dx (target)  -= M (r^-3 w - 3 r^-5 (d.w) d)
//   dL/dX (entry)   += M (r^-3 w - 3 r^-5 (d.w) d)
//   dL/dM (entry)   += r^-3 (d.w)
// A zero separation (the body itself) has no term.  The target side runs
// lane = target over the entries, the entry side lane = entry over the 64
// targets of the group, so neither needs a cross-lane reduction; entry sums
// go to memory with FP64 atomics.
//
// Quadrupole cells add f = 2.5 (d.Qd) r^-7 d - r^-5 Qd (the forward term).
// With q = Qd, s = d.q, u = w.d, p = Qw:
//   d(w.f)/dd = r^-5 [5 u r^-2 q + 2.5 s r^-2 w - p + (5 (w.q) r^-2 - 17.5 s u r^-4) d]
// (target side -=, entry side +=), and w.f is linear in Q:
//   d(w.f)/dQ = 2.5 u r^-7 d d^T - r^-5 (w d^T + d w^T) / 2
// (symmetric; stored xx xy xz yy yz zz, off-diagonals once), summed per entry
// into its node's slots 4..9 and carried to the bodies by the cell's
// moment definition Q = sum m_k (3 s_k s_k^T - |s_k|^2 I), s_k = x_k - X.
template <typename T>
__device__ __forceinline__ void quad_adj_d(T dx, T dy, T dz, T wx, T wy, T wz, const T* Q, T inv2, T inv5, T& gx,
                                           T& gy, T& gz) {
    const T qx = Q[0] * dx + Q[1] * dy + Q[2] * dz;
    const T qy = Q[1] * dx + Q[3] * dy + Q[4] * dz;
    const T qz = Q[2] * dx + Q[4] * dy + Q[5] * dz;
    const T px = Q[0] * wx + Q[1] * wy + Q[2] * wz;
    const T py = Q[1] * wx + Q[3] * wy + Q[4] * wz;
    const T pz = Q[2] * wx + Q[4] * wy + Q[5] * wz;
    const T s = dx * qx + dy * qy + dz * qz;
    const T u = dx * wx + dy * wy + dz * wz;
    const T wq = wx * qx + wy * qy + wz * qz;
    const T cq = (T)5 * u * inv2, cw = (T)2.5 * s * inv2;
    const T cd = ((T)5 * wq - (T)17.5 * s * u * inv2) * inv2;
    gx += inv5 * (cq * qx + cw * wx - px + cd * dx);
    gy += inv5 * (cq * qy + cw * wy - py + cd * dy);
    gz += inv5 * (cq * qz + cw * wz - pz + cd * dz);
}

template <typename T, bool kQuad = false>
__device__ __forceinline__ void adj_self(const T (*ent)[4], const T (*q)[6], int ne, T xr, T yr, T zr, T wx, T wy,
                                         T wz, T eps2, T& gx, T& gy, T& gz) {
    for (int k = 0; k < ne; ++k) {
        const T dx = ent[k][0] - xr, dy = ent[k][1] - yr, dz = ent[k][2] - zr;
        const T d2 = dist2(dx, dy, dz);
        T inv = rsqrt_t<T>(d2 + eps2);
        inv = d2 > (T)0 ? inv : (T)0;
        const T inv2 = inv * inv;
        const T inv3 = inv * inv2;
        const T dw = dx * wx + dy * wy + dz * wz;
        const T c3 = ent[k][3] * inv3;
        const T c5 = (T)3 * c3 * inv2 * dw;
        gx -= c3 * wx - c5 * dx;
        gy -= c3 * wy - c5 * dy;
        gz -= c3 * wz - c5 * dz;
        if constexpr (kQuad) {
            T hx = 0, hy = 0, hz = 0;
            quad_adj_d<T>(dx, dy, dz, wx, wy, wz, q[k], inv2, inv3 * inv2, hx, hy, hz);
            gx -= hx;
            gy -= hy;
            gz -= hz;
        }
    }
}

// Entry side: lane = entry e over the 64 targets; `sink(e, X, Y, Z, h, G)`
// receives the entry's sums in scaled units: dL/dX (X, Y, Z), dL/dM (h) and,
// for quadrupole cells, dL/dQ (G, xx xy xz yy yz zz).
template <typename T, bool kQuad, class Sink>
__device__ __forceinline__ void adj_entry_sums(const T (*ent)[4], const T (*q)[6], int ne, const T (*tg)[6], T eps2,
                                               uint32_t lane, Sink&& sink) {
    for (int e = (int)lane; e < ne; e += 64) {
        const T X = ent[e][0], Y = ent[e][1], Z = ent[e][2];
        T gx = 0, gy = 0, gz = 0, h = 0;
        T hx = 0, hy = 0, hz = 0;                     // quadrupole part of dL/dX
        T G[6] = {0, 0, 0, 0, 0, 0};                   // dL/dQ
        T Q[6] = {0, 0, 0, 0, 0, 0};
        if constexpr (kQuad) {
#pragma unroll
            for (int k = 0; k < 6; ++k) Q[k] = q[e][k];
        }
        for (int t = 0; t < 64; ++t) {
            const T dx = X - tg[t][0], dy = Y - tg[t][1], dz = Z - tg[t][2];
            const T wx = tg[t][3], wy = tg[t][4], wz = tg[t][5];
            const T d2 = dist2(dx, dy, dz);
            T inv = rsqrt_t<T>(d2 + eps2);
            inv = d2 > (T)0 ? inv : (T)0;
            const T inv2 = inv * inv;
            const T inv3 = inv * inv2;
            const T dw = dx * wx + dy * wy + dz * wz;
            const T c5 = (T)3 * inv3 * inv2 * dw;
            gx += inv3 * wx - c5 * dx;
            gy += inv3 * wy - c5 * dy;
            gz += inv3 * wz - c5 * dz;
            h += inv3 * dw;
            if constexpr (kQuad) {
                const T inv5 = inv3 * inv2;
                quad_adj_d<T>(dx, dy, dz, wx, wy, wz, Q, inv2, inv5, hx, hy, hz);
                const T c = (T)2.5 * dw * inv5 * inv2, hh = (T)0.5 * inv5;
                G[0] += c * dx * dx - inv5 * wx * dx;
                G[1] += c * dx * dy - hh * (wx * dy + wy * dx);
                G[2] += c * dx * dz - hh * (wx * dz + wz * dx);
                G[3] += c * dy * dy - inv5 * wy * dy;
                G[4] += c * dy * dz - hh * (wy * dz + wz * dy);
                G[5] += c * dz * dz - inv5 * wz * dz;
            }
        }
        const T M = ent[e][3];
        sink(e, M * gx + hx, M * gy + hy, M * gz + hz, h, G);
    }
}

// Entry side summed into memory: g4 holds `gs` doubles per entry (4: dL/dX,
// dL/dM; 10: and dL/dQ); sx, sm, sq take the sums back to physical units.
template <typename T, bool kQuad = false>
__device__ __forceinline__ void adj_scatter(const T (*ent)[4], const T (*q)[6], const uint32_t* ids, int ne,
                                            const T (*tg)[6], T eps2, uint32_t lane, double* __restrict__ g4,
                                            double sx = 1.0, double sm = 1.0, double sq = 1.0) {
    constexpr int gs = kQuad ? 10 : 4;
    adj_entry_sums<T, kQuad>(ent, q, ne, tg, eps2, lane, [&](int e, T X, T Y, T Z, T h, const T* G) {
        double* g = g4 + gs * (size_t)ids[e];
        unsafeAtomicAdd(g, (double)X * sx);
        unsafeAtomicAdd(g + 1, (double)Y * sx);
        unsafeAtomicAdd(g + 2, (double)Z * sx);
        unsafeAtomicAdd(g + 3, (double)h * sm);
        if constexpr (kQuad) {
#pragma unroll
            for (int k = 0; k < 6; ++k) unsafeAtomicAdd(g + 4 + k, (double)G[k] * sq);
        }
    });
}


// FP32 group walk over the compact layout: 32-byte visit records (one scalar
// burst per visited cell) hold what the opening test needs — the centre
// relative to the system's root centre in FP32, the edge squared, the
// structure — and 32-byte payloads the mass and quadrupole.  An
// entry's group-relative FP32 centre is formed from the FP64 centres when its
// list is evaluated; opened leaves are gathered as 16-byte offsets from the
// leaf's centre of mass plus that centre's offset from the group origin.
// Every lane builds its own position the same way (its leaf's centre offset +
// its own offset), so a body meets itself at exactly d = 0.
//
// Work distribution: consecutive workgroups land on the 8 XCDs in turn, each
// with its own L2.  Group g is given to block (g mod per) * 8 + g / per, so
// XCD x walks the contiguous tree-order slice of groups [x*per, (x+1)*per):
// its L2 holds that region's cells instead of the whole tree.

__device__ __forceinline__ uint32_t xcd_group(uint32_t block, uint32_t groups) {
    const uint32_t per = (groups + 7u) / 8u;
    return (block % 8u) * per + block / 8u;
}

// What a visit reads: one 32-byte visit record, one scalar burst.
struct Node32 {
    float cx, cy, cz, size2;   // centre - root centre; edge squared
    uint32_t meta, skip, start;
};

__device__ __forceinline__ Node32 load_node32(ConstU64 raw, uint32_t p) {
    const size_t b = 4 * (size_t)p;   // one 32-byte WalkVisit32
    const unsigned long long w0 = raw[b], w1 = raw[b + 1], w2 = raw[b + 2], w3 = raw[b + 3];
    Node32 nd;
    nd.cx = __uint_as_float((uint32_t)w0);
    nd.cy = __uint_as_float((uint32_t)(w0 >> 32));
    nd.cz = __uint_as_float((uint32_t)w1);
    nd.meta = (uint32_t)(w1 >> 32);
    nd.skip = (uint32_t)w2;
    nd.start = (uint32_t)(w2 >> 32);
    nd.size2 = __uint_as_float((uint32_t)(w3 >> 32));
    return nd;
}

template <bool kQuad, bool kGuard, bool kPot = false>
__device__ __forceinline__ void eval_list32(const float (*nd)[4], const float (*q)[6], int nn, float xr, float yr,
                                            float zr, float eps2, float& ax, float& ay, float& az, float& phi) {
    // two entries per pass: their LDS reads share one address register
    // (constant offsets) instead of two register moves per entry
#pragma unroll 2
    for (int k = 0; k < nn; ++k) {
        const float dx = nd[k][0] - xr, dy = nd[k][1] - yr, dz = nd[k][2] - zr;
        float inv, D2 = 0.f;
        if (kGuard) {
            const float d2 = dx * dx + dy * dy + dz * dz;
            inv = d2 > 0.f ? __builtin_amdgcn_rsqf(d2 + eps2) : 0.f;
        } else {
            D2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
            inv = __builtin_amdgcn_rsqf(D2);
        }
        const float inv2 = inv * inv;
        float pt = kPot ? nd[k][3] * inv : 0.f;
        if (kQuad) {
            const float q0 = q[k][0], q1 = q[k][1], q2 = q[k][2], q3 = q[k][3], q4 = q[k][4], q5 = q[k][5];
            const float qx = fmaf(q2, dz, fmaf(q1, dy, q0 * dx));
            const float qy = fmaf(q4, dz, fmaf(q3, dy, q1 * dx));
            const float qz = fmaf(q5, dz, fmaf(q4, dy, q2 * dx));
            const float dqd = fmaf(dz, qz, fmaf(dy, qy, dx * qx));
            if (!kGuard) {
                // softened: a += r^-5 (d (M r^2 + 2.5 (d.Q.d) r^-2) - Q.d) with
                // r^2 = |d|^2 + eps^2, the monopole and both quadrupole terms
                // over one r^-5 factor (one multiply fewer per cell)
                const float inv5 = inv * (inv2 * inv2);
                const float cc = fmaf(2.5f * dqd, inv2, nd[k][3] * D2);
                ax = fmaf(inv5, fmaf(dx, cc, -qx), ax);
                ay = fmaf(inv5, fmaf(dy, cc, -qy), ay);
                az = fmaf(inv5, fmaf(dz, cc, -qz), az);
                if (kPot) pt = fmaf(0.5f * dqd, inv5, pt);
            } else {
                const float inv3 = inv * inv2, inv5 = inv3 * inv2;
                const float c = fmaf(2.5f * dqd, inv5 * inv2, nd[k][3] * inv3);
                ax = fmaf(-inv5, qx, ax);
                ay = fmaf(-inv5, qy, ay);
                az = fmaf(-inv5, qz, az);
                ax = fmaf(c, dx, ax);
                ay = fmaf(c, dy, ay);
                az = fmaf(c, dz, az);
                if (kPot) pt = fmaf(0.5f * dqd, inv5, pt);
            }
        } else {
            const float c = nd[k][3] * (inv * inv2);
            ax = fmaf(c, dx, ax);
            ay = fmaf(c, dy, ay);
            az = fmaf(c, dz, az);
        }
        if (kPot) phi -= (dx * dx + dy * dy + dz * dz) > 0.f ? pt : 0.f;
    }
}

// The closing half kick of body o with the acceleration just written for it
// (the same arithmetic as the kick kernel).
__device__ __forceinline__ void walk_kick(const WalkArgs& a, uint32_t o, double ax, double ay, double az) {
    double v0 = a.kvx[o], v1 = a.kvy[o], v2 = a.kvz[o];
    if (a.kick == 2u) {
        double c[3] = {a.kcx[o], a.kcy[o], a.kcz[o]};
        half_kick(v0, v1, v2, c, ax, ay, az, a.kdt);
        a.kcx[o] = c[0]; a.kcy[o] = c[1]; a.kcz[o] = c[2];
    } else {
        half_kick(v0, v1, v2, nullptr, ax, ay, az, a.kdt);
    }
    a.kvx[o] = v0; a.kvy[o] = v1; a.kvz[o] = v2;
    if (a.knonfinite && !(isfinite(v0) && isfinite(v1) && isfinite(v2))) *a.knonfinite = 1u;
}

// List capacities of the FP32 walk: 64 cells per flush (one per lane), 32
// leaf ranges.
constexpr int kCells32 = 64;
constexpr int kRanges32 = 32;
static_assert(kCells32 == 64, "the FP32 cell list is one id per lane");

// ---------------------------------------------------------------- lane-parallel FP32 walk
//
// The shared-list rule of the group walks (a cell enters the group's list
// only when every body of the group accepts it, and is opened as soon as one
// body opens it), with the traversal spread over the lanes instead of one
// cell per visit.  Each step takes up to 64 cells from a stack in LDS, one per lane;
// lane j tests its cell against every body of the group at once through the
// least |d|^2 over the bodies (fl(theta^2 x) is monotone in x, so every body
// accepts exactly when edge^2 < theta^2 min|d|^2); accepted cells join the
// cell list, opened leaves the body list, and the children of opened cells
// (from the octant table, one 32-byte load per opened cell) go on the stack.
// A group's ~2,800 dependent visits become ~50 independent steps.  Should
// the stack not hold the children of every opened cell, the cells left over
// are walked one cell at a time along the pre-order (no cell is lost).
//
// Persistent: the grid is what the chip holds at once.  XCD x (workgroups
// x, x + 8, ... are dispatched to it) owns the contiguous tree-order slice
// of groups [x * per, (x + 1) * per); its waves take groups from that slice
// through an atomic counter, then from the other XCDs' slices, and leave
// when every slice is taken.

// LDS stack of cells to test: 768 entries keep the FP32 walk's LDS at 10 KB
// (16 one-wave workgroups per compute unit, as many as its registers allow;
// 1,024 entries: 14); overflowing cells are walked one at a time
constexpr int kLpStack = 768;
constexpr int kLpCells = 128;      // accepted cell ids waiting for evaluation (64 at a time)
constexpr int kLpBodies = 96;
constexpr int kLpUnd = 128;        // undecided cells (at most 63 + 64 wait at once)

// The next group for this wave (the same value in every lane), or `groups`.
__device__ __forceinline__ uint32_t lp_next_group(uint32_t* wq, uint32_t groups, uint32_t lane) {
    uint32_t g = groups;
    if (lane == 0) {
        const uint32_t per = (groups + 7u) / 8u;
        const uint32_t xcd = blockIdx.x & 7u;
        for (uint32_t k = 0; k < 8; ++k) {
            const uint32_t x = (xcd + k) & 7u;
            const uint32_t lo = x * per;
            const uint32_t len = lo < groups ? (groups - lo < per ? groups - lo : per) : 0u;
            if (len == 0u) continue;
            if (__hip_atomic_load(wq + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len) continue;
            const uint32_t t = atomicAdd(wq + x, 1u);
            if (t < len) {
                g = lo + t;
                break;
            }
        }
    }
    return __shfl(g, 0, 64);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, uint32_t lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if ((int)lane >= off) v += t;
    }
    return v;
}

// |d|^2 of the lane-parallel walk's opening test, in one fixed rounding order:
// fl(fl(fl(dy dy) + dx dx) + dz dz) with fused multiply-adds.  Every place that
// decides a cell (the exact least |d|^2, the box bound, the sample bodies) uses
// this one expression, so their values compare exactly.
__device__ __forceinline__ float lp_d2(float dx, float dy, float dz) {
    return fmaf(dz, dz, fmaf(dx, dx, dy * dy));
}

__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_mind(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_maxd(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Evaluation of the lane-parallel walk's lists.  Sub-group q's cells are the
// K = 64 / S entries nd[q K .. q K + K); lane l (sub-group l / K) walks its
// own K entries.  Bodies are one list for the whole wave; a body entry counts
// for the sub-groups in its mask only.
template <int S, bool kQuad, bool kGuard, bool kPot>
__device__ __forceinline__ void lp_eval_cells(const float (*nd)[4], const float (*q)[6], uint32_t lane, float xr,
                                              float yr, float zr, float eps2, float& ax, float& ay, float& az,
                                              float& phi) {
    constexpr int K = 64 / S;
    const int off = S == 1 ? 0 : (int)(lane / K) * K;
    int n = K;
    asm volatile("" : "+s"(n));   // a loop of K entries, not K unrolled copies (registers)
    eval_list32<kQuad, kGuard, kPot>(nd + off, q + off, n, xr, yr, zr, eps2, ax, ay, az, phi);
}

template <int S, bool kGuard, bool kPot>
__device__ __forceinline__ void lp_eval_bodies(const float (*bd)[4], const uint8_t* bm, int nb, uint32_t lane,
                                               float xr, float yr, float zr, float eps2, float& ax, float& ay,
                                               float& az, float& phi) {
    if constexpr (S == 1) {
        eval_list32<false, kGuard, kPot>(bd, nullptr, nb, xr, yr, zr, eps2, ax, ay, az, phi);
    } else {
        const uint32_t bit = 1u << (lane / (64 / S));
        for (int k = 0; k < nb; ++k) {
            const float m = (bm[k] & bit) ? bd[k][3] : 0.f;
            const float dx = bd[k][0] - xr, dy = bd[k][1] - yr, dz = bd[k][2] - zr;
            float inv;
            if (kGuard) {
                const float d2 = dx * dx + dy * dy + dz * dz;
                inv = d2 > 0.f ? __builtin_amdgcn_rsqf(d2 + eps2) : 0.f;
            } else {
                inv = __builtin_amdgcn_rsqf(fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2))));
            }
            const float c = m * (inv * inv * inv);
            ax = fmaf(c, dx, ax);
            ay = fmaf(c, dy, ay);
            az = fmaf(c, dz, az);
            if (kPot) phi -= (dx * dx + dy * dy + dz * dz) > 0.f ? m * inv : 0.f;
        }
    }
}

template <bool kQuad, bool kGuard, bool kPot, int S, bool kAdj = false>
__global__ __launch_bounds__(64) void k_walk_lp32(WalkArgs a) {
    static_assert(S == 1 || S == 2 || S == 4, "1, 2 or 4 sub-groups per wave");
    static_assert(!kAdj || (S == 1 && !kPot), "the adjoint walk keeps one list per wave and has no potential");
    constexpr int K = 64 / S;                  // bodies per sub-group, cells per sub-group and flush
    constexpr uint32_t kAll = (1u << S) - 1u;
    constexpr int kBodiesLp = kAdj ? 128 : kLpBodies;   // the adjoint's body flushes fill both scatter halves
    __shared__ float s_nd[kCells32][4];
    __shared__ float s_q[kQuad ? kCells32 : 1][6];
    __shared__ float s_bd[kBodiesLp][4];
    __shared__ uint8_t s_bm[S > 1 ? kBodiesLp : 1];
    __shared__ uint32_t s_bid[kAdj ? kBodiesLp : 1];   // adjoint: tree-order index of each body entry
    __shared__ float s_tg[kAdj ? 64 : 1][6];            // adjoint: targets' positions and weights
    __shared__ uint32_t s_rstart[kRanges32], s_roff[kRanges32], s_rleaf[kRanges32];
    __shared__ uint8_t s_rmask[kRanges32];
    __shared__ float s_rdel[kRanges32][3];
    __shared__ float4 s_tgt[64];
    __shared__ uint32_t s_cid[S][kLpCells];
    __shared__ uint32_t s_stack[kLpStack];
    __shared__ uint8_t s_smask[S > 1 ? kLpStack : 1];
    __shared__ uint8_t s_om[S > 1 ? 32 : 1];   // one-cell-at-a-time walk: open mask per level
    __shared__ uint32_t s_und[S == 1 ? kLpUnd : 1];   // cells the bounds leave undecided

    if (*a.abort_flag) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t sg = lane / K;              // this lane's sub-group
    ConstU64 raw = (ConstU64)a.nodes32;
    const uint4* __restrict__ rec4 = reinterpret_cast<const uint4*>(a.nodes32);
    ConstD tx = (ConstD)a.tx;
    ConstD ty = (ConstD)a.ty;
    ConstD tz = (ConstD)a.tz;
    const float th2 = a.theta2f;

    for (;;) {
        const uint32_t grp = lp_next_group(a.wq, a.groups, lane);
        if (grp >= a.groups) break;
        const uint4 gr = a.gtab[grp];
        const uint32_t g0 = gr.x;
        const uint32_t i = g0 + lane;
        const bool valid = i < gr.y;
        const uint64_t vmask = __builtin_amdgcn_ballot_w64(valid);
        uint32_t live_sg = 0;                  // sub-groups with at least one body
#pragma unroll
        for (int q = 0; q < S; ++q)
            if ((vmask >> (q * K)) & (K == 64 ? ~0ull : ((1ull << K) - 1ull))) live_sg |= 1u << q;
        const uint32_t iv = valid ? i : g0;
        const uint32_t rid = a.sys_root[gr.z];
        double ox = tx[g0], oy = ty[g0], oz = tz[g0];
        asm volatile("" : "+v"(ox), "+v"(oy), "+v"(oz));
        const uint32_t own = a.tleaf[iv];
        const float4 ob = a.tb4[iv];
        const double lam = a.scale32[2 * (size_t)gr.z], mu = a.scale32[2 * (size_t)gr.z + 1];
        const float xr = (float)((a.com[3 * (size_t)own] - ox) * lam) + ob.x;
        const float yr = (float)((a.com[3 * (size_t)own + 1] - oy) * lam) + ob.y;
        const float zr = (float)((a.com[3 * (size_t)own + 2] - oz) * lam) + ob.z;
        const float eps2 = (float)(a.eps2 * lam * lam);
        const double* b4s = a.b4 + 4 * (size_t)gr.z;
        const float xs = (float)((tx[iv] - b4s[0]) * lam), ys = (float)((ty[iv] - b4s[1]) * lam),
                    zs = (float)((tz[iv] - b4s[2]) * lam);
        float ax = 0.f, ay = 0.f, az = 0.f, phi = 0.f;
        // adjoint: the body's weight w = dL/da (scaled units cancel: the walk
        // output is linear in w) and, in ax/ay/az, dL/dx of its own position
        float wxr = 0.f, wyr = 0.f, wzr = 0.f;
        if constexpr (kAdj) {
            if (valid) {
                const uint32_t o = a.tperm[i];
                wxr = (float)a.wx[o];
                wyr = (float)a.wy[o];
                wzr = (float)a.wz[o];
            }
        }
        uint32_t cnt = 0;
        uint32_t lc = 0, lb = 0;               // adjoint: cell and body entries stored
        const size_t lbase = kAdj ? (size_t)grp * a.list_cap : 0;
        int nc[S];                             // cells waiting per sub-group
#pragma unroll
        for (int q = 0; q < S; ++q) nc[q] = 0;
        int nr = 0, nq = 0;                    // leaf ranges, bodies in them

        __syncthreads();                       // the previous group is done with the lists
        s_tgt[lane] = valid ? make_float4(xs, ys, zs, 0.f) : make_float4(INFINITY, INFINITY, INFINITY, 0.f);
        if constexpr (kAdj) {
            s_tg[lane][0] = xr; s_tg[lane][1] = yr; s_tg[lane][2] = zr;
            s_tg[lane][3] = wxr; s_tg[lane][4] = wyr; s_tg[lane][5] = wzr;
        }

        // one chunk of every sub-group's cell list: up to K cells each
        auto flush_cells = [&]() {
            __syncthreads();
            const uint32_t k = lane % K;
            int t = 0;
#pragma unroll
            for (int q = 0; q < S; ++q)
                if ((int)sg == q) t = nc[q] < K ? nc[q] : K;
            const bool has = (int)k < t;
            const uint32_t id = has ? s_cid[sg][k] : 0u;
            bool stored = false;
            if constexpr (kAdj) {
                // the adjoint stores the list for the evaluation kernel while it fits
                if (lc + lb + (uint32_t)t <= a.list_cap) {
                    if (has) a.lists[lbase + lc + lane] = id;
                    lc += (uint32_t)t;
                    stored = true;
                }
            }
            if (!stored) {
                {
                    // lane l gathers entry l: sub-group l / K, position l % K
                    const WalkPay32& c = a.pay32[id];
                    s_nd[lane][0] = (float)((a.com[3 * (size_t)id] - ox) * lam);
                    s_nd[lane][1] = (float)((a.com[3 * (size_t)id + 1] - oy) * lam);
                    s_nd[lane][2] = (float)((a.com[3 * (size_t)id + 2] - oz) * lam);
                    s_nd[lane][3] = has ? c.mass : 0.f;      // padding: a massless cell at a finite place
                    if (kQuad) {
#pragma unroll
                        for (int u = 0; u < 6; ++u) s_q[lane][u] = has ? c.q[u] : 0.f;
                    }
                }
                __syncthreads();
                if constexpr (kAdj) {
                    // the list's first t entries: their ids are s_cid[0][0 .. t)
                    adj_self<float, kQuad>(s_nd, s_q, t, xr, yr, zr, wxr, wyr, wzr, eps2, ax, ay, az);
                    adj_scatter<float, kQuad>(s_nd, s_q, s_cid[0], t, s_tg, eps2, lane, a.gnode,
                                              lam * lam * lam / mu, lam * lam, (lam * lam) * (lam * lam));
                } else {
                    lp_eval_cells<S, kQuad, kGuard, kPot>(s_nd, s_q, lane, xr, yr, zr, eps2, ax, ay, az, phi);
                }
            }
            // the rest of each list moves to its front
            uint32_t rest[kLpCells / 64];
#pragma unroll
            for (int h = 0; h < kLpCells / 64; ++h) rest[h] = 0u;
            cnt += (uint32_t)t;
#pragma unroll
            for (int q = 0; q < S; ++q) {
                const int tq = nc[q] < K ? nc[q] : K;
#pragma unroll
                for (int h = 0; h < kLpCells / 64; ++h) {
                    const int e = (int)lane + 64 * h;
                    rest[h] = e + tq < nc[q] ? s_cid[q][e + tq] : 0u;
                }
                __syncthreads();
#pragma unroll
                for (int h = 0; h < kLpCells / 64; ++h) {
                    const int e = (int)lane + 64 * h;
                    if (e + tq < nc[q]) s_cid[q][e] = rest[h];
                }
                nc[q] -= tq;
            }
            __syncthreads();
        };
        auto flush_bodies = [&]() {
            __syncthreads();
            if constexpr (kAdj) {
                if (lc + lb + (uint32_t)nq <= a.list_cap) {
#pragma unroll
                    for (int h = 0; h < (kBodiesLp + 63) / 64; ++h) {
                        const int e = (int)lane + 64 * h;
                        if (e < nq) {
                            int lo = 0, hi = nr - 1;
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if ((int)s_roff[mid] <= e) lo = mid; else hi = mid - 1;
                            }
                            a.lists[lbase + a.list_cap - 1u - (lb + (uint32_t)e)] =
                                s_rstart[lo] + (uint32_t)(e - (int)s_roff[lo]);
                        }
                    }
                    lb += (uint32_t)nq;
                    cnt += (uint32_t)nq;
                    nr = 0;
                    nq = 0;
                    return;
                }
            }
            if ((int)lane < nr) {
                const size_t l3 = 3 * (size_t)s_rleaf[lane];
                s_rdel[lane][0] = (float)((a.com[l3] - ox) * lam);
                s_rdel[lane][1] = (float)((a.com[l3 + 1] - oy) * lam);
                s_rdel[lane][2] = (float)((a.com[l3 + 2] - oz) * lam);
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < (kBodiesLp + 63) / 64; ++h) {
                const int e = (int)lane + 64 * h;
                if (e < nq) {
                    int lo = 0, hi = nr - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if ((int)s_roff[mid] <= e) lo = mid; else hi = mid - 1;
                    }
                    const uint32_t j = s_rstart[lo] + (uint32_t)(e - (int)s_roff[lo]);
                    const float4 b = a.tb4[j];
                    s_bd[e][0] = s_rdel[lo][0] + b.x;
                    s_bd[e][1] = s_rdel[lo][1] + b.y;
                    s_bd[e][2] = s_rdel[lo][2] + b.z;
                    s_bd[e][3] = b.w;
                    if constexpr (S > 1) s_bm[e] = s_rmask[lo];
                    if constexpr (kAdj) s_bid[e] = j;
                }
            }
            __syncthreads();
            if constexpr (kAdj) {
                adj_self<float>(s_bd, s_q, nq, xr, yr, zr, wxr, wyr, wzr, eps2, ax, ay, az);
                adj_scatter<float>(s_bd, s_q, s_bid, nq, s_tg, eps2, lane, a.gbody, lam * lam * lam / mu, lam * lam);
            } else {
                lp_eval_bodies<S, kGuard, kPot>(s_bd, s_bm, nq, lane, xr, yr, zr, eps2, ax, ay, az, phi);
            }
            if constexpr (S == 1) {
                cnt += (uint32_t)nq;
            } else {
                for (int e = 0; e < nq; ++e) cnt += (s_bm[e] >> sg) & 1u;
            }
            nr = 0;
            nq = 0;
        };
        // the bodies of leaf `leafid` (count bodies from `st`) join the body list
        // for the sub-groups in `mask`
        auto add_leaf = [&](uint32_t leafid, uint32_t st, uint32_t count, uint32_t mask) {
            for (uint32_t k = 0; k < count; k += 64) {
                const int m = (int)((count - k) < 64u ? (count - k) : 64u);
                if (nq + m > kBodiesLp || nr == kRanges32) flush_bodies();
                if (lane == 0) {
                    s_rstart[nr] = st + k;
                    s_roff[nr] = (uint32_t)nq;
                    s_rleaf[nr] = leafid;
                    if constexpr (S > 1) s_rmask[nr] = (uint8_t)mask;
                }
                ++nr;
                nq += m;
            }
        };
        // flush while every list has a full chunk, or some list could not take
        // another step's 64 cells
        auto maybe_flush = [&]() {
            for (;;) {
                bool all_full = true, any_high = false;
#pragma unroll
                for (int q = 0; q < S; ++q) {
                    if (((live_sg >> q) & 1u) && nc[q] < K) all_full = false;
                    if (nc[q] > kLpCells - 64) any_high = true;
                }
                if (!(all_full || any_high)) break;
                flush_cells();
            }
        };
        // the subtree below an opened cell p0 (open for the sub-groups in m0),
        // one cell at a time (the stack is full)
        auto walk_subtree = [&](uint32_t p0, uint32_t end, uint32_t m0) {
            const uint32_t L0 = raw[4 * (size_t)p0 + 1] >> 32 & 31u;   // meta of p0: its level
            if constexpr (S > 1) {
                if (lane == 0) s_om[L0] = (uint8_t)m0;
                __syncthreads();
            }
            uint32_t q = p0 + 1;
            while (q < end) {
                q = __builtin_amdgcn_readfirstlane(q);
                const Node32 nd = load_node32(raw, q);
                const uint32_t nskip = nd.skip > q ? nd.skip : q + 1;
                const uint32_t L = nd.meta & 31u;
                const uint32_t A = S > 1 ? (uint32_t)s_om[L - 1] : 1u;   // open mask of its parent
                const float dx = nd.cx - xs, dy = nd.cy - ys, dz = nd.cz - zs;
                const float d2 = dx * dx + dy * dy + dz * dz;
                const uint64_t ob = __builtin_amdgcn_ballot_w64(!(nd.size2 < th2 * d2)) & vmask;
                uint32_t O = 0;                // sub-groups that open it
#pragma unroll
                for (int s2 = 0; s2 < S; ++s2)
                    if ((ob >> (s2 * K)) & (K == 64 ? ~0ull : ((1ull << K) - 1ull))) O |= 1u << s2;
                O &= A;
                const uint32_t acc = A & ~O & live_sg;
                if (acc) {
#pragma unroll
                    for (int s2 = 0; s2 < S; ++s2)
                        if ((acc >> s2) & 1u) {
                            if (lane == 0) s_cid[s2][nc[s2]] = q;
                            ++nc[s2];
                        }
                    maybe_flush();
                }
                if (O == 0u) {
                    q = nskip;
                } else if (nd.meta & 32u) {
                    add_leaf(q, nd.start, nd.meta >> 6, O);
                    q = nskip;
                } else {
                    if constexpr (S > 1) {
                        __syncthreads();
                        if (lane == 0) s_om[L] = (uint8_t)O;
                        __syncthreads();
                    }
                    q = q + 1;
                }
            }
        };

        int top = 0;
        {
            const Node32 root = load_node32(raw, rid);
            if (root.meta & 32u) {
                add_leaf(rid, root.start, root.meta >> 6, live_sg);   // a one-leaf tree: all its bodies
            } else {
                const uint32_t cb = __builtin_amdgcn_readfirstlane(rec4[2 * (size_t)rid + 1].z);
                if (lane < 8u && ((cb >> lane) & 1u)) {
                    const uint32_t slot = __popc(cb & ((1u << lane) - 1u));
                    s_stack[slot] = a.child[8 * (size_t)rid + lane];
                    if constexpr (S > 1) s_smask[slot] = (uint8_t)live_sg;
                }
                top = __popc(cb);
            }
        }
        __syncthreads();
        // One wave, one list (S == 1): most cells are decided without the least
        // |d|^2 over all 64 bodies.  With lo/hi the group's bounding box, the
        // per-axis distance fl(c - clamp(c, lo, hi)) is at most |fl(c - x_b)| for
        // every body b (rounding is monotone), so lp_d2 of it is at most every
        // body's lp_d2: edge^2 < theta^2 d2_box means every body accepts.  A body's
        // own lp_d2 is at least the least one: edge^2 >= theta^2 d2_b for one of
        // four sample bodies means some body opens.  Cells neither bound decides
        // wait in s_und and get the exact least |d|^2, 64 at a time.  The
        // decisions are the exact rule's; only the order cells join the lists
        // differs.
        float blo_x = 0.f, blo_y = 0.f, blo_z = 0.f, bhi_x = 0.f, bhi_y = 0.f, bhi_z = 0.f;
        float4 smp[4];
        if constexpr (S == 1) {
            blo_x = wave_minf(valid ? xs : INFINITY);
            blo_y = wave_minf(valid ? ys : INFINITY);
            blo_z = wave_minf(valid ? zs : INFINITY);
            bhi_x = wave_maxf(valid ? xs : -INFINITY);
            bhi_y = wave_maxf(valid ? ys : -INFINITY);
            bhi_z = wave_maxf(valid ? zs : -INFINITY);
            const int nv = __popcll(vmask);    // the valid bodies are lanes 0 .. nv - 1
#pragma unroll
            for (int h = 0; h < 4; ++h) smp[h] = s_tgt[(nv * h) / 4 + nv / 8];
        }
        int nund = 0;                          // cells in s_und
        while (top > 0 || nund > 0) {
            // the exact test once 64 cells wait for it, or when nothing else is left
            const bool exact = S == 1 && (nund >= 64 || top == 0);
            const int have = exact ? nund : top;
            const int k = have < 64 ? have : 64;
            const int base = have - k;
            const bool live = (int)lane < k;
            uint32_t p, A;
            if (exact) {
                p = live ? s_und[base + lane] : rid;
                A = live ? 1u : 0u;
                nund = base;
            } else {
                p = live ? s_stack[base + lane] : rid;
                A = live ? (S > 1 ? (uint32_t)s_smask[base + lane] : 1u) : 0u;
                top = base;
            }
            const uint4 r0 = rec4[2 * (size_t)p], r1 = rec4[2 * (size_t)p + 1];
            const float cx = __uint_as_float(r0.x), cy = __uint_as_float(r0.y), cz = __uint_as_float(r0.z);
            const uint32_t meta = r0.w, skip = r1.x, start = r1.y, cb = r1.z;
            const float size2 = __uint_as_float(r1.w);
            uint32_t P = 0;                    // sub-groups where every body accepts the cell
            if (S > 1 || exact) {
                // least |d|^2 over each sub-group's bodies (absent bodies sit at infinity)
#pragma unroll
                for (int q = 0; q < S; ++q) {
                    float m = INFINITY;
#pragma unroll 8
                    for (int b = 0; b < K; ++b) {
                        const float4 t = s_tgt[q * K + b];
                        m = fminf(m, lp_d2(cx - t.x, cy - t.y, cz - t.z));
                    }
                    if (size2 < th2 * m) P |= 1u << q;
                }
            } else {
                const float d2b = lp_d2(cx - __builtin_amdgcn_fmed3f(blo_x, cx, bhi_x),
                                        cy - __builtin_amdgcn_fmed3f(blo_y, cy, bhi_y),
                                        cz - __builtin_amdgcn_fmed3f(blo_z, cz, bhi_z));
                float ds = INFINITY;
#pragma unroll
                for (int h = 0; h < 4; ++h) ds = fminf(ds, lp_d2(cx - smp[h].x, cy - smp[h].y, cz - smp[h].z));
                const bool acc_sure = size2 < th2 * d2b;
                const bool open_sure = !(size2 < th2 * ds);
                const bool und = live && !acc_sure && !open_sure;
                const uint64_t mu = __builtin_amdgcn_ballot_w64(und);
                if (und)
                    s_und[nund + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mu >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)mu, 0u))] = p;
                nund += __popcll(mu);
                if (und) A = 0u;
                if (acc_sure) P = 1u;
            }
            const uint32_t acc = A & P;
            const uint32_t O = A & ~P;          // sub-groups that open it
            const bool leaf = (meta & 32u) != 0u;
            const bool open_in = O != 0u && !leaf;
            // children of the opened cells onto the stack, as far as it holds them
            const uint32_t nch = open_in ? (uint32_t)__popc(cb) : 0u;
            const uint32_t incl = wave_incl_scan(nch, lane);
            const uint32_t room = (uint32_t)(kLpStack - top);
            const uint64_t m_fit = __builtin_amdgcn_ballot_w64(incl <= room);   // a prefix of the lanes
            const int nfit = __popcll(m_fit);
            const bool push = open_in && ((m_fit >> lane) & 1ull);
            const uint64_t m_over = __builtin_amdgcn_ballot_w64(open_in && !push);
            const uint64_t m_leaf = __builtin_amdgcn_ballot_w64(O != 0u && leaf);
            const uint32_t pushed = nfit > 0 ? __shfl(incl, nfit - 1, 64) : 0u;
            __syncthreads();                   // every lane has read its cell before the stack is written
            if (push) {
                const uint4 ca = *reinterpret_cast<const uint4*>(a.child + 8 * (size_t)p);
                const uint4 cbv = *reinterpret_cast<const uint4*>(a.child + 8 * (size_t)p + 4);
                const uint32_t ids[8] = {ca.x, ca.y, ca.z, ca.w, cbv.x, cbv.y, cbv.z, cbv.w};
                uint32_t w = (uint32_t)top + incl - nch;
#pragma unroll
                for (int o = 0; o < 8; ++o)
                    if ((cb >> o) & 1u) {
                        if constexpr (S > 1) s_smask[w] = (uint8_t)O;
                        s_stack[w++] = ids[o];
                    }
            }
            top += (int)pushed;
            // accepted cells join their sub-groups' lists
#pragma unroll
            for (int q = 0; q < S; ++q) {
                const bool aq = (acc >> q) & 1u;
                const uint64_t mq = __builtin_amdgcn_ballot_w64(aq);
                if (aq)
                    s_cid[q][nc[q] + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mq >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mq, 0u))] = p;
                nc[q] += __popcll(mq);
            }
            __syncthreads();
            maybe_flush();
            // opened leaves: their bodies, for the sub-groups that opened them
            for (uint64_t m = m_leaf; m; m &= m - 1) {
                const int l = __builtin_ctzll(m);
                const uint32_t lid = __shfl(p, l, 64), lst = __shfl(start, l, 64), lme = __shfl(meta, l, 64);
                add_leaf(lid, lst, lme >> 6, __shfl(O, l, 64));
            }
            // opened cells whose children found no room: one cell at a time
            for (uint64_t m = m_over; m; m &= m - 1) {
                const int l = __builtin_ctzll(m);
                const uint32_t p0 = __shfl(p, l, 64), end = __shfl(skip, l, 64);
                walk_subtree(p0, end > p0 ? end : p0 + 1, __shfl(O, l, 64));
            }
            __syncthreads();
        }
        for (;;) {
            bool any = false;
#pragma unroll
            for (int q = 0; q < S; ++q) any = any || nc[q] > 0;
            if (!any) break;
            flush_cells();
        }
        if (nq > 0) flush_bodies();
        if constexpr (kAdj) {
            if (lane == 0) a.list_n[grp] = make_uint2(lc, lb);
        }
        if (valid) {
            const uint32_t o = a.tperm[i];
            if constexpr (kAdj) {
                // dL/dx of the own position (the lists evaluated here; the
                // evaluation kernel adds the stored ones) back to physical
                // units (exact): lam^3 / mu
                const double back = lam * lam * lam / mu;
                a.gself[3 * (size_t)o] = (double)ax * back;
                a.gself[3 * (size_t)o + 1] = (double)ay * back;
                a.gself[3 * (size_t)o + 2] = (double)az * back;
            } else {
                const double back = lam * lam / mu;
                const double gx = (double)ax * back, gy = (double)ay * back, gz = (double)az * back;
                a.ax[o] = gx;
                a.ay[o] = gy;
                a.az[o] = gz;
                if constexpr (kPot) a.pot[o] = (double)phi * (lam / mu);
                if (a.kick) walk_kick(a, o, gx, gy, gz);
            }
            if (a.counts) a.counts[o] = cnt;
        }
    }
}

// The stored adjoint lists of k_walk_lp32 / k_walk_lp64<..., kAdj> evaluated with W waves
// per group: wave w takes every W-th chunk of 64 cells, then of 64 bodies,
// gathers it into its own LDS slots, and runs both adjoint passes on it (lane
// = target over the chunk; lane = entry over the 64 targets, whose sums are
// written as the entry's record at its place in the concatenated lists, to be
// added up per entry in a fixed order after a sort by entry).  The waves only
// meet at the end, where their own-position sums are added in wave order and
// added to what the traversal evaluated.  Entry positions are formed exactly
// as the traversal forms them.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <typename T, bool kQuad, int W>
__global__ __launch_bounds__(64 * W) void k_adj_lists(WalkArgs a) {
    constexpr bool k32 = sizeof(T) == 4;       // FP32: scaled group-relative units; FP64: absolute
    __shared__ T s_tg[64][6];
    __shared__ T s_nd[W][64][4];
    __shared__ T s_q[W][kQuad ? 64 : 1][6];
    __shared__ uint32_t s_id[W][64];
    __shared__ T s_red[W][64][3];
    const uint32_t grp = xcd_group(blockIdx.x, a.groups);
    if (grp >= a.groups) return;               // the whole workgroup
    const uint32_t w = threadIdx.x / 64u, lane = threadIdx.x % 64u;
    const uint4 gr = a.gtab[grp];
    const uint32_t g0 = gr.x;
    const uint32_t i = g0 + lane;
    const bool valid = i < gr.y;
    const uint32_t iv = valid ? i : g0;
    ConstD tx = (ConstD)a.tx;
    ConstD ty = (ConstD)a.ty;
    ConstD tz = (ConstD)a.tz;
    double ox = k32 ? tx[g0] : 0.0, oy = k32 ? ty[g0] : 0.0, oz = k32 ? tz[g0] : 0.0;
    asm volatile("" : "+v"(ox), "+v"(oy), "+v"(oz));
    double lam = 1.0, mu = 1.0;
    T xr, yr, zr, eps2;
    if constexpr (k32) {
        const uint32_t own = a.tleaf[iv];
        const float4 ob = a.tb4[iv];
        lam = a.scale32[2 * (size_t)gr.z];
        mu = a.scale32[2 * (size_t)gr.z + 1];
        xr = (float)((a.com[3 * (size_t)own] - ox) * lam) + ob.x;
        yr = (float)((a.com[3 * (size_t)own + 1] - oy) * lam) + ob.y;
        zr = (float)((a.com[3 * (size_t)own + 2] - oz) * lam) + ob.z;
        eps2 = (float)(a.eps2 * lam * lam);
    } else {
        xr = tx[iv];
        yr = ty[iv];
        zr = tz[iv];
        eps2 = a.eps2;
    }
    T wxr = 0, wyr = 0, wzr = 0;
    uint32_t o = 0;
    if (valid) {
        o = a.tperm[i];
        wxr = (T)a.wx[o];
        wyr = (T)a.wy[o];
        wzr = (T)a.wz[o];
    }
    if (w == 0) {
        s_tg[lane][0] = xr; s_tg[lane][1] = yr; s_tg[lane][2] = zr;
        s_tg[lane][3] = wxr; s_tg[lane][4] = wyr; s_tg[lane][5] = wzr;
    }
    __syncthreads();
    const uint2 cnt = a.list_n[grp];
    const size_t base = (size_t)grp * a.list_cap;
    const size_t off = a.list_off[grp];
    constexpr size_t kPs = kQuad ? 10 : 4;
    // back to physical units (exact powers of two; 1 for FP64)
    const double sx = lam * lam * lam / mu, sm = lam * lam, sq = (lam * lam) * (lam * lam);
    T ax = 0, ay = 0, az = 0;
    for (uint32_t c0 = 64u * w; c0 < cnt.x; c0 += 64u * W) {
        const int ne = (int)(cnt.x - c0 < 64u ? cnt.x - c0 : 64u);
        if ((int)lane < ne) {
            const uint32_t id = a.lists[base + c0 + lane];
            if constexpr (k32) {
                const WalkPay32& c = a.pay32[id];
                s_nd[w][lane][0] = (float)((a.com[3 * (size_t)id] - ox) * lam);
                s_nd[w][lane][1] = (float)((a.com[3 * (size_t)id + 1] - oy) * lam);
                s_nd[w][lane][2] = (float)((a.com[3 * (size_t)id + 2] - oz) * lam);
                s_nd[w][lane][3] = c.mass;
                if constexpr (kQuad) {
#pragma unroll
                    for (int u = 0; u < 6; ++u) s_q[w][lane][u] = c.q[u];
                }
            } else {
                const WalkNode nd = a.nodes[id];
                s_nd[w][lane][0] = nd.cx;
                s_nd[w][lane][1] = nd.cy;
                s_nd[w][lane][2] = nd.cz;
                s_nd[w][lane][3] = nd.mass;
                if constexpr (kQuad) {
                    const bool many = nd.count > 1;
#pragma unroll
                    for (int u = 0; u < 6; ++u) s_q[w][lane][u] = many ? a.quad[6 * (size_t)id + u] : 0.0;
                }
            }
            s_id[w][lane] = id;
        }
        wave_lds_sync();
        adj_self<T, kQuad>(s_nd[w], s_q[w], ne, xr, yr, zr, wxr, wyr, wzr, eps2, ax, ay, az);
        adj_entry_sums<T, kQuad>(s_nd[w], s_q[w], ne, s_tg, eps2, lane,
                                 [&](int e, T X, T Y, T Z, T h, const T* G) {
            const size_t pos = off + c0 + (uint32_t)e;
            double* p = a.part + kPs * pos;
            p[0] = (double)X * sx;
            p[1] = (double)Y * sx;
            p[2] = (double)Z * sx;
            p[3] = (double)h * sm;
            if constexpr (kQuad) {
#pragma unroll
                for (int k = 0; k < 6; ++k) p[4 + k] = (double)G[k] * sq;
            }
            a.pkey[pos] = s_id[w][e];
            a.pval[pos] = (uint32_t)pos;
        });
        wave_lds_sync();
    }
    for (uint32_t b0 = 64u * w; b0 < cnt.y; b0 += 64u * W) {
        const int ne = (int)(cnt.y - b0 < 64u ? cnt.y - b0 : 64u);
        if ((int)lane < ne) {
            const uint32_t j = a.lists[base + a.list_cap - 1u - (b0 + lane)];
            if constexpr (k32) {
                const size_t l3 = 3 * (size_t)a.tleaf[j];
                const float4 b = a.tb4[j];
                s_nd[w][lane][0] = (float)((a.com[l3] - ox) * lam) + b.x;
                s_nd[w][lane][1] = (float)((a.com[l3 + 1] - oy) * lam) + b.y;
                s_nd[w][lane][2] = (float)((a.com[l3 + 2] - oz) * lam) + b.z;
                s_nd[w][lane][3] = b.w;
            } else {
                s_nd[w][lane][0] = tx[j];
                s_nd[w][lane][1] = ty[j];
                s_nd[w][lane][2] = tz[j];
                s_nd[w][lane][3] = a.tm[j];
            }
            s_id[w][lane] = j;
        }
        wave_lds_sync();
        adj_self<T>(s_nd[w], s_q[w], ne, xr, yr, zr, wxr, wyr, wzr, eps2, ax, ay, az);
        adj_entry_sums<T, false>(s_nd[w], s_q[w], ne, s_tg, eps2, lane,
                                 [&](int e, T X, T Y, T Z, T h, const T*) {
            const size_t pos = off + cnt.x + b0 + (uint32_t)e;
            double* p = a.part + kPs * pos;
            p[0] = (double)X * sx;
            p[1] = (double)Y * sx;
            p[2] = (double)Z * sx;
            p[3] = (double)h * sm;
            a.pkey[pos] = (uint64_t)a.node_cap + s_id[w][e];
            a.pval[pos] = (uint32_t)pos;
        });
        wave_lds_sync();
    }
    s_red[w][lane][0] = ax;
    s_red[w][lane][1] = ay;
    s_red[w][lane][2] = az;
    __syncthreads();
    if (w == 0 && valid) {
        T gx = s_red[0][lane][0], gy = s_red[0][lane][1], gz = s_red[0][lane][2];
#pragma unroll
        for (int k = 1; k < W; ++k) {
            gx += s_red[k][lane][0];
            gy += s_red[k][lane][1];
            gz += s_red[k][lane][2];
        }
        a.gself[3 * (size_t)o] += (double)gx * sx;
        a.gself[3 * (size_t)o + 1] += (double)gy * sx;
        a.gself[3 * (size_t)o + 2] += (double)gz * sx;
    }
}

// Places of the groups' lists in their concatenation: exclusive scan of the
// stored entry counts (one workgroup); the total goes to off[groups].
__global__ __launch_bounds__(1024) void k_list_scan(const uint2* __restrict__ list_n, uint32_t groups,
                                                    uint32_t* __restrict__ off) {
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_carry;
    const uint32_t t = threadIdx.x, lane = t % 64u, wv = t / 64u;
    if (t == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < groups; b += 1024u) {
        const uint32_t g = b + t;
        const uint2 c = g < groups ? list_n[g] : make_uint2(0u, 0u);
        const uint32_t v = c.x + c.y;
        const uint32_t incl = wave_incl_scan(v, lane);
        if (lane == 63u) s_w[wv] = incl;
        __syncthreads();
        uint32_t wb = 0, tot = 0;
        for (uint32_t k = 0; k < 16u; ++k) {
            if (k < wv) wb += s_w[k];
            tot += s_w[k];
        }
        const uint32_t carry = s_carry;
        if (g < groups) off[g] = carry + wb + incl - v;
        __syncthreads();
        if (t == 0) s_carry = carry + tot;
        __syncthreads();
    }
    if (t == 0) off[groups] = s_carry;
}

// Entry sums after the sort by entry.  k_run_bounds marks where each key's
// run of records starts and ends in the sorted order (keys without records
// keep the empty range the fill left); k_run_sum gives each key one wave,
// whose lanes add every 64th record of the run in sorted order (the order of
// the places, the sort being stable) and then combine the 64 lane sums in a
// fixed butterfly order, and adds the result to the entry's accumulator:
// gnode (`gs` per cell: dL/dX, dL/dM, dL/dQ) or gbody (4 per body).  The
// sums do not depend on timing: the gradients repeat bit for bit.
__global__ __launch_bounds__(256) void k_run_bounds(const uint64_t* __restrict__ key, const uint32_t* __restrict__ ntot,
                                                    uint2* __restrict__ run) {
    const uint32_t n = *ntot;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const uint64_t k = key[i];
        if (i == 0 || key[i - 1] != k) run[k].x = i;
        if (i + 1 == n || key[i + 1] != k) run[k].y = i + 1;
    }
}

__global__ __launch_bounds__(256) void k_fill_runs(uint2* __restrict__ run, size_t nkeys) {
    for (size_t i = blockIdx.x * 256u + threadIdx.x; i < nkeys; i += (size_t)gridDim.x * 256u) run[i] = make_uint2(0u, 0u);
}

template <int GS>
__global__ __launch_bounds__(256) void k_run_sum(const uint2* __restrict__ run, const uint32_t* __restrict__ val,
                                                 const double* __restrict__ part, uint32_t node_cap, size_t nkeys,
                                                 double* __restrict__ gnode, double* __restrict__ gbody) {
    const uint32_t lane = threadIdx.x % 64u;
    for (size_t k = (blockIdx.x * 256u + threadIdx.x) / 64u; k < nkeys; k += (size_t)gridDim.x * 4u) {
        const uint2 r = run[k];
        if (r.y <= r.x) continue;              // the whole wave: no records for this key
        const bool cell = k < node_cap;
        double acc[GS];
#pragma unroll
        for (int u = 0; u < GS; ++u) acc[u] = 0.0;
        for (uint32_t j = r.x + lane; j < r.y; j += 64u) {
            const double* p = part + (size_t)GS * val[j];
            if (cell) {
#pragma unroll
                for (int u = 0; u < GS; ++u) acc[u] += p[u];
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) acc[u] += p[u];
            }
        }
#pragma unroll
        for (int u = 0; u < GS; ++u) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[u] += __shfl_xor(acc[u], off, 64);
        }
        if (lane == 0) {
            if (cell) {
                double* g = gnode + (size_t)GS * k;
#pragma unroll
                for (int u = 0; u < GS; ++u) g[u] += acc[u];
            } else {
                double* g = gbody + 4 * (k - node_cap);
#pragma unroll
                for (int u = 0; u < 4; ++u) g[u] += acc[u];
            }
        }
    }
}

// ---------------------------------------------------------------- lane-parallel FP64 walk
//
// k_walk_lp32's traversal and persistent XCD-sliced queue over the FP64 walk
// records: the reference's opening test size^2 / d^2 < theta^2 in FP64 from
// absolute coordinates (fl(s^2 / x) is non-increasing in x, so every body
// accepts exactly when the test passes at the least |d|^2), and the
// reference's term with 1/sqrt(D2) and 1/D2 within ~1 and ~2 ulp.  One list
// per 64-body wave.
constexpr int kLp64Cells = 64;
constexpr int kLp64Bodies = 64;
constexpr uint64_t kDigestCell = 0x243f6a8885a308d3ull, kDigestLeaf = 0x13198a2e03707344ull;
constexpr uint64_t kDigestFrame = 0x9e3779b97f4a7c15ull;

template <bool kQuad, bool kGuard, bool kPot = false, bool kAdj = false>
__global__ __launch_bounds__(64) void k_walk_lp64(WalkArgs a) {
    static_assert(!(kAdj && kPot), "the adjoint walk has no potential");
    __shared__ double s_nd[kLp64Cells][4];
    __shared__ double s_q[kQuad ? kLp64Cells : 1][6];
    __shared__ double s_bd[kLp64Bodies][4];
    __shared__ uint32_t s_bid[kAdj ? kLp64Bodies : 1];   // adjoint: tree-order index of each body entry
    __shared__ double s_tg[kAdj ? 64 : 1][6];            // adjoint: targets' positions and weights
    __shared__ uint32_t s_rstart[kRanges32], s_roff[kRanges32];
    __shared__ double4 s_tgt[64];
    __shared__ uint32_t s_cid[kLpCells];
    __shared__ uint32_t s_stack[kLpStack];
    __shared__ uint32_t s_und[kLpUnd];        // cells the bounds leave undecided

    if (*a.abort_flag) return;
    const uint32_t lane = threadIdx.x;
    ConstU64 raw = (ConstU64)a.nodes;
    const uint4* __restrict__ rec4 = reinterpret_cast<const uint4*>(a.nodes);
    ConstF qf = (ConstF)a.quadf;
    (void)qf;
    const double th2 = a.theta2;
    const double eps2 = a.eps2;
    // decision digest: every list entry hashed with its group and the frame
    const bool dig = a.digest && a.digest[1] != 0ull;
    const uint64_t salt = dig ? (uint64_t)a.digest[2] * kDigestFrame : 0ull;
    uint64_t hsum = 0;

    for (;;) {
        const uint32_t grp = lp_next_group(a.wq, a.groups, lane);
        if (grp >= a.groups) break;
        const uint4 gr = a.gtab[grp];
        const uint32_t g0 = gr.x;
        const uint32_t i = g0 + lane;
        const bool valid = i < gr.y;
        const uint64_t vmask = __builtin_amdgcn_ballot_w64(valid);
        const uint32_t iv = valid ? i : g0;
        const uint32_t rid = a.sys_root[gr.z];
        const double x = a.tx[iv], y = a.ty[iv], z = a.tz[iv];
        double ax = 0, ay = 0, az = 0, phi = 0;
        uint32_t cnt = 0;
        int nc = 0, nr = 0, nq = 0;
        // adjoint: the body's weight w = dL/da; ax/ay/az collect dL/dx of its own
        // position; the lists are stored (cell ids from the front of the group's
        // region, tree-order body indices from the back) while they fit
        double wxr = 0, wyr = 0, wzr = 0;
        uint32_t lc = 0, lb = 0;
        const size_t lbase = kAdj ? (size_t)grp * a.list_cap : 0;
        if constexpr (kAdj) {
            if (valid) {
                const uint32_t o = a.tperm[i];
                wxr = a.wx[o];
                wyr = a.wy[o];
                wzr = a.wz[o];
            }
        }

        __syncthreads();
        s_tgt[lane] = valid ? make_double4(x, y, z, 0.0) : make_double4(INFINITY, INFINITY, INFINITY, 0.0);
        if constexpr (kAdj) {
            s_tg[lane][0] = x; s_tg[lane][1] = y; s_tg[lane][2] = z;
            s_tg[lane][3] = wxr; s_tg[lane][4] = wyr; s_tg[lane][5] = wzr;
        }

        auto flush_cells = [&](int nn) {
            __syncthreads();
            if constexpr (kAdj) {
                if (lc + lb + (uint32_t)nn <= a.list_cap) {
                    if ((int)lane < nn) a.lists[lbase + lc + lane] = s_cid[lane];
                    lc += (uint32_t)nn;
                    cnt += (uint32_t)nn;
                    return;
                }
            }
            if ((int)lane < nn) {
                const uint32_t id = s_cid[lane];
                if (dig) hsum += mix64((((uint64_t)grp << 32) | id) ^ kDigestCell ^ salt);
                const WalkNode w = a.nodes[id];
                s_nd[lane][0] = w.cx;
                s_nd[lane][1] = w.cy;
                s_nd[lane][2] = w.cz;
                s_nd[lane][3] = w.mass;
                if (kQuad) {
                    const bool many = w.count > 1;
#pragma unroll
                    for (int u = 0; u < 6; ++u) s_q[lane][u] = many ? a.quad[6 * (size_t)id + u] : 0.0;
                }
            }
            __syncthreads();
            if constexpr (kAdj) {
                adj_self<double, kQuad>(s_nd, s_q, nn, x, y, z, wxr, wyr, wzr, eps2, ax, ay, az);
                adj_scatter<double, kQuad>(s_nd, s_q, s_cid, nn, s_tg, eps2, lane, a.gnode);
            } else {
                eval_nodes<double, kQuad, kGuard, kPot>(s_nd, s_q, nn, x, y, z, eps2, ax, ay, az, phi);
            }
            cnt += (uint32_t)nn;
        };
        auto flush_64 = [&]() {
            flush_cells(64);
            const uint32_t rest = (int)lane + 64 < nc ? s_cid[64 + lane] : 0u;
            __syncthreads();
            if ((int)lane + 64 < nc) s_cid[lane] = rest;
            nc -= 64;
            __syncthreads();
        };
        auto flush_bodies = [&]() {
            __syncthreads();
            const int e = (int)lane;
            uint32_t j = 0;
            if (e < nq) {
                int lo = 0, hi = nr - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if ((int)s_roff[mid] <= e) lo = mid; else hi = mid - 1;
                }
                j = s_rstart[lo] + (uint32_t)(e - (int)s_roff[lo]);
            }
            if constexpr (kAdj) {
                if (lc + lb + (uint32_t)nq <= a.list_cap) {
                    if (e < nq) a.lists[lbase + a.list_cap - 1u - (lb + (uint32_t)e)] = j;
                    lb += (uint32_t)nq;
                    cnt += (uint32_t)nq;
                    nr = 0;
                    nq = 0;
                    return;
                }
            }
            if (e < nq) {
                s_bd[e][0] = a.tx[j];
                s_bd[e][1] = a.ty[j];
                s_bd[e][2] = a.tz[j];
                s_bd[e][3] = a.tm[j];
                if constexpr (kAdj) s_bid[e] = j;
            }
            __syncthreads();
            if constexpr (kAdj) {
                adj_self<double>(s_bd, s_q, nq, x, y, z, wxr, wyr, wzr, eps2, ax, ay, az);
                adj_scatter<double>(s_bd, s_q, s_bid, nq, s_tg, eps2, lane, a.gbody);
            } else {
                eval_bodies<double, kGuard, kPot>(s_bd, nq, x, y, z, eps2, ax, ay, az, phi);
            }
            cnt += (uint32_t)nq;
            nr = 0;
            nq = 0;
        };
        auto add_leaf = [&](uint32_t st, uint32_t count) {
            if (dig && lane == 0) hsum += mix64((((uint64_t)grp << 32) | st) ^ kDigestLeaf ^ salt);
            for (uint32_t k = 0; k < count; k += 64) {
                const int m = (int)((count - k) < 64u ? (count - k) : 64u);
                if (nq + m > kLp64Bodies || nr == kRanges32) flush_bodies();
                if (lane == 0) {
                    s_rstart[nr] = st + k;
                    s_roff[nr] = (uint32_t)nq;
                }
                ++nr;
                nq += m;
            }
        };
        auto add_cell = [&](uint32_t id) {
            if (lane == 0) s_cid[nc] = id;
            ++nc;
            if (nc == 64) flush_64();
        };
        auto walk_subtree = [&](uint32_t p0, uint32_t end) {
            uint32_t q = p0 + 1;
            while (q < end) {
                q = __builtin_amdgcn_readfirstlane(q);
                const WalkNode nd = load_node(raw, q);
                const uint32_t nskip = nd.skip > q ? nd.skip : q + 1;
                const double dx = nd.cx - x, dy = nd.cy - y, dz = nd.cz - z;
                const double d2 = dist2(dx, dy, dz);
                const bool opened = (__builtin_amdgcn_ballot_w64(!mac_t(nd.size2, d2, th2, 0.f, 0.f)) & vmask) != 0ull;
                if (!opened) {
                    add_cell(q);
                    q = nskip;
                } else if (nd.level_leaf & 256u) {
                    add_leaf(nd.start, nd.count);
                    q = nskip;
                } else {
                    q = q + 1;
                }
            }
        };

        int top = 0;
        {
            const WalkNode root = load_node(raw, rid);
            if (root.level_leaf & 256u) {
                add_leaf(root.start, root.count);
            } else {
                const uint32_t cb = __builtin_amdgcn_readfirstlane(a.child_byte[rid]);
                if (lane < 8u && ((cb >> lane) & 1u))
                    s_stack[__popc(cb & ((1u << lane) - 1u))] = a.child[8 * (size_t)rid + lane];
                top = __popc(cb);
            }
        }
        __syncthreads();
        // Most cells need no least |d|^2 over the 64 bodies (the FP32 walk's
        // rule): with lo/hi the group's bounding box, fl(c - clamp(c, lo, hi))
        // per axis is at most |fl(c - x_b)| for every body (rounding is
        // monotone), so dist2 of it is at most every body's, and fl(s^2 / x) is
        // non-increasing in x: the test passing at the box means every body
        // accepts.  Failing at one of four sample bodies means some body opens.
        // The other cells wait in s_und for the exact test, 64 at a time.
        const double blo_x = wave_mind(valid ? x : INFINITY), bhi_x = wave_maxd(valid ? x : -INFINITY);
        const double blo_y = wave_mind(valid ? y : INFINITY), bhi_y = wave_maxd(valid ? y : -INFINITY);
        const double blo_z = wave_mind(valid ? z : INFINITY), bhi_z = wave_maxd(valid ? z : -INFINITY);
        const int nv = __popcll(vmask);        // the valid bodies are lanes 0 .. nv - 1
        int nund = 0;
        while (top > 0 || nund > 0) {
            const bool exact = nund >= 64 || top == 0;
            const int have = exact ? nund : top;
            const int k = have < 64 ? have : 64;
            const int base = have - k;
            const bool live = (int)lane < k;
            uint32_t p;
            if (exact) {
                p = live ? s_und[base + lane] : rid;
                nund = base;
            } else {
                p = live ? s_stack[base + lane] : rid;
                top = base;
            }
            const uint4 r0 = rec4[4 * (size_t)p], r1 = rec4[4 * (size_t)p + 1], r2 = rec4[4 * (size_t)p + 2];
            const double cx = __hiloint2double((int)r0.y, (int)r0.x), cy = __hiloint2double((int)r0.w, (int)r0.z);
            const double cz = __hiloint2double((int)r1.y, (int)r1.x);
            const double size2 = __hiloint2double((int)r2.y, (int)r2.x);
            const uint32_t skip = r2.z, start = r2.w;
            const uint4 r3 = rec4[4 * (size_t)p + 3];
            const uint32_t count = r3.x, ll = r3.y;
            const uint32_t cb = live ? a.child_byte[p] : 0u;
            bool pass = false, decided = live;
            if (exact) {
                double md2 = INFINITY;
#pragma unroll 8
                for (int b = 0; b < 64; ++b) {
                    const double4 t = s_tgt[b];
                    const double dx = cx - t.x, dy = cy - t.y, dz = cz - t.z;
                    md2 = fmin(md2, dist2(dx, dy, dz));
                }
                pass = live && mac_t(size2, md2, th2, 0.f, 0.f);
            } else {
                const double d2b = dist2(cx - fmin(fmax(cx, blo_x), bhi_x), cy - fmin(fmax(cy, blo_y), bhi_y),
                                         cz - fmin(fmax(cz, blo_z), bhi_z));
                bool open_sure = false;
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    const double4 t = s_tgt[(nv * h) / 4 + nv / 8];
                    open_sure = open_sure || !mac_t(size2, dist2(cx - t.x, cy - t.y, cz - t.z), th2, 0.f, 0.f);
                }
                const bool acc_sure = mac_t(size2, d2b, th2, 0.f, 0.f);
                const bool und = live && !acc_sure && !open_sure;
                const uint64_t mu = __builtin_amdgcn_ballot_w64(und);
                if (und)
                    s_und[nund + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mu >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((uint32_t)mu, 0u))] = p;
                nund += __popcll(mu);
                pass = live && acc_sure;
                decided = live && !und;
            }
            const bool leaf = (ll & 256u) != 0u;
            const bool open_in = decided && !pass && !leaf;
            const uint64_t m_acc = __builtin_amdgcn_ballot_w64(pass);
            const uint64_t m_leaf = __builtin_amdgcn_ballot_w64(decided && !pass && leaf);
            const uint32_t nch = open_in ? (uint32_t)__popc(cb) : 0u;
            const uint32_t incl = wave_incl_scan(nch, lane);
            const uint32_t room = (uint32_t)(kLpStack - top);
            const uint64_t m_fit = __builtin_amdgcn_ballot_w64(incl <= room);
            const int nfit = __popcll(m_fit);
            const bool push = open_in && ((m_fit >> lane) & 1ull);
            const uint64_t m_over = __builtin_amdgcn_ballot_w64(open_in && !push);
            const uint32_t pushed = nfit > 0 ? __shfl(incl, nfit - 1, 64) : 0u;
            __syncthreads();
            if (push) {
                const uint4 ca = *reinterpret_cast<const uint4*>(a.child + 8 * (size_t)p);
                const uint4 cbv = *reinterpret_cast<const uint4*>(a.child + 8 * (size_t)p + 4);
                const uint32_t ids[8] = {ca.x, ca.y, ca.z, ca.w, cbv.x, cbv.y, cbv.z, cbv.w};
                uint32_t w = (uint32_t)top + incl - nch;
#pragma unroll
                for (int o = 0; o < 8; ++o)
                    if ((cb >> o) & 1u) s_stack[w++] = ids[o];
            }
            top += (int)pushed;
            if (pass) s_cid[nc + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m_acc >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)m_acc, 0u))] = p;
            nc += __popcll(m_acc);
            __syncthreads();
            if (nc >= 64) flush_64();
            for (uint64_t m = m_leaf; m; m &= m - 1) {
                const int l = __builtin_ctzll(m);
                add_leaf(__shfl(start, l, 64), __shfl(count, l, 64));
            }
            for (uint64_t m = m_over; m; m &= m - 1) {
                const int l = __builtin_ctzll(m);
                const uint32_t p0 = __shfl(p, l, 64), end = __shfl(skip, l, 64);
                walk_subtree(p0, end > p0 ? end : p0 + 1);
            }
            __syncthreads();
        }
        if (nc > 0) {
            flush_cells(nc);
            nc = 0;
        }
        if (nq > 0) flush_bodies();
        if constexpr (kAdj) {
            if (lane == 0) a.list_n[grp] = make_uint2(lc, lb);
        }
        if (valid) {
            const uint32_t o = a.tperm[i];
            if constexpr (kAdj) {
                a.gself[3 * (size_t)o] = ax;
                a.gself[3 * (size_t)o + 1] = ay;
                a.gself[3 * (size_t)o + 2] = az;
            } else {
                a.ax[o] = ax;
                a.ay[o] = ay;
                a.az[o] = az;
                if constexpr (kPot) a.pot[o] = phi;
                if (a.kick) walk_kick(a, o, ax, ay, az);
            }
            if (a.counts) a.counts[o] = cnt;
        }
    }
    if (dig) {
        // wrapping sums commute: the digest does not depend on which wave walked which group
        for (int off = 32; off > 0; off >>= 1) hsum += (uint64_t)__shfl_xor((unsigned long long)hsum, off, 64);
        if (lane == 0) atomicAdd(a.digest, (unsigned long long)hsum);
    }
}

__global__ void k_fill_wq(uint32_t* __restrict__ wq) {
    if (threadIdx.x < 8) wq[threadIdx.x] = 0u;
}

// FP64 direct sum over all bodies (array order) for a list of targets; the
// same softened monopole term, j != i by index.  pot = -sum m_j / sqrt(r^2+eps^2).
__global__ __launch_bounds__(kBlock) void k_direct(const double* __restrict__ px, const double* __restrict__ py,
                                                   const double* __restrict__ pz, const double* __restrict__ m,
                                                   SysView sv, const int64_t* __restrict__ targets, size_t nt,
                                                   double eps2, double* __restrict__ acc, double* __restrict__ pot) {
#pragma clang fp contract(off)
    __shared__ double sx[kBlock], sy[kBlock], sz[kBlock], sm[kBlock];
    __shared__ unsigned long long s_lo, s_hi;
    const size_t k = blockIdx.x * (size_t)kBlock + threadIdx.x;
    const bool valid = k < nt;
    const int64_t ti = valid ? targets[k] : 0;
    const uint32_t sys = sys_of(sv, (uint32_t)ti);
    // sources: the target's own system; the block streams the union of its targets' ranges
    const size_t lo = sv.off[sys], hi = sv.off[sys + 1];
    if (threadIdx.x == 0) { s_lo = ~0ull; s_hi = 0ull; }
    __syncthreads();
    if (valid) {
        atomicMin(&s_lo, (unsigned long long)lo);
        atomicMax(&s_hi, (unsigned long long)hi);
    }
    __syncthreads();
    const size_t blo = s_lo, bhi = s_hi;
    const double x = px[ti], y = py[ti], z = pz[ti];
    double ax = 0, ay = 0, az = 0, ph = 0;
    for (size_t b = blo; b < bhi; b += kBlock) {
        const size_t j = b + threadIdx.x;
        __syncthreads();
        sx[threadIdx.x] = j < bhi ? px[j] : 0.0;
        sy[threadIdx.x] = j < bhi ? py[j] : 0.0;
        sz[threadIdx.x] = j < bhi ? pz[j] : 0.0;
        sm[threadIdx.x] = j < bhi ? m[j] : 0.0;
        __syncthreads();
        const int lim = (int)((bhi - b) < (size_t)kBlock ? (bhi - b) : (size_t)kBlock);
        for (int t = 0; t < lim; ++t) {
            const size_t jj = b + t;
            if ((int64_t)jj == ti || jj < lo || jj >= hi) continue;
            const double dx = sx[t] - x, dy = sy[t] - y, dz = sz[t] - z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            term64(ax, ay, az, dx, dy, dz, d2, sm[t], eps2, nullptr);
            ph -= sm[t] / sqrt(d2 + eps2);
        }
    }
    if (valid) {
        acc[3 * k] = ax;
        acc[3 * k + 1] = ay;
        acc[3 * k + 2] = az;
        if (pot) pot[k] = ph;
    }
}

// FP64 kinetic energy, momentum and angular momentum, two-stage reduction in a
// fixed order (deterministic).
__global__ __launch_bounds__(kBlock) void k_diag_part(const double* __restrict__ px, const double* __restrict__ py,
                                                      const double* __restrict__ pz, const double* __restrict__ vx,
                                                      const double* __restrict__ vy, const double* __restrict__ vz,
                                                      const double* __restrict__ m, size_t n, double* __restrict__ part) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const double mi = m[i], x = px[i], y = py[i], z = pz[i], u = vx[i], v = vy[i], w = vz[i];
        acc[0] += 0.5 * mi * (u * u + v * v + w * w);
        acc[1] += mi * u;
        acc[2] += mi * v;
        acc[3] += mi * w;
        acc[4] += mi * (y * w - z * v);
        acc[5] += mi * (z * u - x * w);
        acc[6] += mi * (x * v - y * u);
    }
    __shared__ double red[7][kBlock];
    for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 7; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 7) part[7 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_diag_final(const double* __restrict__ part, int nparts, double* __restrict__ out) {
    if (threadIdx.x < 7) {
        double s = 0.0;
        for (int b = 0; b < nparts; ++b) s += part[7 * b + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// Workgroups of a persistent walk: as many one-wave workgroups as the chip
// holds at once (a multiple of 8, one share per XCD), at most one per group.
// The occupancy query is made once per kernel (the kernels share one
// signature, so the cache is keyed by the kernel's address).
template <class K>
unsigned persistent_grid(K kernel, unsigned groups) {
    struct Entry { const void* k; int per_cu, ncu; };
    static Entry cache[32];
    static int used = 0;
    static std::mutex mu;
    int per_cu = 0, ncu = 0;
    {
        std::lock_guard<std::mutex> lock(mu);
        for (int e = 0; e < used; ++e)
            if (cache[e].k == (const void*)kernel) { per_cu = cache[e].per_cu; ncu = cache[e].ncu; }
        if (per_cu == 0) {
            int dev = 0, pc = 0, nc = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, kernel, 64, 0) != hipSuccess ||
                hipDeviceGetAttribute(&nc, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                pc = nc = 0;
            ncu = nc > 0 ? nc : 256;
            per_cu = pc > 0 ? pc : 1;
            if (used < 32) cache[used++] = Entry{(const void*)kernel, per_cu, ncu};
        }
    }
    unsigned g = (unsigned)(per_cu * ncu);
    g = (g + 7u) & ~7u;
    const unsigned need = 8u * ((groups + 7u) / 8u);
    return g < need ? g : need;
}

#define BHX_LP32S(Q, G, P, S)                                                                                  \
    hipLaunchKernelGGL((k_walk_lp32<Q, G, P, S>), dim3(persistent_grid(k_walk_lp32<Q, G, P, S>, groups)), dim3(64), \
                       0, s, a)
#define BHX_LP32(Q, G, P)                                                 \
    do {                                                                   \
        hipLaunchKernelGGL(k_fill_wq, dim3(1), dim3(64), 0, s, a.wq);      \
        if (sub == 4) BHX_LP32S(Q, G, P, 4);                               \
        else if (sub == 2) BHX_LP32S(Q, G, P, 2);                          \
        else BHX_LP32S(Q, G, P, 1);                                        \
    } while (0)

}  // namespace

hipError_t launch_walk(bhx_ctx* c, hipStream_t s) {
    return launch_walk_to(c, s, c->cfg.group_walk != 0, c->cfg.precision, c->ax, c->ay, c->az, c->wcount);
}

// The walk and the frame's closing half kick: the group walks kick each body
// as they write its acceleration; after a per-body walk the kick kernel runs.
hipError_t launch_walk_kick(bhx_ctx* c, double dt, hipStream_t s, uint32_t* nonfinite) {
    if (c->cfg.group_walk == 0) {
        hipError_t e = launch_walk(c, s);
        return e != hipSuccess ? e : launch_kick(c, dt, s, nonfinite);
    }
    WalkKick k{dt, nonfinite};
    return launch_walk_to(c, s, true, c->cfg.precision, c->ax, c->ay, c->az, c->wcount, &k);
}

hipError_t launch_walk_to(bhx_ctx* c, hipStream_t s, bool group, int precision, double* ox, double* oy, double* oz,
                          uint32_t* counts, const WalkKick* kick, uint32_t* qcounts) {
    if (c->n == 0) return hipSuccess;
    if (kick && !group) return hipErrorInvalidValue;   // only the group walks kick
    WalkArgs a{};
    a.qcounts = qcounts;
    if (kick) {
        a.kick = c->cfg.compensated ? 2u : 1u;
        a.kdt = kick->dt;
        a.kvx = c->vx; a.kvy = c->vy; a.kvz = c->vz;
        a.kcx = c->cvx; a.kcy = c->cvy; a.kcz = c->cvz;
        a.knonfinite = kick->nonfinite;
    }
    a.nodes = c->tree.walk;
    a.nodes32 = c->tree.walk32;
    a.pay32 = c->tree.pay32;
    a.scale32 = c->sys_scale;
    a.com = c->tree.com;
    a.tb4 = c->tb4;
    a.tleaf = c->tleaf;
    a.groups = c->ngroups;
    a.gtab = c->gtab;
    a.sys_root = c->sys_root;
    a.abort_flag = &c->ctr->node_overflow;
    a.b4 = c->bounds_slot(c->tree_slot);
    a.quad = c->tree.quad;
    a.quadf = c->tree.quadf;
    a.tx = c->tx; a.ty = c->ty; a.tz = c->tz; a.tm = c->tm;
    a.tperm = c->tperm;
    a.n = (uint32_t)c->n;
    a.quad_on = (uint32_t)c->cfg.quadrupole;
    a.theta2 = c->cfg.theta * c->cfg.theta;
    a.eps2 = c->cfg.eps * c->cfg.eps;
    a.theta2f = (float)a.theta2;
    a.eps2f = (float)a.eps2;
    a.tscale = c->tree_tscale;
    a.ax = ox; a.ay = oy; a.az = oz;
    a.counts = counts;
    a.wq = c->wq;
    a.child = c->tree.child;
    a.child_byte = c->tree.child_byte;
    a.digest = c->digest;
    const int sub = c->walk_sub;
    const unsigned groups = c->ngroups;
    const unsigned grid = (groups + kWaves - 1) / kWaves;
    const bool quad = c->cfg.quadrupole != 0;
    if (group) {
        const bool guard = !(a.eps2 > 0.0);
#define BHX_GROUP(T, Q, G)                                                                                   \
    do {                                                                                                      \
        hipLaunchKernelGGL(k_fill_wq, dim3(1), dim3(64), 0, s, a.wq);                                         \
        hipLaunchKernelGGL((k_walk_lp64<Q, G>), dim3(persistent_grid(k_walk_lp64<Q, G>, groups)), dim3(64), 0, s, \
                           a);                                                                                \
    } while (0)
        if (precision == BHX_FP32) {
            if (quad) { if (guard) BHX_LP32(true, true, false); else BHX_LP32(true, false, false); }
            else { if (guard) BHX_LP32(false, true, false); else BHX_LP32(false, false, false); }
        } else {
            if (quad) { if (guard) BHX_GROUP(double, true, true); else BHX_GROUP(double, true, false); }
            else { if (guard) BHX_GROUP(double, false, true); else BHX_GROUP(double, false, false); }
        }
#undef BHX_GROUP
    } else if (precision == BHX_FP32) {
        if (quad) hipLaunchKernelGGL(k_walk32<true>, dim3(grid), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL(k_walk32<false>, dim3(grid), dim3(kBlock), 0, s, a);
    } else {
        if (quad) hipLaunchKernelGGL(k_walk64<true>, dim3(grid), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL(k_walk64<false>, dim3(grid), dim3(kBlock), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_walk_potential(bhx_ctx* c, hipStream_t s, double* ox, double* oy, double* oz, double* pot) {
    if (c->n == 0) return hipSuccess;
    WalkArgs a{};
    a.nodes = c->tree.walk;
    a.nodes32 = c->tree.walk32;
    a.pay32 = c->tree.pay32;
    a.scale32 = c->sys_scale;
    a.com = c->tree.com;
    a.tb4 = c->tb4;
    a.tleaf = c->tleaf;
    a.groups = c->ngroups;
    a.gtab = c->gtab;
    a.sys_root = c->sys_root;
    a.abort_flag = &c->ctr->node_overflow;
    a.b4 = c->bounds_slot(c->tree_slot);
    a.quad = c->tree.quad;
    a.quadf = c->tree.quadf;
    a.tx = c->tx; a.ty = c->ty; a.tz = c->tz; a.tm = c->tm;
    a.tperm = c->tperm;
    a.n = (uint32_t)c->n;
    a.quad_on = (uint32_t)c->cfg.quadrupole;
    a.theta2 = c->cfg.theta * c->cfg.theta;
    a.eps2 = c->cfg.eps * c->cfg.eps;
    a.theta2f = (float)a.theta2;
    a.eps2f = (float)a.eps2;
    a.tscale = c->tree_tscale;
    a.ax = ox; a.ay = oy; a.az = oz;
    a.pot = pot;
    a.wq = c->wq;
    a.child = c->tree.child;
    a.child_byte = c->tree.child_byte;
    const int sub = c->walk_sub;
    const unsigned groups = c->ngroups;
    const bool quad = c->cfg.quadrupole != 0, guard = !(a.eps2 > 0.0);
    if (c->cfg.precision == BHX_FP32) {
        if (quad) { if (guard) BHX_LP32(true, true, true); else BHX_LP32(true, false, true); }
        else { if (guard) BHX_LP32(false, true, true); else BHX_LP32(false, false, true); }
    } else {
#define BHX_POT64(Q, G)                                                                                        \
    do {                                                                                                        \
        hipLaunchKernelGGL(k_fill_wq, dim3(1), dim3(64), 0, s, a.wq);                                           \
        hipLaunchKernelGGL((k_walk_lp64<Q, G, true>), dim3(persistent_grid(k_walk_lp64<Q, G, true>, groups)),   \
                           dim3(64), 0, s, a);                                                                  \
    } while (0)
        if (quad) { if (guard) BHX_POT64(true, true); else BHX_POT64(true, false); }
        else { if (guard) BHX_POT64(false, true); else BHX_POT64(false, false); }
#undef BHX_POT64
    }
    return hipGetLastError();
}

hipError_t launch_walk_adjoint(bhx_ctx* c, const WalkTree& t, int precision, const double* wx, const double* wy,
                               const double* wz, double* gself, double* gnode, double* gbody, const AdjLists* L,
                               hipStream_t s) {
    if (c->n == 0 || c->ngroups == 0) return hipSuccess;
    WalkArgs a{};
    a.nodes = t.walk;
    a.nodes32 = t.walk32;
    a.pay32 = t.pay32;
    a.scale32 = t.scale32;
    a.com = t.com;
    a.tb4 = t.tb4;
    a.tleaf = t.tleaf;
    a.groups = c->ngroups;
    a.gtab = c->gtab;
    a.sys_root = t.sys_root;
    a.abort_flag = &c->ctr->pad[0];   // never written: stored trees are complete
    a.b4 = t.b4;
    a.tx = t.tx; a.ty = t.ty; a.tz = t.tz; a.tm = t.tm;
    a.tperm = t.tperm;
    a.n = (uint32_t)c->n;
    a.theta2 = c->cfg.theta * c->cfg.theta;
    a.eps2 = c->cfg.eps * c->cfg.eps;
    a.theta2f = (float)a.theta2;
    a.eps2f = (float)a.eps2;
    a.tscale = c->tree_tscale;
    a.wx = wx; a.wy = wy; a.wz = wz;
    a.gself = gself;
    a.gnode = gnode;
    a.gbody = gbody;
    a.quad = t.quad;
    a.child = t.child;
    a.child_byte = t.child_byte;
    // the adjoint stream runs beside the main stream's walks: counters of its own
    a.wq = c->wq + 8;
    if (!L) return hipErrorInvalidValue;
    {
        a.lists = L->lists;
        a.list_n = L->list_n;
        a.list_cap = L->cap;
        a.list_off = L->off;
        a.part = L->part;
        a.pkey = L->key;
        a.pval = L->val;
        a.node_cap = L->node_cap;
    }
    const unsigned groups = c->ngroups;
    const unsigned bgrid = 8u * ((groups + 7u) / 8u);
    const unsigned rgrid = L ? (unsigned)std::min<size_t>((L->max_pairs + 255) / 256, 8192) : 1u;
    const unsigned kgrid = L ? (unsigned)std::min<size_t>((L->nkeys + 255) / 256, 4096) : 1u;
    const unsigned wgrid = L ? (unsigned)std::min<size_t>((L->nkeys + 3) / 4, 65536) : 1u;
    hipError_t e = hipSuccess;
    const bool guard = !(a.eps2 > 0.0);
    const bool quad = c->cfg.quadrupole != 0;
#define BHX_ADJ(Q, G)                                                                                             \
    do {                                                                                                          \
        hipLaunchKernelGGL(k_fill_wq, dim3(1), dim3(64), 0, s, a.wq);                                             \
        if (precision == BHX_FP32)                                                                                \
            hipLaunchKernelGGL((k_walk_lp32<Q, G, false, 1, true>),                                               \
                               dim3(persistent_grid(k_walk_lp32<Q, G, false, 1, true>, groups)), dim3(64), 0, s, a); \
        else                                                                                                      \
            hipLaunchKernelGGL((k_walk_lp64<Q, G, false, true>),                                                  \
                               dim3(persistent_grid(k_walk_lp64<Q, G, false, true>, groups)), dim3(64), 0, s, a); \
        hipLaunchKernelGGL(k_list_scan, dim3(1), dim3(1024), 0, s, a.list_n, groups, L->off);                     \
        if (precision == BHX_FP32) {                                                                              \
            if (groups < 8192u)                                                                                   \
                hipLaunchKernelGGL((k_adj_lists<float, Q, 8>), dim3(bgrid), dim3(512), 0, s, a);                  \
            else                                                                                                  \
                hipLaunchKernelGGL((k_adj_lists<float, Q, 4>), dim3(bgrid), dim3(256), 0, s, a);                  \
        } else {                                                                                                  \
            hipLaunchKernelGGL((k_adj_lists<double, Q, 4>), dim3(bgrid), dim3(256), 0, s, a);                     \
        }                                                                                                         \
        {                                                                                                         \
            e = radix_sort_pairs(L->key, L->val, L->key_alt, L->val_alt, L->max_pairs, L->hist, L->status,       \
                                 L->tile_ctr, c->ctr, s, L->passes, L->off + groups);                             \
            if (e != hipSuccess) return e;                                                                        \
            hipLaunchKernelGGL(k_fill_runs, dim3(kgrid), dim3(256), 0, s, L->run, L->nkeys);                      \
            hipLaunchKernelGGL(k_run_bounds, dim3(rgrid), dim3(256), 0, s, (L->passes & 1) ? L->key_alt : L->key,  \
                               L->off + groups, L->run);                                                          \
            hipLaunchKernelGGL((k_run_sum<Q ? 10 : 4>), dim3(wgrid), dim3(256), 0, s, L->run,                     \
                               (L->passes & 1) ? L->val_alt : L->val, L->part, L->node_cap, L->nkeys, gnode,       \
                               gbody);                                                                            \
        }                                                                                                         \
    } while (0)
    if (quad) { if (guard) BHX_ADJ(true, true); else BHX_ADJ(true, false); }
    else { if (guard) BHX_ADJ(false, true); else BHX_ADJ(false, false); }
#undef BHX_ADJ
    return hipGetLastError();
}

hipError_t launch_direct(bhx_ctx* c, const int64_t* targets, size_t nt, double eps, double* acc, double* pot,
                         hipStream_t s) {
    const unsigned grid = (unsigned)((nt + kBlock - 1) / kBlock);
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_direct, dim3(grid), dim3(kBlock), 0, s, c->px, c->py, c->pz, c->m, c->sys(), targets, nt,
                       eps * eps, acc, pot);
    return hipGetLastError();
}

hipError_t launch_diag(bhx_ctx* c, double* out7, hipStream_t s) {
    // part buffer: reuse the staging area (needs 7 * parts doubles)
    const int parts = 256;
    double* part = c->stage;
    hipLaunchKernelGGL(k_diag_part, dim3(parts), dim3(kBlock), 0, s, c->px, c->py, c->pz, c->vx, c->vy, c->vz, c->m,
                       c->n, part);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, part, parts, out7);
    return hipGetLastError();
}

}  // namespace bhx
