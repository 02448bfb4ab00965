// bhx_grad.hip — reverse-mode differentiation of a recorded trajectory.
//
// No reference counterpart: the reference integrates forward only.  The
// forward frames are the engine's own (bhx_step order, replayed from the
// captured frame graphs).  A frame maps (x, v, a) to
//     sort:   x', v', a', m' = gather by the Morton permutation P
//     kick:   h = v' + a' dt/2
//     drift:  y = x' + h dt
//     walk:   b = A(y, m')            (tree and interaction lists of y)
//     kick:   u = h + b dt/2
// and its adjoint, for incoming (Y, U, B) = dL/d(y, u, b), is
//     w = B + U dt/2                  (weights of the walk's output)
//     Yt = Y + (dA/dy)^T w,  Mt += (dA/dm')^T w     (adjoint walk, frozen lists)
//     H = U + Yt dt
//     (X, V, Ab, M) = scatter by P of (Yt, H, H dt/2, Mt).
// The walk's decisions (tree, lists, orders) are frozen at those of the
// forward frame: the backward pass re-runs the frames of each segment from a
// checkpoint — bitwise the same frames — and keeps what the adjoint walk
// reads (body order, tree, node moments) in one store slot per frame.  The
// cells' centres of mass and masses are functions of their bodies, so a cell
// term's adjoint (dL/dX_c, dL/dM_c) is pushed down to every body of the cell:
//     dL/dx_k += (m_k / M_c) dL/dX_c
//     dL/dm_k += dL/dM_c + dL/dX_c . (x_k - X_c) / M_c.
// With quadrupoles, a cell of more than one body also carries G = dL/dQ_c
// (symmetric), and Q_c = sum m_k (3 s_k s_k^T - |s_k|^2 I), s_k = x_k - X_c,
// gives
//     dL/dx_k += m_k (6 G s_k - 2 tr(G) s_k)
//     dL/dm_k += 3 s_k . G s_k - tr(G) |s_k|^2
// (the path through X_c vanishes: sum m_k s_k = 0).
// Compensated summation in the leapfrog is treated as exact arithmetic.
#include <algorithm>
#include <chrono>
#include <vector>

#include "bhx_internal.hpp"

using namespace bhx;

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kMaxJobs = 20;

unsigned grid_of(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    if (g == 0) g = 1;
    if (g > 65535u * 16u) g = 65535u * 16u;
    return (unsigned)g;
}

struct CopyJob {
    const uint32_t* src;
    uint32_t* dst;
    size_t words;
};
struct CopyJobs {
    CopyJob job[kMaxJobs];
    int n;
};

// Several device-to-device copies in one launch.
__global__ __launch_bounds__(kBlock) void k_copy_jobs(CopyJobs J) {
    const size_t stride = (size_t)gridDim.x * kBlock;
    const size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x;
    for (int k = 0; k < J.n; ++k) {
        const CopyJob j = J.job[k];
        for (size_t w = t; w < j.words; w += stride) j.dst[w] = j.src[w];
    }
}

__global__ __launch_bounds__(kBlock) void k_iota32(uint32_t* __restrict__ v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) v[i] = (uint32_t)i;
}

// Which starting body each slot holds after one more frame: out[i] = in[perm[i]].
__global__ __launch_bounds__(kBlock) void k_compose(const uint32_t* __restrict__ in, const uint32_t* __restrict__ perm,
                                                    size_t n, uint32_t* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) out[i] = in[perm[i]];
}

__global__ void k_zero_flags(Counters* __restrict__ ctr) {
    if (threadIdx.x == 0) {
        ctr->max_nodes = 0;
        ctr->overflow_seen = 0;
        ctr->nonfinite = 0;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_d(double* __restrict__ p, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) p[i] = v;
}

// Caller's per-body vectors (starting order, xyz interleaved) into slot order.
__global__ __launch_bounds__(kBlock) void k_from_start(const double* __restrict__ in, const uint32_t* __restrict__ ident,
                                                       size_t n, double* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const size_t j = ident[i];
        out[3 * i] = in[3 * j];
        out[3 * i + 1] = in[3 * j + 1];
        out[3 * i + 2] = in[3 * j + 2];
    }
}

// w = B + U * hdt: the walk output's weights (hdt = dt/2: the second half-kick's share).
__global__ __launch_bounds__(kBlock) void k_adj_weights(const double* __restrict__ B, const double* __restrict__ U,
                                                        size_t n, double hdt, double* __restrict__ wx,
                                                        double* __restrict__ wy, double* __restrict__ wz) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        wx[i] = B[3 * i] + U[3 * i] * hdt;
        wy[i] = B[3 * i + 1] + U[3 * i + 1] * hdt;
        wz[i] = B[3 * i + 2] + U[3 * i + 2] * hdt;
    }
}

// Cell adjoints pushed down to the bodies (tree order j), written in array
// order.  kQuad: gnode holds 10 per node and the cells' body counts come
// from the walk records (walk or walk32, whichever the slot holds); a cell of
// one body has no quadrupole.
template <bool kQuad>
__global__ __launch_bounds__(kBlock) void k_adj_push(const uint32_t* __restrict__ tleaf, const uint32_t* __restrict__ parent,
                                                     const double* __restrict__ mass, const double* __restrict__ com,
                                                     const double* __restrict__ gnode, const double* __restrict__ gbody,
                                                     const double* __restrict__ tx, const double* __restrict__ ty,
                                                     const double* __restrict__ tz, const double* __restrict__ tm,
                                                     const uint32_t* __restrict__ tperm, size_t n,
                                                     const WalkNode* __restrict__ walk,
                                                     const WalkVisit32* __restrict__ walk32,
                                                     double* __restrict__ push) {
    constexpr int gs = kQuad ? 10 : 4;
    for (size_t j = blockIdx.x * (size_t)kBlock + threadIdx.x; j < n; j += (size_t)gridDim.x * kBlock) {
        double ux = 0.0, uy = 0.0, uz = 0.0, sm = 0.0;
        double qx = 0.0, qy = 0.0, qz = 0.0;
        const double x = tx[j], y = ty[j], z = tz[j];
        for (uint32_t c = tleaf[j]; c != kNone; c = parent[c]) {
            const double M = mass[c];
            if (!(M > 0.0)) continue;
            const double* g = gnode + gs * (size_t)c;
            const double* X = com + 3 * (size_t)c;
            const double iM = 1.0 / M;
            ux += g[0] * iM;
            uy += g[1] * iM;
            uz += g[2] * iM;
            sm += g[3] - (g[0] * X[0] + g[1] * X[1] + g[2] * X[2]) * iM;
            if constexpr (kQuad) {
                const uint32_t cnt = walk ? walk[c].count : (walk32[c].meta >> 6);
                if (cnt > 1u) {
                    const double* G = g + 4;
                    const double sx = x - X[0], sy = y - X[1], sz = z - X[2];
                    const double gx = G[0] * sx + G[1] * sy + G[2] * sz;
                    const double gy = G[1] * sx + G[3] * sy + G[4] * sz;
                    const double gz = G[2] * sx + G[4] * sy + G[5] * sz;
                    const double tr = G[0] + G[3] + G[5];
                    qx += 6.0 * gx - 2.0 * tr * sx;
                    qy += 6.0 * gy - 2.0 * tr * sy;
                    qz += 6.0 * gz - 2.0 * tr * sz;
                    sm += 3.0 * (sx * gx + sy * gy + sz * gz) - tr * (sx * sx + sy * sy + sz * sz);
                }
            }
        }
        const double* gb = gbody + 4 * j;
        const double m = tm[j];
        const size_t o = tperm[j];
        push[4 * o] = m * (ux + qx) + gb[0];
        push[4 * o + 1] = m * (uy + qy) + gb[1];
        push[4 * o + 2] = m * (uz + qz) + gb[2];
        push[4 * o + 3] = sm + x * ux + y * uy + z * uz + gb[3];
    }
}

// Back through the drift, the first half-kick and the frame's body sort.
__global__ __launch_bounds__(kBlock) void k_adj_kdk(const double* __restrict__ Y, const double* __restrict__ U,
                                                    const double* __restrict__ Mt, const double* __restrict__ gself,
                                                    const double* __restrict__ push, const uint32_t* __restrict__ perm,
                                                    size_t n, double dt, double* __restrict__ X, double* __restrict__ V,
                                                    double* __restrict__ Ab, double* __restrict__ Mo) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const size_t j = perm[i];
        for (int a = 0; a < 3; ++a) {
            const double yt = Y[3 * i + a] + gself[3 * i + a] + push[4 * i + a];
            const double h = U[3 * i + a] + yt * dt;
            X[3 * j + a] = yt;
            V[3 * j + a] = h;
            Ab[3 * j + a] = h * (0.5 * dt);
        }
        Mo[j] = Mt[i] + push[4 * i + 3];
    }
}

// The initial walk a_0 = A(x_0, m): its weights are dL/da_0.
__global__ __launch_bounds__(kBlock) void k_adj_init(double* __restrict__ X, double* __restrict__ Mb,
                                                     const double* __restrict__ gself, const double* __restrict__ push,
                                                     size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        for (int a = 0; a < 3; ++a) X[3 * i + a] += gself[3 * i + a] + push[4 * i + a];
        Mb[i] += push[4 * i + 3];
    }
}

// Decision digest, tree half: every body's starting identity, tree slot, leaf
// id and leaf key hashed with the frame index, summed (wrapping) into
// digest[0]; then the frame index the next walk mixes in.  ident maps array
// slots to starting bodies (null: the array order is the starting order).
// The leaf keys fix the node set (every node is a prefix of a leaf key) and
// the tree slots the groups, so equal sums mean equal trees and groups.  A
// build that ran out of nodes left no tree (the recording repeats with more
// room): nothing is read then, and ids out of range are never followed.
constexpr uint64_t kDigestTree = 0xa4093822299f31d0ull, kDigestFrame = 0x9e3779b97f4a7c15ull;
__global__ __launch_bounds__(kBlock) void k_tree_digest(const uint32_t* __restrict__ tperm,
                                                        const uint32_t* __restrict__ tleaf,
                                                        const uint64_t* __restrict__ key,
                                                        const uint32_t* __restrict__ ident, size_t n, uint64_t frame,
                                                        const Counters* __restrict__ ctr, size_t cap,
                                                        unsigned long long* __restrict__ digest) {
    if (ctr->node_overflow) return;
    uint64_t h = 0;
    for (size_t j = blockIdx.x * (size_t)kBlock + threadIdx.x; j < n; j += (size_t)gridDim.x * kBlock) {
        const uint32_t leaf = tleaf[j];
        const uint32_t o = tperm[j];
        const uint64_t body = ident ? (o < n ? ident[o] : 0xFFFFFFFFull) : o;
        const uint64_t lk = leaf < cap ? key[leaf] : 0ull;
        h += mix64(mix64(lk ^ kDigestTree) + ((body << 32) | (uint64_t)j) + (uint64_t)leaf * 0x632be59bd9b4e019ull +
                   frame * kDigestFrame);
    }
    for (int off = 32; off > 0; off >>= 1) h += (uint64_t)__shfl_xor((unsigned long long)h, off, 64);
    if ((threadIdx.x & 63u) == 0) atomicAdd(digest, (unsigned long long)h);
    if (blockIdx.x == 0 && threadIdx.x == 0) digest[2] = frame + 1;
}

__global__ void k_digest_set(unsigned long long* __restrict__ digest, unsigned long long on, unsigned long long frame) {
    if (threadIdx.x == 0) {
        digest[0] = 0ull;
        digest[1] = on;
        digest[2] = frame;
    }
}

}  // namespace

// What the adjoint walk of one frame reads.
struct GradSlot {
    uint32_t* perm = nullptr;                  // the frame's body sort
    double *tx = nullptr, *ty = nullptr, *tz = nullptr, *tm = nullptr;
    uint32_t* tperm = nullptr;
    float4* tb4 = nullptr;
    uint32_t* tleaf = nullptr;
    WalkNode* walk = nullptr;
    WalkVisit32* walk32 = nullptr;
    WalkPay32* pay32 = nullptr;
    uint32_t* parent = nullptr;
    double* com = nullptr;
    double* mass = nullptr;
    uint32_t* sys_root = nullptr;
    double* b4 = nullptr;
    double* scale32 = nullptr;
    double* quad = nullptr;                    // FP64 walk with quadrupoles: 6 per node
    uint32_t* child = nullptr;                 // octant table, 8 per node
    uint32_t* child_byte = nullptr;            // octant mask, 1 per node
};

struct GradRecord {
    int nsteps = 0, every = 0, init_acc = 0, precision = 0, quad = 0;
    double dt = 0.0;
    size_t n = 0, nsys = 0, state_len = 0;
    size_t node_cap = 0;                       // node records per store slot
    std::vector<double*> ckpt;                 // state at frames 0, every, 2 every, ...
    double* final_state = nullptr;
    uint32_t* ident = nullptr;                 // slot -> starting body, after the last frame
    uint32_t* ident_tmp = nullptr;
    void* init_store = nullptr;                // the initial walk's tree
    GradSlot init_slot;
    void* seg_store = nullptr;                 // two banks of `every` frame slots
    std::vector<GradSlot> slots;
    void* all_store = nullptr;                 // every recorded frame's slot, when it fits the budget
    std::vector<GradSlot> all_slots;
    size_t all_nodes = 0;                      // node records per slot of all_store
    // the reverse pass re-runs segment k-1 on the main stream while the
    // adjoint of segment k runs on the context's adjoint stream
    bool events = false;
    hipEvent_t ev_stored[2] = {}, ev_consumed[2] = {}, ev_ready = nullptr;
    double* work = nullptr;                    // adjoint vectors
    size_t work_nodes = 0;
    bhx::AdjLists lists;                       // FP32 adjoint: stored group lists and their evaluation
    void* lists_mem = nullptr;
    double record_ms = 0.0, adjoint_ms = 0.0;
};

namespace bhx {

void free_grad(bhx_ctx* c) {
    GradRecord* g = c->grad;
    if (!g) return;
    for (double* p : g->ckpt)
        if (p) (void)hipFree(p);
    if (c->s_adj) (void)hipStreamSynchronize(c->s_adj);
    void* ptrs[] = {g->final_state, g->ident, g->ident_tmp, g->init_store, g->seg_store, g->work, g->lists_mem,
                    g->all_store};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : {g->ev_stored[0], g->ev_stored[1], g->ev_consumed[0], g->ev_consumed[1], g->ev_ready})
        if (e) (void)hipEventDestroy(e);
    delete g;
    c->grad = nullptr;
}

}  // namespace bhx

namespace {

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// FP64 quadrupoles ride along only for the FP64 walk (the FP32 payload holds its own).
bool slot_quad(int precision, int quad) { return quad && precision != BHX_FP32; }

size_t slot_bytes(size_t n, size_t nodes, size_t nsys, bool quad) {
    return align_up(4 * n) + 4 * align_up(8 * n) + align_up(4 * n) + align_up(16 * n) + align_up(4 * n) +
           align_up(64 * nodes) + align_up(4 * nodes) + align_up(24 * nodes) + align_up(8 * nodes) +
           align_up(4 * nsys) + align_up(32 * nsys) + align_up(16 * nsys) + (quad ? align_up(48 * nodes) : 0) +
           align_up(32 * nodes) + align_up(4 * nodes);
}

GradSlot carve_slot(char*& p, size_t n, size_t nodes, size_t nsys, int precision, bool quad) {
    GradSlot s;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += align_up(bytes);
        return r;
    };
    s.perm = (uint32_t*)take(4 * n);
    s.tx = (double*)take(8 * n);
    s.ty = (double*)take(8 * n);
    s.tz = (double*)take(8 * n);
    s.tm = (double*)take(8 * n);
    s.tperm = (uint32_t*)take(4 * n);
    s.tb4 = (float4*)take(16 * n);
    s.tleaf = (uint32_t*)take(4 * n);
    char* w = take(64 * nodes);
    if (precision == BHX_FP32) {
        s.walk32 = (WalkVisit32*)w;
        s.pay32 = (WalkPay32*)(w + 32 * nodes);
    } else {
        s.walk = (WalkNode*)w;
    }
    s.parent = (uint32_t*)take(4 * nodes);
    s.com = (double*)take(24 * nodes);
    s.mass = (double*)take(8 * nodes);
    s.sys_root = (uint32_t*)take(4 * nsys);
    s.b4 = (double*)take(32 * nsys);
    s.scale32 = (double*)take(16 * nsys);
    if (quad) s.quad = (double*)take(48 * nodes);
    s.child = (uint32_t*)take(32 * nodes);
    s.child_byte = (uint32_t*)take(4 * nodes);
    return s;
}

void add_job(CopyJobs& J, const void* src, void* dst, size_t bytes) {
    J.job[J.n].src = (const uint32_t*)src;
    J.job[J.n].dst = (uint32_t*)dst;
    J.job[J.n].words = bytes / 4;
    ++J.n;
}

hipError_t run_jobs(const CopyJobs& J, size_t max_words, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_jobs, dim3(grid_of(max_words < 1024 * 1024 ? max_words : 1024 * 1024)), dim3(kBlock), 0,
                       s, J);
    return hipGetLastError();
}

double** state_arrays(bhx_ctx* c, double** a) {
    double* v[16] = {c->px, c->py, c->pz, c->vx, c->vy, c->vz, c->m, c->cpx, c->cpy, c->cpz,
                     c->cvx, c->cvy, c->cvz, c->ax, c->ay, c->az};
    for (int k = 0; k < 16; ++k) a[k] = v[k];
    return a;
}

// The body state (16 arrays) and the bounds the next frame starts from, into
// / out of `buf`.  The bounds go to whichever slot is current when loading:
// the host's slot parity may differ from the one at saving.
int save_state(bhx_ctx* c, double* buf, hipStream_t s) {
    double* a[16];
    state_arrays(c, a);
    CopyJobs J{};
    for (int k = 0; k < 16; ++k) add_job(J, a[k], buf + k * c->n, 8 * c->n);
    add_job(J, c->bounds_slot(c->bslot), buf + 16 * c->n, 8 * 4 * c->nsys);
    HIPCHK(run_jobs(J, 2 * c->n, s));
    return BHX_OK;
}

int load_state(bhx_ctx* c, const double* buf, hipStream_t s) {
    double* a[16];
    state_arrays(c, a);
    CopyJobs J{};
    for (int k = 0; k < 16; ++k) add_job(J, buf + k * c->n, a[k], 8 * c->n);
    add_job(J, buf + 16 * c->n, c->bounds_slot(c->bslot), 8 * 4 * c->nsys);
    HIPCHK(run_jobs(J, 2 * c->n, s));
    c->bounds_valid = true;
    c->tree_valid = c->tree_built = false;
    return BHX_OK;
}

// What the adjoint walk of the frame just run (or of the tree just built) reads.
int store_frame(bhx_ctx* c, const GradSlot& d, size_t nodes, int precision, bool with_perm, hipStream_t s) {
    const size_t n = c->n;
    CopyJobs J{};
    if (with_perm) add_job(J, c->vals, d.perm, 4 * n);
    add_job(J, c->tx, d.tx, 8 * n);
    add_job(J, c->ty, d.ty, 8 * n);
    add_job(J, c->tz, d.tz, 8 * n);
    add_job(J, c->tm, d.tm, 8 * n);
    add_job(J, c->tperm, d.tperm, 4 * n);
    add_job(J, c->tb4, d.tb4, 16 * n);
    add_job(J, c->tleaf, d.tleaf, 4 * n);
    if (precision == BHX_FP32) {
        add_job(J, c->tree.walk32, d.walk32, 32 * nodes);
        add_job(J, c->tree.pay32, d.pay32, 32 * nodes);
    } else {
        add_job(J, c->tree.walk, d.walk, 64 * nodes);
    }
    add_job(J, c->tree.parent, d.parent, 4 * nodes);
    add_job(J, c->tree.com, d.com, 24 * nodes);
    add_job(J, c->tree.mass, d.mass, 8 * nodes);
    add_job(J, c->sys_root, d.sys_root, 4 * c->nsys);
    add_job(J, c->bounds_slot(c->tree_slot), d.b4, 32 * c->nsys);
    add_job(J, c->sys_scale, d.scale32, 16 * c->nsys);
    if (d.quad) add_job(J, c->tree.quad, d.quad, 48 * nodes);
    add_job(J, c->tree.child, d.child, 32 * nodes);
    add_job(J, c->tree.child_byte, d.child_byte, 4 * nodes);
    HIPCHK(run_jobs(J, 16 * (n > nodes ? n : nodes), s));
    return BHX_OK;
}

WalkTree tree_of(const GradSlot& d) {
    WalkTree t;
    t.walk = d.walk;
    t.walk32 = d.walk32;
    t.pay32 = d.pay32;
    t.scale32 = d.scale32;
    t.com = d.com;
    t.tb4 = d.tb4;
    t.tleaf = d.tleaf;
    t.tx = d.tx; t.ty = d.ty; t.tz = d.tz; t.tm = d.tm;
    t.tperm = d.tperm;
    t.sys_root = d.sys_root;
    t.b4 = d.b4;
    t.quad = d.quad;
    t.child = d.child;
    t.child_byte = d.child_byte;
    return t;
}

// Adjoint vectors, all per body in slot order (xyz interleaved): two sets of
// (X, V, A, M), the walk weights, the own-position term, the pushed-down
// cell terms, and the per-entry accumulators.
struct Work {
    double *X[2], *V[2], *A[2], *M[2];
    double *wx, *wy, *wz, *gself, *push, *gbody, *gnode;
};

Work carve_work(double* w, size_t n, size_t nodes) {
    Work k;
    for (int b = 0; b < 2; ++b) {
        k.X[b] = w; w += 3 * n;
        k.V[b] = w; w += 3 * n;
        k.A[b] = w; w += 3 * n;
        k.M[b] = w; w += n;
    }
    k.wx = w; w += n;
    k.wy = w; w += n;
    k.wz = w; w += n;
    k.gself = w; w += 3 * n;
    k.push = w; w += 4 * n;
    k.gbody = w; w += 4 * n;
    k.gnode = w;
    (void)nodes;
    return k;
}

size_t work_len(size_t n, size_t nodes, bool quad) {
    return 2 * 10 * n + 3 * n + 3 * n + 4 * n + 4 * n + (quad ? 10 : 4) * nodes;
}

// One adjoint walk on stored tree d with weights (B + U hdt); own-position
// terms into k.gself, cell and body terms pushed down into k.push.
int adjoint_walk(bhx_ctx* c, GradRecord* g, const Work& k, const GradSlot& d, const double* B, const double* U,
                 double hdt, hipStream_t s) {
    const size_t n = c->n;
    hipLaunchKernelGGL(k_adj_weights, dim3(grid_of(n)), dim3(kBlock), 0, s, B, U, n, hdt, k.wx, k.wy, k.wz);
    hipLaunchKernelGGL(k_fill_d, dim3(grid_of(4 * n)), dim3(kBlock), 0, s, k.gbody, 4 * n, 0.0);
    const size_t gs = g->quad ? 10 : 4;
    hipLaunchKernelGGL(k_fill_d, dim3(grid_of(gs * g->work_nodes)), dim3(kBlock), 0, s, k.gnode, gs * g->work_nodes, 0.0);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_walk_adjoint(c, tree_of(d), g->precision, k.wx, k.wy, k.wz, k.gself, k.gnode, k.gbody, &g->lists,
                               s));
    if (g->quad)
        hipLaunchKernelGGL(k_adj_push<true>, dim3(grid_of(n)), dim3(kBlock), 0, s, d.tleaf, d.parent, d.mass, d.com,
                           k.gnode, k.gbody, d.tx, d.ty, d.tz, d.tm, d.tperm, n, d.walk, d.walk32, k.push);
    else
        hipLaunchKernelGGL(k_adj_push<false>, dim3(grid_of(n)), dim3(kBlock), 0, s, d.tleaf, d.parent, d.mass, d.com,
                           k.gnode, k.gbody, d.tx, d.ty, d.tz, d.tm, d.tperm, n, d.walk, d.walk32, k.push);
    HIPCHK(hipGetLastError());
    return BHX_OK;
}

// The adjoint's list regions and entry records, in one allocation.  A
// region per walk group of 8,192 entries (about twice a typical 64-body group
// list), fewer when there are many groups (at most 2^26 entries in all); what
// does not fit is evaluated by the traversal itself.
int alloc_adj_lists(bhx_ctx* c, GradRecord* g) {
    AdjLists& L = g->lists;
    const size_t groups = c->ngroups > 0 ? c->ngroups : 1;
    size_t cap = ((size_t)1 << 26) / groups;
    cap = cap > 8192 ? 8192 : cap;
    cap = cap < 1024 ? 1024 : cap & ~(size_t)63;
    if (c->adj_list_cap) cap = c->adj_list_cap;
    const size_t pairs = groups * cap;
    const size_t ps = g->quad ? 10 : 4;
    // sort keys: cell ids below node_cap (every stored tree's node ids), bodies
    // node_cap + tree-order index
    const size_t kmax = g->work_nodes + c->n;
    int bits = 1;
    while (bits < 64 && (((size_t)1 << bits) <= kmax)) ++bits;
    const int passes = (bits + 7) / 8;
    const size_t tiles = sort_tiles_for(pairs);
    const size_t bytes[] = {groups * cap * 4, groups * 8, (groups + 1) * 4, pairs * ps * 8, pairs * 8, pairs * 8,
                            pairs * 4, pairs * 4, 8 * 256 * 4, (size_t)passes * tiles * 256 * 4, 8 * 4, kmax * 8};
    size_t total = 0;
    for (size_t b : bytes) total += align_up(b);
    if (hipMalloc(&g->lists_mem, total) != hipSuccess) {
        g->lists_mem = nullptr;
        return fail(c, BHX_E_NOMEM, "adjoint list allocation");
    }
    char* p = (char*)g->lists_mem;
    void** dst[] = {(void**)&L.lists, (void**)&L.list_n, (void**)&L.off, (void**)&L.part, (void**)&L.key,
                    (void**)&L.key_alt, (void**)&L.val, (void**)&L.val_alt, (void**)&L.hist, (void**)&L.status,
                    (void**)&L.tile_ctr, (void**)&L.run};
    for (size_t k = 0; k < sizeof bytes / sizeof bytes[0]; ++k) {
        *dst[k] = p;
        p += align_up(bytes[k]);
    }
    L.cap = (uint32_t)cap;
    L.node_cap = (uint32_t)g->work_nodes;
    L.max_pairs = pairs;
    L.groups = c->ngroups;
    L.passes = passes;
    L.nkeys = kmax;
    return BHX_OK;
}

}  // namespace

extern "C" {

int bhx_record(bhx_ctx* c, int nsteps, int checkpoint_every, int init_acc) {
    if (!c) return BHX_E_NULL;
    if (nsteps < 1 || checkpoint_every < 1) return fail(c, BHX_E_ARG, "step count and checkpoint interval must be >= 1");
    const bhx_config& cf = c->cfg;
    if (!cf.group_walk || cf.mode != BHX_MODE_INTENDED || !cf.permute_acc)
        return fail(c, BHX_E_STATE, "differentiation needs group_walk=1, the intended payload, permute_acc=1");
    if (c->walk_sub != 1)
        return fail(c, BHX_E_STATE, "differentiation needs one interaction list per 64-body group (group_size 64)");
    if (c->n == 0) return fail(c, BHX_E_STATE, "no bodies");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipStreamSynchronize(c->s_side));
    free_grad(c);
    GradRecord* g = new GradRecord();
    c->grad = g;
    g->nsteps = nsteps;
    g->every = checkpoint_every;
    g->init_acc = init_acc ? 1 : 0;
    g->precision = cf.precision;
    g->quad = cf.quadrupole ? 1 : 0;
    g->dt = cf.dt;
    g->n = c->n;
    g->nsys = c->nsys;
    g->state_len = 16 * c->n + 4 * c->nsys;
    const size_t n = c->n;
    const int nck = (nsteps + checkpoint_every - 1) / checkpoint_every;
    int rc = BHX_OK;
    // the decision digest is collected by the FP64 group walk
    const bool digest = cf.precision != BHX_FP32;
    c->digest_valid = false;
    auto bail = [&](int code) {
        hipLaunchKernelGGL(k_digest_set, dim3(1), dim3(64), 0, s, c->digest, 0ull, 0ull);
        free_grad(c);
        return code;
    };
    uint64_t init_digest = 0;
    g->ckpt.assign(nck, nullptr);
    for (double*& p : g->ckpt)
        if (dalloc(&p, g->state_len) != hipSuccess) return bail(fail(c, BHX_E_NOMEM, "checkpoint allocation"));
    if (dalloc(&g->final_state, g->state_len) != hipSuccess || dalloc(&g->ident, n) != hipSuccess ||
        dalloc(&g->ident_tmp, n) != hipSuccess)
        return bail(fail(c, BHX_E_NOMEM, "record allocation"));
    if (g->init_acc) {
        // a_0 = A(x_0): the tree of the starting positions in their own bounds
        HIPCHK(launch_bounds(c, c->bounds_slot(c->bslot), s));
        c->bounds_valid = true;
        rc = build_on_device(c, s);
        if (rc != BHX_OK) return bail(rc);
        if (digest) hipLaunchKernelGGL(k_digest_set, dim3(1), dim3(64), 0, s, c->digest, 1ull, 0ull);
        HIPCHK(launch_walk(c, s));
        if (digest) {
            hipLaunchKernelGGL(k_tree_digest, dim3(grid_of(n)), dim3(kBlock), 0, s, c->tperm, c->tleaf, c->tree.key,
                               nullptr, n, 0ull, c->ctr, c->tree.cap, c->digest);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(&init_digest, c->digest, sizeof init_digest, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        const size_t nodes = c->n_nodes;
        void* st = nullptr;
        if (hipMalloc(&st, slot_bytes(n, nodes, c->nsys, slot_quad(g->precision, g->quad))) != hipSuccess)
            return bail(fail(c, BHX_E_NOMEM, "store allocation"));
        g->init_store = st;
        char* p = (char*)st;
        g->init_slot = carve_slot(p, n, nodes, c->nsys, g->precision, slot_quad(g->precision, g->quad));
        rc = store_frame(c, g->init_slot, nodes, g->precision, false, s);
        if (rc != BHX_OK) return bail(rc);
        g->work_nodes = nodes;
    }
    if (!c->bounds_valid) {
        HIPCHK(launch_bounds(c, c->bounds_slot(c->bslot), s));
        c->bounds_valid = true;
    }
    bool done = false;
    for (int attempt = 0; attempt < 3 && !done; ++attempt) {
        if (attempt > 0) {
            rc = load_state(c, g->ckpt[0], s);
            if (rc != BHX_OK) return bail(rc);
        }
        // every frame's walk data, when it fits the memory budget: the reverse
        // pass then reads it instead of re-running the frames from checkpoints
        if (g->all_store) {
            HIPCHK(hipStreamSynchronize(s));
            (void)hipFree(g->all_store);
            g->all_store = nullptr;
            g->all_slots.clear();
        }
        {
            const size_t one = slot_bytes(n, c->tree.cap, c->nsys, slot_quad(g->precision, g->quad));
            uint64_t budget = c->record_memory;
            if (budget == UINT64_MAX) {
                size_t fr = 0, tot = 0;
                budget = hipMemGetInfo(&fr, &tot) == hipSuccess ? fr / 2 : 0;
            }
            if (one <= budget / (uint64_t)nsteps) {
                if (hipMalloc(&g->all_store, one * (size_t)nsteps) == hipSuccess) {
                    char* p = (char*)g->all_store;
                    g->all_slots.resize((size_t)nsteps);
                    for (auto& sl : g->all_slots)
                        sl = carve_slot(p, n, c->tree.cap, c->nsys, g->precision, slot_quad(g->precision, g->quad));
                    g->all_nodes = c->tree.cap;
                } else {
                    (void)hipGetLastError();
                    g->all_store = nullptr;
                }
            }
        }
        hipLaunchKernelGGL(k_iota32, dim3(grid_of(n)), dim3(kBlock), 0, s, g->ident, n);
        hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, s, c->ctr);   // count nodes from here
        if (digest) hipLaunchKernelGGL(k_digest_set, dim3(1), dim3(64), 0, s, c->digest, 1ull, 1ull);
        for (int f = 0; f < nsteps; ++f) {
            if (f % checkpoint_every == 0) {
                rc = save_state(c, g->ckpt[f / checkpoint_every], s);
                if (rc != BHX_OK) return bail(rc);
            }
            rc = graph_frame(c, s);
            if (rc != BHX_OK) return bail(rc);
            if (g->all_store) {
                rc = store_frame(c, g->all_slots[(size_t)f], g->all_nodes, g->precision, true, s);
                if (rc != BHX_OK) return bail(rc);
            }
            hipLaunchKernelGGL(k_compose, dim3(grid_of(n)), dim3(kBlock), 0, s, g->ident, c->vals, n, g->ident_tmp);
            std::swap(g->ident, g->ident_tmp);
            if (digest)
                hipLaunchKernelGGL(k_tree_digest, dim3(grid_of(n)), dim3(kBlock), 0, s, c->tperm, c->tleaf,
                                   c->tree.key, g->ident, n, (uint64_t)f + 1, c->ctr, c->tree.cap, c->digest);
        }
        rc = read_counters(c, s);
        if (rc != BHX_OK) return bail(rc);
        if (!c->ctr_host->overflow_seen) {
            done = true;
            break;
        }
        const size_t need = c->ctr_host->max_nodes;
        HIPCHK(alloc_tree(c, need + need / 4 + 1024));
    }
    if (!done) return bail(fail(c, BHX_E_NOMEM, "node capacity"));
    g->node_cap = c->ctr_host->max_nodes;
    if (g->node_cap > g->work_nodes) g->work_nodes = g->node_cap;
    rc = save_state(c, g->final_state, s);
    if (rc != BHX_OK) return bail(rc);
    uint64_t frames_digest = 0;
    if (digest) {
        HIPCHK(hipMemcpyAsync(&frames_digest, c->digest, sizeof frames_digest, hipMemcpyDeviceToHost, s));
        hipLaunchKernelGGL(k_digest_set, dim3(1), dim3(64), 0, s, c->digest, 0ull, 0ull);   // stop collecting
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    c->digest_host = init_digest + frames_digest;
    c->digest_valid = digest;
    c->n_nodes = c->ctr_host->total_nodes;
    c->tree_valid = c->tree_built = true;
    c->tree_n = c->n;
    g->record_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return BHX_OK;
}

int bhx_recorded_bodies(bhx_ctx* c, double* pos, double* vel, double* mass, double* acc) {
    if (!c) return BHX_E_NULL;
    GradRecord* g = c->grad;
    if (!g) return fail(c, BHX_E_STATE, "no recorded trajectory");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    const size_t n = g->n;
    std::vector<uint32_t> id(n);
    std::vector<double> st(16 * n);
    HIPCHK(hipMemcpyAsync(id.data(), g->ident, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(st.data(), g->final_state, 16 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // saved arrays: pos 0-2, vel 3-5, mass 6, compensation 7-12, acc 13-15 (slot order)
    for (size_t i = 0; i < n; ++i) {
        const size_t j = id[i];
        for (int a = 0; a < 3; ++a) {
            if (pos) pos[3 * j + a] = st[(0 + a) * n + i];
            if (vel) vel[3 * j + a] = st[(3 + a) * n + i];
            if (acc) acc[3 * j + a] = st[(13 + a) * n + i];
        }
        if (mass) mass[j] = st[6 * n + i];
    }
    return BHX_OK;
}

int bhx_record_digest(bhx_ctx* c, uint64_t* out) {
    if (!c || !out) return c ? fail(c, BHX_E_NULL, "digest output is a null pointer") : BHX_E_NULL;
    if (!c->grad) return fail(c, BHX_E_STATE, "no recorded trajectory");
    if (!c->digest_valid) return fail(c, BHX_E_STATE, "the decision digest is collected by the FP64 walk only");
    *out = c->digest_host;
    return BHX_OK;
}

int bhx_adjoint(bhx_ctx* c, const double* gx, const double* gv, const double* ga, double* gx0, double* gv0,
                double* gm, double* ga0) {
    if (!c) return BHX_E_NULL;
    GradRecord* g = c->grad;
    if (!g) return fail(c, BHX_E_STATE, "no recorded trajectory");
    if (g->n != c->n || g->nsys != c->nsys) return fail(c, BHX_E_STATE, "bodies changed since the recording");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipStreamSynchronize(c->s_side));
    const size_t n = c->n;
    const int every = g->every;
    const bool kept = g->all_store != nullptr;   // every frame stored while recording: nothing to re-run
    if (!kept && !g->seg_store) {
        const size_t one = slot_bytes(n, g->node_cap, c->nsys, slot_quad(g->precision, g->quad));
        void* st = nullptr;
        if (hipMalloc(&st, 2 * one * (size_t)every) != hipSuccess) return fail(c, BHX_E_NOMEM, "store allocation");
        g->seg_store = st;
        char* p = (char*)st;
        g->slots.resize(2 * (size_t)every);
        for (auto& sl : g->slots) sl = carve_slot(p, n, g->node_cap, c->nsys, g->precision, slot_quad(g->precision, g->quad));
    }
    if (!g->events) {
        g->events = true;
        for (int b = 0; b < 2; ++b) {
            HIPCHK(hipEventCreateWithFlags(&g->ev_stored[b], hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&g->ev_consumed[b], hipEventDisableTiming));
        }
        HIPCHK(hipEventCreateWithFlags(&g->ev_ready, hipEventDisableTiming));
    }
    if (!g->work) {
        if (dalloc(&g->work, work_len(n, g->work_nodes, g->quad != 0)) != hipSuccess) return fail(c, BHX_E_NOMEM, "adjoint allocation");
    }
    if (g->lists_mem && (g->lists.groups != c->ngroups ||
                         (c->adj_list_cap != 0 && g->lists.cap != c->adj_list_cap))) {
        (void)hipStreamSynchronize(c->s_adj);
        (void)hipFree(g->lists_mem);
        g->lists_mem = nullptr;
    }
    if (!g->lists_mem) {
        int rc = alloc_adj_lists(c, g);
        if (rc != BHX_OK) return rc;
    }
    hipStream_t sa = c->s_adj;
    Work k = carve_work(g->work, n, g->work_nodes);
    hipLaunchKernelGGL(k_zero_flags, dim3(1), dim3(64), 0, s, c->ctr);
    // incoming adjoints (starting order) into the slot order of the last frame; dL/dm starts at 0
    {
        const double* ins[3] = {gx, gv, ga};
        double* outs[3] = {k.X[0], k.V[0], k.A[0]};
        for (int q = 0; q < 3; ++q) {
            if (ins[q]) {
                HIPCHK(hipMemcpyAsync(c->stage, ins[q], 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(k_from_start, dim3(grid_of(n)), dim3(kBlock), 0, s, c->stage, g->ident, n, outs[q]);
            } else {
                hipLaunchKernelGGL(k_fill_d, dim3(grid_of(3 * n)), dim3(kBlock), 0, s, outs[q], 3 * n, 0.0);
            }
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(s));   // the stage is reused by the next upload
        }
        hipLaunchKernelGGL(k_fill_d, dim3(grid_of(n)), dim3(kBlock), 0, s, k.M[0], n, 0.0);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(g->ev_ready, s));
        HIPCHK(hipStreamWaitEvent(sa, g->ev_ready, 0));
    }
    int cur = 0;
    const int nseg = (int)g->ckpt.size();
    int rc = BHX_OK;
    bool bank_read[2] = {false, false};
    // the frames of segment `seg` again, bitwise as recorded, into store bank `bank`
    auto rerun = [&](int seg) -> int {
        const int bank = seg & 1;
        const int f0 = seg * every, f1 = std::min(g->nsteps, f0 + every);
        if (bank_read[bank]) HIPCHK(hipStreamWaitEvent(s, g->ev_consumed[bank], 0));
        int r = load_state(c, g->ckpt[seg], s);
        if (r != BHX_OK) return r;
        for (int f = f0; f < f1; ++f) {
            // linear frames: the adjoint stream runs beside them
            r = graph_frame(c, s, true);
            if (r != BHX_OK) return r;
            r = store_frame(c, g->slots[(size_t)bank * every + (f - f0)], g->node_cap, g->precision, true, s);
            if (r != BHX_OK) return r;
        }
        HIPCHK(hipEventRecord(g->ev_stored[bank], s));
        return BHX_OK;
    };
    if (!kept) {
        rc = rerun(nseg - 1);
        if (rc != BHX_OK) return rc;
    }
    for (int seg = nseg - 1; seg >= 0; --seg) {
        const int bank = seg & 1;
        const int f0 = seg * every, f1 = std::min(g->nsteps, f0 + every);
        if (!kept) {
            if (seg > 0) {   // queued ahead: runs while this segment's adjoint walks do
                rc = rerun(seg - 1);
                if (rc != BHX_OK) return rc;
            }
            HIPCHK(hipStreamWaitEvent(sa, g->ev_stored[bank], 0));
        }
        for (int f = f1 - 1; f >= f0; --f) {
            const GradSlot& d = kept ? g->all_slots[(size_t)f] : g->slots[(size_t)bank * every + (f - f0)];
            rc = adjoint_walk(c, g, k, d, k.A[cur], k.V[cur], 0.5 * g->dt, sa);
            if (rc != BHX_OK) return rc;
            hipLaunchKernelGGL(k_adj_kdk, dim3(grid_of(n)), dim3(kBlock), 0, sa, k.X[cur], k.V[cur], k.M[cur],
                               k.gself, k.push, d.perm, n, g->dt, k.X[cur ^ 1], k.V[cur ^ 1], k.A[cur ^ 1],
                               k.M[cur ^ 1]);
            HIPCHK(hipGetLastError());
            cur ^= 1;
        }
        if (!kept) {
            HIPCHK(hipEventRecord(g->ev_consumed[bank], sa));
            bank_read[bank] = true;
        }
    }
    if (g->init_acc) {
        // a_0 = A(x_0, m): its adjoint weights are dL/da_0
        rc = adjoint_walk(c, g, k, g->init_slot, k.A[cur], k.A[cur], 0.0, sa);
        if (rc != BHX_OK) return rc;
        hipLaunchKernelGGL(k_adj_init, dim3(grid_of(n)), dim3(kBlock), 0, sa, k.X[cur], k.M[cur], k.gself, k.push, n);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_fill_d, dim3(grid_of(3 * n)), dim3(kBlock), 0, sa, k.A[cur], 3 * n, 0.0);
        HIPCHK(hipGetLastError());
    }
    // the starting slot order is the caller's order
    double* res[4] = {gx0, gv0, ga0, gm};
    const double* src[4] = {k.X[cur], k.V[cur], k.A[cur], k.M[cur]};
    const size_t len[4] = {3 * n, 3 * n, 3 * n, n};
    for (int q = 0; q < 4; ++q)
        if (res[q]) HIPCHK(hipMemcpyAsync(res[q], src[q], len[q] * sizeof(double), hipMemcpyDeviceToHost, sa));
    HIPCHK(hipStreamSynchronize(sa));
    // leave the context at the end of the trajectory
    rc = load_state(c, g->final_state, s);
    if (rc != BHX_OK) return rc;
    rc = read_counters(c, s);
    if (rc != BHX_OK) return rc;
    if (c->ctr_host->overflow_seen) return fail(c, BHX_E_STATE, "a re-run frame differs from the recording");
    g->adjoint_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return BHX_OK;
}

int bhx_set_record_memory(bhx_ctx* c, uint64_t bytes) {
    if (!c) return BHX_E_NULL;
    c->record_memory = bytes;
    return BHX_OK;
}

int bhx_set_adjoint_list_capacity(bhx_ctx* c, uint32_t entries) {
    if (!c) return BHX_E_NULL;
    if (entries != 0 && (entries < 64 || entries % 64 != 0))
        return fail(c, BHX_E_ARG, "adjoint list capacity must be 0 or a multiple of 64, at least 64");
    c->adj_list_cap = entries;
    return BHX_OK;
}

int bhx_record_keeps_frames(bhx_ctx* c, int* kept) {
    if (!c || !kept) return c ? fail(c, BHX_E_NULL, "output is a null pointer") : BHX_E_NULL;
    if (!c->grad) return fail(c, BHX_E_STATE, "no recorded trajectory");
    *kept = c->grad->all_store != nullptr;
    return BHX_OK;
}

int bhx_grad_timing(bhx_ctx* c, double* record_ms, double* adjoint_ms) {
    if (!c) return BHX_E_NULL;
    if (!c->grad) return fail(c, BHX_E_STATE, "no recorded trajectory");
    if (record_ms) *record_ms = c->grad->record_ms;
    if (adjoint_ms) *adjoint_ms = c->grad->adjoint_ms;
    return BHX_OK;
}

}  // extern "C"
