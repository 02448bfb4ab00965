// bhx_internal.hpp — shared declarations of the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bhx.h"

namespace bhx {

constexpr int kMaxLevel = 21;   // 1 sentinel bit + 21 x 3 path bits = 64-bit key
constexpr int kMaxLeafCap = 64;
constexpr size_t kMaxSystems = 65536;   // system ids sort in two byte passes

// Walk record of one node, 64 bytes so one wave-uniform node fetch is a
// single scalar load burst.  Pre-order layout: the first child of an internal
// node is the next record, `skip` is the first record outside its subtree.
struct alignas(64) WalkNode {
    double cx, cy, cz;      // centre of mass (barycentre)
    double mass;
    double size2;           // nodeBounds.size * nodeBounds.size (FP64, as the MAC forms it)
    uint32_t skip;
    uint32_t start;         // first body in tree order
    uint32_t count;         // bodies in the cell
    uint32_t level_leaf;    // level | (leaf << 8)
    float size2f;           // size^2 rounded to FP32 (FP32 opening test)
    float massf;
};
static_assert(sizeof(WalkNode) == 64, "walk node is one 64-byte record");

// FP32-walk records, two arrays indexed by node.  A visit reads one
// WalkVisit32 (one 32-byte scalar load; the pre-order traversal moves
// through consecutive records, two per 64-byte line): the centre of mass
// relative to the system's root centre in FP32 (for the opening test only;
// evaluated entries take their group-relative FP32 centre from the FP64
// centres), the structure and the cell edge squared.  The payload an
// evaluated entry needs, mass and quadrupole, sits in WalkPay32, gathered
// lane-parallel when a list is flushed.
struct alignas(32) WalkVisit32 {
    float cx, cy, cz;       // centre of mass - root centre (FP32)
    uint32_t meta;          // level (bits 0-4) | leaf (bit 5) | min(count, 2^26-1) << 6
    uint32_t skip;
    uint32_t start;         // first body in tree order
    uint32_t child_byte;    // bit o: a child in octant o (its id in the octant table)
    float size2;            // cell edge squared
};
static_assert(sizeof(WalkVisit32) == 32, "FP32 visit record is 32 bytes");
struct alignas(32) WalkPay32 {
    float mass;
    float q[6];             // Qxx Qxy Qxz Qyy Qyz Qzz (0 for one-body cells)
    uint32_t pad;
};
static_assert(sizeof(WalkPay32) == 32, "FP32 payload record is 32 bytes");

struct DeviceTree {
    size_t cap = 0;          // node capacity
    uint64_t* key = nullptr;
    uint32_t* start = nullptr;
    uint32_t* count = nullptr;
    uint32_t* skip = nullptr;
    uint32_t* parent = nullptr;
    uint32_t* child_byte = nullptr;
    uint32_t* level = nullptr;
    uint32_t* first = nullptr;       // min array index among members
    double* center = nullptr;        // 3 per node
    double* com = nullptr;           // 3 per node
    double* mass = nullptr;
    double* quad = nullptr;          // 6 per node
    float* quadf = nullptr;          // 6 per node, FP32 copy for the walk
    WalkNode* walk = nullptr;
    WalkVisit32* walk32 = nullptr;
    WalkPay32* pay32 = nullptr;
    uint32_t* child = nullptr;       // 8 per node: child id by octant (valid where child_byte has the bit)
    uint32_t* by_level = nullptr;    // moments: nodes left for the second phase
    uint32_t* lvl = nullptr;         // moments: their count per level, then the split phase's hand-over counts
    // hash table
    size_t hcap = 0;
    uint64_t* hkey = nullptr;
    uint32_t* hval = nullptr;
};

// One piece of a 1024-element tile of the bounds' sequential sum: a run of
// additions whose partial sums share one binade (added at once, exactly), or
// one plain addition.
struct SumPiece {
    double sum;       // run: sum of the increments in units of the binade's ulp; single: the value
    double lo, hi;    // run: least and greatest prefix of those increments
    int16_t e;        // run: binade exponent of its partial sums; -32768: one plain addition;
                      // -32767: a whole tile added element by element
    uint16_t len;     // elements
    uint32_t start;   // first element (body index)
};

struct Counters {          // small device-side status block
    uint32_t total_nodes;
    uint32_t unsplit;
    uint32_t hash_fail;
    uint32_t sort_fail;
    uint32_t nonfinite;       // a step met a non-finite position or velocity
    uint32_t node_overflow;   // the last build asked for more nodes than the tree holds
    uint32_t max_nodes;       // largest node count a build asked for (sticky)
    uint32_t overflow_seen;   // some build overflowed (sticky until the host clears it)
    uint32_t nodes_kept;      // min(total_nodes, capacity): the nodes the last build wrote
    uint32_t pad[7];
};

// Independent systems: contiguous body ranges that are bounded, sorted, built
// and walked each on its own (one system is the plain case).  Bodies never
// leave their system's range, so the system of a body position is fixed.
struct SysView {
    const uint32_t* off;      // nsys + 1 body offsets
    const uint32_t* id;       // system of each body position; null when nsys == 1
    uint32_t nsys;
};

__device__ __forceinline__ uint32_t sys_of(const SysView& v, uint32_t i) { return v.id ? v.id[i] : 0u; }

// s += inc keeping the rounding error in c (Kahan), fed into the next addition.
__device__ __forceinline__ void kahan_add(double& s, double& c, double inc) {
    const double y = inc - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
}

// The closing half kick of a frame for one body — ComputeVelocityAndPosition,
// Body.hpp:317-324: v = v + (a*dt)*0.5, or compensated (c: the velocity's
// running rounding error).  One definition for every kernel that applies it,
// so the bits do not depend on which one did.
__device__ __forceinline__ void half_kick(double& v0, double& v1, double& v2, double* c, double a0, double a1,
                                          double a2, double dt) {
    if (!c) {
        v0 = v0 + (a0 * dt) * 0.5;
        v1 = v1 + (a1 * dt) * 0.5;
        v2 = v2 + (a2 * dt) * 0.5;
    } else {
        kahan_add(v0, c[0], (a0 * dt) * 0.5);
        kahan_add(v1, c[1], (a1 * dt) * 0.5);
        kahan_add(v2, c[2], (a2 * dt) * 0.5);
    }
}

// Node hash table: open addressing with linear probing, an empty slot holds
// key 0 (every valid key carries the root sentinel bit).  A deleted node
// keeps its key in its slot (so probe sequences stay intact) with this value.
constexpr uint32_t kHashErased = 0xFFFFFFFFu;

__device__ __forceinline__ uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// Home slot of (system, key); system 0 hashes the key alone.
__device__ __forceinline__ uint64_t hash_home(uint64_t k, uint32_t sys) {
    return mix64(k ^ ((uint64_t)sys * 0x9e3779b97f4a7c15ull));
}

// The FP32 walks work in scaled units: a system's lengths times lam (the power
// of two nearest below the inverse root edge) and masses times mu (the same
// for its first body's mass).  Power-of-two scalings are exact, so wherever
// the unscaled FP32 values stay in range every rounding and every opening
// decision is unchanged; in any units r^-5, r^-7 and the mass products then
// stay in range.  Results go back to physical units in FP64.
__device__ __forceinline__ double pow2_inverse_near(double v) {
    if (!(v > 0.0) || !(v < 1.0e300)) return 1.0;
    int e = ilogb(v);
    e = e < -1000 ? -1000 : (e > 1000 ? 1000 : e);
    return ldexp(1.0, -e);
}


}  // namespace bhx

namespace bhx {
// Working memory of one radix sort (sort_pairs); the segmented second stage
// (several systems) uses skey..sidx_alt.
struct SortScratch {
    uint64_t *keys_alt = nullptr, *skey = nullptr, *skey_alt = nullptr;
    uint32_t *vals_alt = nullptr, *hist = nullptr, *status = nullptr, *tile_ctr = nullptr;
    uint32_t *sidx = nullptr, *sidx_alt = nullptr;
};
}  // namespace bhx

struct bhx_ctx {
    bhx_config cfg;
    int device = 0;
    hipStream_t s_main = nullptr;
    hipStream_t s_side = nullptr;
    // the reverse pass's adjoint walks (bhx_adjoint).  Created with the other
    // two so that the three take distinct hardware queues of the process's
    // pool: a stream created later can share the main stream's queue, and its
    // kernels then serialise with the main stream's
    hipStream_t s_adj = nullptr;
    // device memory a recording may use to keep every frame's walk data, so
    // that the reverse pass re-runs no frames (bhx_set_record_memory);
    // UINT64_MAX: half of the free device memory when recording
    uint64_t record_memory = UINT64_MAX;
    // entries per walk group the adjoint stores for its evaluation kernel
    // (bhx_set_adjoint_list_capacity); 0: the default
    uint32_t adj_list_cap = 0;
    // the tree branch of a frame (path keys of the drifted positions and
    // their sort) beside the body sort; created after the other three
    hipStream_t s_tree = nullptr;
    hipEvent_t ev_drift = nullptr;
    hipEvent_t ev_bounds = nullptr;
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_tree = nullptr;
    hipEvent_t ev[12] = {};
    bool profiling = false;
    bhx_timing timing{};
    std::string err;

    size_t cap = 0;
    size_t n = 0;
    // bodies, array order, SoA FP64; *_alt are the ping-pong targets of a gather
    double *px = nullptr, *py = nullptr, *pz = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *m = nullptr;
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    double *cpx = nullptr, *cpy = nullptr, *cpz = nullptr;   // compensation terms
    double *cvx = nullptr, *cvy = nullptr, *cvz = nullptr;
    double *alt[16] = {};                     // gather targets: pos 0-2, vel 3-5, m 6, comp 7-12, acc 13-15
    // sort buffers
    uint64_t *keys = nullptr, *keys_alt = nullptr;
    uint32_t *vals = nullptr, *vals_alt = nullptr;
    uint64_t *mkeys = nullptr;                 // Morton keys of the last bhx_morton_keys
    // tree order copies for the walk
    double *tx = nullptr, *ty = nullptr, *tz = nullptr, *tm = nullptr;
    uint64_t* tkeys = nullptr;                 // sorted tree path keys
    uint32_t* tperm = nullptr;                 // tree order -> array index
    uint32_t* tperm_old = nullptr;             // tree order -> array index before the frame's body sort
    uint32_t* inv = nullptr;                   // array index before the body sort -> after
    float4* tb4 = nullptr;                     // tree order: offset from the leaf's centre of mass, mass
    uint32_t* tleaf = nullptr;                 // tree order: node id of the body's leaf
    // build scratch
    uint32_t *nn = nullptr, *noff = nullptr, *nlo = nullptr;
    uint32_t* scan_tmp = nullptr;
    // sort scratch
    uint32_t* sort_hist = nullptr;             // 8 x 256
    uint32_t* sort_status = nullptr;           // 8 x tiles x 256
    uint32_t* sort_tile_ctr = nullptr;         // 8
    size_t sort_tiles = 0;
    // walk scratch
    uint32_t* wcount = nullptr;                // interaction counts per body (array order)
    uint32_t* wq = nullptr;                    // persistent walks: 8 per-XCD group counters
    // decision digest of the FP64 walk (recordings): [0] sum of hashed list
    // entries and tree memberships, [1] non-zero while a recording collects
    // it, [2] the frame index mixed into the hashes
    unsigned long long* digest = nullptr;
    uint64_t digest_host = 0;                  // the last recording's digest
    bool digest_valid = false;
    int walk_sub = 1;                          // FP32 group walk: lists per 64-body wave (1, 2 or 4)
    // bounds: two slots of (cx, cy, cz, size)
    // (nsys x 4 doubles per slot; slots 0 and 1 alternate between frames, slot 2 is scratch)
    double* bounds = nullptr;
    int bslot = 0;
    int tree_slot = 0;                         // bounds slot the current tree was built with
    double tree_tscale = 1.0;                  // root edge / bounds size of the current tree
    bool bounds_valid = false;
    // staging
    double* stage = nullptr;                   // 3 * cap doubles
    bhx::Counters* ctr = nullptr;
    bhx::Counters* ctr_host = nullptr;         // pinned
    bhx::DeviceTree tree;
    size_t n_nodes = 0;
    bool tree_valid = false;   // tree built on the current bodies (walk allowed)
    bool tree_built = false;   // node table of the last build still intact (downloads, lookups)
    bool tree_edited = false;  // nodes put or erased since the build (the walks' records are stale)
    bool hash_valid = false;   // the node hash table holds the current tree (built on first use)
    size_t tree_n = 0;         // bodies the last build covered
    // independent systems
    uint32_t nsys = 1;
    uint32_t max_sys_n = 0;                    // bodies in the largest system
    size_t sys_cap = 0;                        // systems the per-system buffers hold
    std::vector<uint32_t> sys_off_h;           // nsys + 1 body offsets (host copy)
    uint32_t* sys_off = nullptr;               // device copy
    uint32_t* sysid = nullptr;                 // system of each body position (null for one system)
    uint32_t* sys_root = nullptr;              // root node of each system in the current tree
    double* sys_scale = nullptr;               // per system: FP32 walk length and mass factors
    uint4* gtab = nullptr;                     // walk groups: {first body, end body, system, 0}
    // the sequential mean of one system: 1024-body tiles
    size_t sum_tiles_cap = 0;
    double* sum_tile = nullptr;                // tile sums (3 per tile), then their exclusive prefix
    uint32_t* sum_cnt = nullptr;               // pieces per tile and axis, then their exclusive prefix
    bhx::SumPiece* sum_piece = nullptr;        // 64 per tile and axis, then the same compacted per axis
    size_t gtab_cap = 0;
    uint32_t ngroups = 0;
    std::vector<double> tree_bounds_h;         // (centre, size) per system the last tree was built with
    // captured one-step graphs (bhx_step_graph): one per parity of the body
    // ping-pong buffers / bounds slot, each valid for the state it starts from
    // Frame graphs, captured in pairs (the frame from the current buffer
    // parity and bounds slot, and the frame after it).  Two pairs are kept:
    // state restores (checkpoints, bhx_set_bodies) can leave the buffers and
    // the bounds slot in the other pairing, and a single pair would then be
    // captured again at every switch.
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        const double* bufs[32] = {};           // body arrays and gather targets at capture
        int bslot = -1;
        bool linear = false;                   // next bounds computed in line (no side branch)
    } graphs[4];
    int graph_pair_next = 0;                   // pair replaced by the next capture
    bool graph_valid = false;
    struct GradRecord* grad = nullptr;         // recorded trajectory for bhx_adjoint
    // device snapshots of the body state: [0] where a graph run starts
    // (restored on node overflow), [1] where the autotuner starts
    struct Snapshot {
        double* buf = nullptr;
        size_t len = 0;
        double* ptr[26] = {};
        int bslot = 0;
        bool bounds_valid = false;
    } snaps[2];
    // segmented sort scratch (several systems only)
    uint64_t *skey = nullptr, *skey_alt = nullptr;
    uint32_t *sidx = nullptr, *sidx_alt = nullptr;
    // the tree branch's own sort scratch (it runs beside the body sort)
    bhx::SortScratch sort2;

    bhx::SysView sys() const { return bhx::SysView{sys_off, nsys > 1 ? sysid : nullptr, nsys}; }
    // root edge / bounds size of the configured cell geometry (exact: 1 or 2)
    double cfg_tscale() const { return cfg.geometry ? 2.0 : 1.0; }
    double* bounds_slot(int k) const { return bounds + 4 * sys_cap * (size_t)k; }
};

namespace bhx {

// Records `what` (and the HIP error) as the context's last error; returns code.
inline int fail(bhx_ctx* c, int code, const char* what, hipError_t e = hipSuccess) {
    if (c) {
        c->err = what;
        if (e != hipSuccess) {
            c->err += ": ";
            c->err += hipGetErrorString(e);
        }
    }
    return code;
}

#define HIPCHK(expr)                                                 \
    do {                                                             \
        hipError_t e_ = (expr);                                      \
        if (e_ != hipSuccess) return fail(c, BHX_E_HIP, #expr, e_); \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void**)p, count * sizeof(T));
}

// --- context services (bhx_api.hip) ------------------------------------------
void drop_graphs(bhx_ctx* c);
hipError_t alloc_tree(bhx_ctx* c, size_t cap);
int build_on_device(bhx_ctx* c, hipStream_t s);       // tree of the resident bodies, bounds slot c->bslot
int read_counters(bhx_ctx* c, hipStream_t s);         // waits; reports sort / hash failures
// One-frame graph for the current buffer parity.  linear: the next frame's
// bounds are computed in line on the frame's stream instead of on a side
// branch — for callers that run other streams beside the frames (a graph's
// parallel branch executes on a runtime-chosen stream that may share a
// hardware queue with them and wait behind their kernels).
int step_graph_for(bhx_ctx* c, hipGraphExec_t* out, bool linear = false);
int graph_frame(bhx_ctx* c, hipStream_t s, bool linear = false);   // enqueue one frame from its graph
void free_grad(bhx_ctx* c);                           // recorded trajectory (bhx_grad.hip)

// --- launchers implemented in the .hip files (all enqueue on `s`) ----------
// OctantBounds of the resident bodies of every system into bounds4 (centre,
// size: 4 doubles per system).  Uses the context's sum_* scratch: one bounds
// computation at a time per context.
hipError_t launch_bounds(bhx_ctx* c, double* bounds4, hipStream_t s);
// tiles of the one-system sequential mean for n bodies
size_t sum_tiles_for(size_t n);
// nonfinite (optional): set to 1 when some position is not finite
hipError_t launch_morton(const double* px, const double* py, const double* pz, size_t n, const uint32_t* sysid,
                         const double* bounds4, uint64_t* keys, uint32_t* vals, hipStream_t s,
                         uint32_t* nonfinite = nullptr);
// tree path keys; the root edge is tscale * (bounds size)
hipError_t launch_treekeys(const double* px, const double* py, const double* pz, size_t n, const uint32_t* sysid,
                           const double* bounds4, double tscale, uint64_t* keys, uint32_t* vals, hipStream_t s);
hipError_t launch_iota(uint32_t* vals, size_t n, hipStream_t s);

constexpr size_t kSortTileKeys = 8192;   // keys one radix-sort workgroup ranks per pass
size_t sort_tiles_for(size_t n);
// `ndev` non-null: the pair count is read on the device from *ndev (at most
// n; the scratch and the grid are sized for n).  An even number of passes
// leaves the sorted pairs in keys/vals, an odd number in keys_alt/vals_alt.
hipError_t radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                            size_t n, uint32_t* hist, uint32_t* status, uint32_t* tile_ctr,
                            Counters* ctr, hipStream_t s, int passes = 8, const uint32_t* ndev = nullptr);
// stable sort of (key, value) pairs within each system's body range
// Stable sort of (key, value) pairs, system by system; `sc` null: the
// context's main sort scratch.
hipError_t sort_pairs(bhx_ctx* c, uint64_t* keys, uint32_t* vals, hipStream_t s, const SortScratch* sc = nullptr);

hipError_t launch_gather(double* const* src, double* const* dst, int narr, const uint32_t* perm,
                         size_t n, hipStream_t s);
hipError_t launch_permute_kick_drift(bhx_ctx* c, const uint32_t* perm, double dt, hipStream_t s,
                                     uint32_t* inv = nullptr);
// The tree branch of a frame (accelerations moving with their bodies): path
// keys of the drifted positions into tkeys, identity into tperm_old, in the
// order before the body sort; then, after the body sort and kick/drift,
// launch_tree_order forms tperm in the new order and the tree-order copies.
hipError_t launch_drift_treekeys(bhx_ctx* c, double dt, const double* b4, hipStream_t s);
hipError_t launch_tree_order(bhx_ctx* c, hipStream_t s);
hipError_t launch_kick_drift(bhx_ctx* c, double dt, hipStream_t s);
// nonfinite (optional): set to 1 when some kicked velocity is not finite
hipError_t launch_kick(bhx_ctx* c, double dt, hipStream_t s, uint32_t* nonfinite = nullptr);
hipError_t launch_tree_gather(bhx_ctx* c, hipStream_t s);
// exchange the body arrays with their gather targets, as a permute/kick/drift does
void swap_body_buffers(bhx_ctx* c);

hipError_t build_tree(bhx_ctx* c, hipStream_t s, size_t* total_out, bool sync);
hipError_t launch_lookup(bhx_ctx* c, uint32_t sys, const uint64_t* keys, size_t n, int64_t* ids, hipStream_t s);
// The node hash table of the current tree: built on the first lookup or edit
// after a build (the frames themselves never read it).
hipError_t ensure_hash(bhx_ctx* c, hipStream_t s);

hipError_t launch_walk(bhx_ctx* c, hipStream_t s);
// the walk with the frame's closing half kick (v += (a dt)/2) fused where it can be
hipError_t launch_walk_kick(bhx_ctx* c, double dt, hipStream_t s, uint32_t* nonfinite);
struct WalkKick {
    double dt;
    uint32_t* nonfinite;
};
hipError_t launch_walk_to(bhx_ctx* c, hipStream_t s, bool group, int precision, double* ox, double* oy, double* oz,
                          uint32_t* counts, const WalkKick* kick = nullptr, uint32_t* qcounts = nullptr);
hipError_t launch_direct(bhx_ctx* c, const int64_t* targets, size_t nt, double eps, double* acc,
                         double* pot, hipStream_t s);

// The arrays a walk reads: the live tree or a stored copy of one.
struct WalkTree {
    const WalkNode* walk;
    const WalkVisit32* walk32;
    const WalkPay32* pay32;
    const double* scale32;  // per system: FP32 length and mass factors
    const double* com;      // FP64 centres of mass, 3 per node
    const float4* tb4;
    const uint32_t* tleaf;
    const double *tx, *ty, *tz, *tm;
    const uint32_t* tperm;
    const uint32_t* sys_root;
    const double* b4;
    const double* quad;     // FP64 quadrupoles, 6 per node (FP64 walk with quadrupoles)
    const uint32_t* child;  // octant table: 8 child ids per node (the adjoint's traversal)
    const uint32_t* child_byte;  // octant mask per node (the FP64 adjoint's traversal)
};
// Adjoint of the group walk on tree t with frozen lists: per body weights w
// (array order) -> dL/dx of the own position (gself, 3 per body, array
// order) and per-entry (dL/dX, dL/dM) accumulated into gnode (4 per node;
// 10 with quadrupoles: dL/dQ xx xy xz yy yz zz after them) and gbody (4 per
// body, tree order), which must start at zero.
// group walk of the current tree that also writes each body's potential
// (array order); accelerations go to ox/oy/oz
hipError_t launch_walk_potential(bhx_ctx* c, hipStream_t s, double* ox, double* oy, double* oz, double* pot);
// Working memory of the FP32 adjoint walk: each walk group's list region
// (`cap` entries: cell ids from the front, tree-order body indices from the
// back) and stored counts; the places of the lists in their concatenation
// (groups + 1, the last is the total); per stored entry a record of 4 (10
// with quadrupoles) doubles and a (key, place) pair, sorted by key with
// radix-sort scratch for `max_pairs` = groups x cap pairs in `passes` passes,
// and each key's run of records (nkeys = node_cap + bodies).
struct AdjLists {
    uint32_t* lists = nullptr;
    uint2* list_n = nullptr;
    uint32_t* off = nullptr;
    double* part = nullptr;
    uint64_t *key = nullptr, *key_alt = nullptr;
    uint32_t *val = nullptr, *val_alt = nullptr;
    uint32_t *hist = nullptr, *status = nullptr, *tile_ctr = nullptr;
    uint2* run = nullptr;                      // per key: its records' range in the sorted order
    uint32_t cap = 0, node_cap = 0;
    size_t max_pairs = 0, groups = 0, nkeys = 0;
    int passes = 2;
};
// `L`: the walk's stored lists and their evaluation (both precisions).
hipError_t launch_walk_adjoint(bhx_ctx* c, const WalkTree& t, int precision, const double* wx, const double* wy,
                               const double* wz, double* gself, double* gnode, double* gbody, const AdjLists* L,
                               hipStream_t s);
hipError_t launch_diag(bhx_ctx* c, double* out7, hipStream_t s);

// the reference's per-body tree API (bhx_dropin.hip); one system
hipError_t launch_ref_lists(bhx_ctx* c, const double* pts, size_t nq, double theta, const uint64_t* offsets,
                            uint64_t* counts, int64_t* entries, uint32_t* err, hipStream_t s);
hipError_t launch_force_lists(bhx_ctx* c, const double* pts, size_t nq, const uint64_t* offsets,
                              const int64_t* entries, double* acc, uint32_t* err, hipStream_t s);
hipError_t launch_mac(bhx_ctx* c, const int64_t* ids, const double* pts, size_t n, double theta, uint8_t* out,
                      uint32_t* err, hipStream_t s);
hipError_t launch_node_put(bhx_ctx* c, const bhx_node& rec, uint32_t id, bool insert, uint32_t* err,
                           hipStream_t s);
hipError_t launch_node_erase(bhx_ctx* c, uint64_t key, uint32_t* done, hipStream_t s);
hipError_t launch_prune(bhx_ctx* c, uint64_t root_key, uint32_t root_level, uint32_t* removed, hipStream_t s);
hipError_t launch_decode(const uint64_t* keys, size_t n, double size, double* out, hipStream_t s);

hipError_t scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* tmp,
                              uint32_t* total_dev, hipStream_t s);

}  // namespace bhx
