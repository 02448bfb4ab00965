// bhx_api.hip — context management and the C ABI (include/bhx.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "bhx_internal.hpp"

using namespace bhx;

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    if (g == 0) g = 1;
    if (g > 65535u * 16u) g = 65535u * 16u;
    return (unsigned)g;
}

__global__ void k_aos_to_soa(const double* __restrict__ aos, size_t n, double* __restrict__ x, double* __restrict__ y,
                             double* __restrict__ z) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        x[i] = aos[3 * i];
        y[i] = aos[3 * i + 1];
        z[i] = aos[3 * i + 2];
    }
}

__global__ void k_soa_to_aos(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                             size_t n, double* __restrict__ aos) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        aos[3 * i] = x[i];
        aos[3 * i + 1] = y[i];
        aos[3 * i + 2] = z[i];
    }
}

__global__ void k_zero_d(double* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

__global__ void k_u32_to_u64(const uint32_t* __restrict__ a, size_t n, uint64_t* __restrict__ b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

__global__ void k_u64_to_u32(const uint64_t* __restrict__ a, size_t n, uint32_t* __restrict__ b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (uint32_t)a[i];
}

__global__ void k_copy_u64(const uint64_t* __restrict__ a, size_t n, uint64_t* __restrict__ b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

}  // namespace

namespace bhx {

void drop_graphs(bhx_ctx* c) {
    for (auto& g : c->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        g = bhx_ctx::StepGraph{};
    }
    c->graph_pair_next = 0;
    c->graph_valid = false;
}

void free_tree(bhx_ctx* c) {
    DeviceTree& t = c->tree;
    void* ptrs[] = {t.key, t.start, t.count, t.skip, t.parent, t.child_byte, t.level, t.first, t.center,
                    t.com, t.mass, t.quad, t.quadf, t.walk, t.walk32, t.pay32, t.child, t.by_level, t.lvl, t.hkey, t.hval};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    t = DeviceTree{};
}

hipError_t alloc_tree(bhx_ctx* c, size_t cap) {
    drop_graphs(c);   // captured launches hold the old node arrays
    free_tree(c);
    DeviceTree& t = c->tree;
    hipError_t e;
    if ((e = dalloc(&t.key, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.start, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.count, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.skip, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.parent, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.child_byte, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.level, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.first, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.center, 3 * cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.com, 3 * cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.mass, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.quad, 6 * cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.quadf, 6 * cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.walk, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.walk32, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.pay32, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.child, 8 * cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.by_level, cap)) != hipSuccess) return e;
    if ((e = dalloc(&t.lvl, 3 * (kMaxLevel + 1) + 1)) != hipSuccess) return e;
    // the moments' per-level counts (cleared by every build) and, after them, the
    // hand-over counts of the split phase (zero between builds)
    if ((e = hipMemset(t.lvl, 0, (3 * (kMaxLevel + 1) + 1) * sizeof(uint32_t))) != hipSuccess) return e;
    size_t h = 1024;
    while (h < 2 * cap) h <<= 1;
    if ((e = dalloc(&t.hkey, h)) != hipSuccess) return e;
    if ((e = dalloc(&t.hval, h)) != hipSuccess) return e;
    t.cap = cap;
    t.hcap = h;
    return hipSuccess;
}

void free_all(bhx_ctx* c) {
    free_grad(c);
    drop_graphs(c);
    free_tree(c);
    for (auto& sn : c->snaps)
        if (sn.buf) { (void)hipFree(sn.buf); sn.buf = nullptr; }
    double** dp[] = {&c->px, &c->py, &c->pz, &c->vx, &c->vy, &c->vz, &c->m, &c->ax, &c->ay, &c->az,
                     &c->cpx, &c->cpy, &c->cpz, &c->cvx, &c->cvy, &c->cvz, &c->tx, &c->ty, &c->tz, &c->tm,
                     &c->bounds, &c->stage};
    for (double** p : dp)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    for (double*& p : c->alt)
        if (p) { (void)hipFree(p); p = nullptr; }
    void* vp[] = {c->keys, c->keys_alt, c->vals, c->vals_alt, c->mkeys, c->tkeys, c->tperm, c->nn, c->noff,
                  c->nlo, c->scan_tmp, c->sort_hist, c->sort_status, c->sort_tile_ctr, c->wcount, c->wq, c->digest, c->ctr, c->tb4, c->tleaf,
                  c->sys_off, c->sysid, c->sys_root, c->sys_scale, c->gtab, c->skey, c->skey_alt, c->sidx, c->sidx_alt,
                  c->tperm_old, c->inv, c->sort2.keys_alt, c->sort2.vals_alt, c->sort2.hist, c->sort2.tile_ctr,
                  c->sort2.status, c->sort2.skey, c->sort2.skey_alt, c->sort2.sidx, c->sort2.sidx_alt,
                  c->sum_tile, c->sum_cnt, c->sum_piece};
    for (void* p : vp)
        if (p) (void)hipFree(p);
    if (c->ctr_host) (void)hipHostFree(c->ctr_host);
    for (hipEvent_t& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->ev_drift) (void)hipEventDestroy(c->ev_drift);
    if (c->ev_bounds) (void)hipEventDestroy(c->ev_bounds);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_tree) (void)hipEventDestroy(c->ev_tree);
    if (c->s_tree) (void)hipStreamDestroy(c->s_tree);
    if (c->s_main) (void)hipStreamDestroy(c->s_main);
    if (c->s_side) (void)hipStreamDestroy(c->s_side);
    if (c->s_adj) (void)hipStreamDestroy(c->s_adj);
}

int check_cfg(const bhx_config* cfg) {
    if (!(cfg->theta >= 0.0) || !(cfg->eps >= 0.0)) return BHX_E_ARG;
    if (cfg->leaf_capacity < 1 || cfg->leaf_capacity > kMaxLeafCap) return BHX_E_ARG;
    if (cfg->mode != BHX_MODE_AS_RUN && cfg->mode != BHX_MODE_INTENDED) return BHX_E_ARG;
    if (cfg->precision != BHX_FP64 && cfg->precision != BHX_FP32) return BHX_E_ARG;
    if (cfg->geometry != 0 && cfg->geometry != 1) return BHX_E_ARG;
    if (cfg->group_size != 0 && cfg->group_size != 16 && cfg->group_size != 32 && cfg->group_size != 64)
        return BHX_E_ARG;
    return BHX_OK;
}

// Lists per 64-body wave of the FP32 group walk for a configured group size.
int walk_lists_for(const bhx_config* cfg) { return cfg->group_size == 16 ? 4 : (cfg->group_size == 32 ? 2 : 1); }

// Per-system buffers for `nsys` systems (bounds slots, offsets, roots).
hipError_t ensure_sys_cap(bhx_ctx* c, size_t nsys) {
    if (nsys <= c->sys_cap && c->bounds) return hipSuccess;
    size_t cap = c->sys_cap ? c->sys_cap : 1;
    while (cap < nsys) cap *= 2;
    double *nb = nullptr, *ns = nullptr;
    uint32_t *no = nullptr, *nr = nullptr;
    hipError_t e = dalloc(&nb, 12 * cap);
    if (e == hipSuccess) e = dalloc(&no, cap + 1);
    if (e == hipSuccess) e = dalloc(&nr, cap);
    if (e == hipSuccess) e = dalloc(&ns, 2 * cap);
    if (e != hipSuccess) {
        if (nb) (void)hipFree(nb);
        if (no) (void)hipFree(no);
        if (nr) (void)hipFree(nr);
        return e;
    }
    if (c->bounds) (void)hipFree(c->bounds);
    if (c->sys_off) (void)hipFree(c->sys_off);
    if (c->sys_root) (void)hipFree(c->sys_root);
    if (c->sys_scale) (void)hipFree(c->sys_scale);
    c->bounds = nb;
    c->sys_off = no;
    c->sys_root = nr;
    c->sys_scale = ns;
    c->sys_cap = cap;
    c->bounds_valid = false;
    return hipSuccess;
}

// Splits the bodies into systems by `off` (nsys + 1 offsets, first 0, last n,
// strictly increasing) and lays out the walk groups: runs of up to 64 bodies
// that start at each system's first body.
int set_layout(bhx_ctx* c, const std::vector<uint32_t>& off) {
    const uint32_t nsys = (uint32_t)(off.size() - 1);
    HIPCHK(ensure_sys_cap(c, nsys));
    std::vector<uint4> groups;
    uint32_t maxn = 0;
    for (uint32_t k = 0; k < nsys; ++k) {
        const uint32_t a = off[k], b = off[k + 1];
        if (b - a > maxn) maxn = b - a;
        for (uint32_t g = a; g < b; g += 64) groups.push_back(make_uint4(g, g + 64 < b ? g + 64 : b, k, 0u));
    }
    const size_t gcap = c->cap / 64 + 1 + nsys;
    if (!c->gtab || c->gtab_cap < gcap) {
        if (c->gtab) (void)hipFree(c->gtab);
        c->gtab = nullptr;
        c->gtab_cap = 0;
        HIPCHK(dalloc(&c->gtab, gcap));
        c->gtab_cap = gcap;
    }
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    HIPCHK(hipMemcpyAsync(c->sys_off, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (!groups.empty())
        HIPCHK(hipMemcpyAsync(c->gtab, groups.data(), groups.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
    if (nsys > 1) {
        if (!c->sysid) {
            HIPCHK(dalloc(&c->sysid, c->cap));
            HIPCHK(dalloc(&c->skey, c->cap));
            HIPCHK(dalloc(&c->skey_alt, c->cap));
            HIPCHK(dalloc(&c->sidx, c->cap));
            HIPCHK(dalloc(&c->sidx_alt, c->cap));
            HIPCHK(dalloc(&c->sort2.skey, c->cap));
            HIPCHK(dalloc(&c->sort2.skey_alt, c->cap));
            HIPCHK(dalloc(&c->sort2.sidx, c->cap));
            HIPCHK(dalloc(&c->sort2.sidx_alt, c->cap));
        }
        std::vector<uint32_t> id(off[nsys]);
        for (uint32_t k = 0; k < nsys; ++k)
            for (uint32_t i = off[k]; i < off[k + 1]; ++i) id[i] = k;
        if (!id.empty())
            HIPCHK(hipMemcpyAsync(c->sysid, id.data(), id.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    } else {
        HIPCHK(hipStreamSynchronize(s));
    }
    if (off != c->sys_off_h) c->tree_built = false;
    drop_graphs(c);
    c->sys_off_h = off;
    c->nsys = nsys;
    c->max_sys_n = maxn;
    c->ngroups = (uint32_t)groups.size();
    c->bounds_valid = false;
    c->tree_valid = false;
    c->graph_valid = false;
    return BHX_OK;
}

// tree keys -> sort -> tree-order copies -> node set, hash, moments
int build_on_device(bhx_ctx* c, hipStream_t s) {
    c->tree_valid = false;
    c->tree_built = false;
    if (c->n == 0) {
        c->n_nodes = 0;
        c->tree_valid = c->tree_built = true;
        c->tree_n = 0;
        return BHX_OK;
    }
    const double* b4 = c->bounds_slot(c->bslot);
    HIPCHK(launch_treekeys(c->px, c->py, c->pz, c->n, c->sys().id, b4, c->cfg_tscale(), c->tkeys, c->tperm, s));
    HIPCHK(sort_pairs(c, c->tkeys, c->tperm, s));
    HIPCHK(launch_tree_gather(c, s));
    size_t total = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        HIPCHK(build_tree(c, s, &total, true));
        if (total <= c->tree.cap) break;
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(alloc_tree(c, total + total / 4 + 1024));
    }
    if (total > c->tree.cap) return fail(c, BHX_E_NOMEM, "node capacity");
    c->n_nodes = total;
    c->tree_valid = c->tree_built = true;
    c->tree_edited = false;
    c->tree_n = c->n;
    return BHX_OK;
}

__global__ void k_fill_u32_api(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Steps flag a non-finite position (at each frame's keys) or velocity (at
// each closing kick); the flag is cleared when a stepping call starts and when
// it has been reported.
hipError_t clear_nonfinite(bhx_ctx* c, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u32_api, dim3(1), dim3(64), 0, s, &c->ctr->nonfinite, (size_t)1, 0u);
    return hipGetLastError();
}

int read_counters(bhx_ctx* c, hipStream_t s) {
    HIPCHK(hipMemcpyAsync(c->ctr_host, c->ctr, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (c->ctr_host->nonfinite) {
        HIPCHK(clear_nonfinite(c, s));
        HIPCHK(hipStreamSynchronize(s));
        return fail(c, BHX_E_NONFINITE, "a step produced a non-finite position or velocity");
    }
    if (c->ctr_host->sort_fail) return fail(c, BHX_E_SORT, "radix sort look-back did not complete");
    if (c->ctr_host->hash_fail) return fail(c, BHX_E_HASH_FULL, "node hash table full");
    return BHX_OK;
}

__global__ void k_pick3(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                        const int64_t* __restrict__ idx, size_t n, double* __restrict__ out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int64_t j = idx[i];
        out[3 * i] = x[j];
        out[3 * i + 1] = y[j];
        out[3 * i + 2] = z[j];
    }
}

// The tree branch of a frame on its own stream: path keys of the drifted
// positions (in the order before the body sort) and their sort.
hipError_t fork_tree_branch(bhx_ctx* c, const double* b_cur, double dt, hipStream_t s) {
    hipError_t e = hipEventRecord(c->ev_fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->s_tree, c->ev_fork, 0);
    if (e == hipSuccess) e = launch_drift_treekeys(c, dt, b_cur, c->s_tree);
    if (e == hipSuccess) e = sort_pairs(c, c->tkeys, c->tperm_old, c->s_tree, &c->sort2);
    if (e == hipSuccess) e = hipEventRecord(c->ev_tree, c->s_tree);
    return e;
}

// The frame's tree order after its kick/drift: joined from the tree branch,
// or path keys of the new positions sorted in line.
hipError_t tree_order(bhx_ctx* c, const double* b_cur, bool branch, hipStream_t s) {
    hipError_t e;
    if (branch) {
        e = hipStreamWaitEvent(s, c->ev_tree, 0);
        if (e == hipSuccess) e = launch_tree_order(c, s);
        return e;
    }
    e = launch_treekeys(c->px, c->py, c->pz, c->n, c->sys().id, b_cur, c->cfg_tscale(), c->tkeys, c->tperm, s);
    if (e == hipSuccess) e = sort_pairs(c, c->tkeys, c->tperm, s);
    if (e == hipSuccess) e = launch_tree_gather(c, s);
    return e;
}

// One frame in the reference order with no host wait: the tree is built into
// the arrays it has (an overflow is flagged on the device, the walk then does
// nothing), and the side branch computing the next bounds is joined at the
// end so one frame is one closed unit of work.
int enqueue_step_async(bhx_ctx* c, hipStream_t s, bool linear) {
    const int cur = c->bslot;
    double* b_cur = c->bounds_slot(cur);
    double* b_next = c->bounds_slot(cur ^ 1);
    const SysView sv = c->sys();
    const double dt = c->cfg.dt;
    // accelerations that move with their bodies make the drift independent of
    // the body sort: the tree branch then runs beside it
    const bool branch = !linear && c->cfg.permute_acc;
    if (branch) HIPCHK(fork_tree_branch(c, b_cur, dt, s));
    HIPCHK(launch_morton(c->px, c->py, c->pz, c->n, sv.id, b_cur, c->keys, c->vals, s, &c->ctr->nonfinite));
    HIPCHK(sort_pairs(c, c->keys, c->vals, s));
    HIPCHK(launch_permute_kick_drift(c, c->vals, dt, s, branch ? c->inv : nullptr));
    if (linear) {
        HIPCHK(launch_bounds(c, b_next, s));
    } else {
        HIPCHK(hipEventRecord(c->ev_drift, s));
        HIPCHK(hipStreamWaitEvent(c->s_side, c->ev_drift, 0));
        HIPCHK(launch_bounds(c, b_next, c->s_side));
        HIPCHK(hipEventRecord(c->ev_bounds, c->s_side));
    }
    HIPCHK(tree_order(c, b_cur, branch, s));
    size_t total = 0;
    HIPCHK(build_tree(c, s, &total, false));
    HIPCHK(launch_walk_kick(c, dt, s, &c->ctr->nonfinite));
    if (!linear) HIPCHK(hipStreamWaitEvent(s, c->ev_bounds, 0));
    c->bslot = cur ^ 1;
    return BHX_OK;
}

void buffer_set(const bhx_ctx* c, const double** out) {
    const double* a[16] = {c->px, c->py, c->pz, c->vx, c->vy, c->vz, c->m, c->cpx, c->cpy, c->cpz,
                           c->cvx, c->cvy, c->cvz, c->ax, c->ay, c->az};
    for (int k = 0; k < 16; ++k) {
        out[k] = a[k];
        out[16 + k] = c->alt[k];
    }
}

// Captures one frame starting from the current buffer parity into `g`.
int capture_step(bhx_ctx* c, bhx_ctx::StepGraph& g, bool linear) {
    buffer_set(c, g.bufs);
    g.bslot = c->bslot;
    g.linear = linear;
    HIPCHK(hipStreamBeginCapture(c->s_main, hipStreamCaptureModeRelaxed));
    const int rc = enqueue_step_async(c, c->s_main, linear);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->s_main, &graph);
    if (rc != BHX_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc != BHX_OK ? rc : fail(c, BHX_E_HIP, "hipStreamEndCapture", e);
    }
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        g.exec = nullptr;
        return fail(c, BHX_E_HIP, "hipGraphInstantiate", ei);
    }
    return BHX_OK;
}

// The 16 state arrays of the bodies (positions, velocities, masses,
// compensation terms, accelerations) and their gather twins.
double** state_ptrs(bhx_ctx* c, double** out) {
    double** a[16] = {&c->px, &c->py, &c->pz, &c->vx, &c->vy, &c->vz, &c->m, &c->cpx, &c->cpy, &c->cpz,
                      &c->cvx, &c->cvy, &c->cvz, &c->ax, &c->ay, &c->az};
    for (int k = 0; k < 16; ++k) out[k] = *a[k];
    return out;
}

int take_snapshot(bhx_ctx* c, bhx_ctx::Snapshot& sn, hipStream_t s) {
    const size_t len = 16 * c->cap + 12 * c->sys_cap;
    if (!sn.buf || sn.len < len) {
        if (sn.buf) (void)hipFree(sn.buf);
        sn.buf = nullptr;
        sn.len = 0;
        HIPCHK(dalloc(&sn.buf, len));
        sn.len = len;
    }
    double* cur[16];
    state_ptrs(c, cur);
    for (int k = 0; k < 16; ++k)
        HIPCHK(hipMemcpyAsync(sn.buf + k * c->cap, cur[k], c->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(sn.buf + 16 * c->cap, c->bounds, 12 * c->sys_cap * sizeof(double), hipMemcpyDeviceToDevice,
                          s));
    for (int k = 0; k < 13; ++k) {
        sn.ptr[k] = cur[k];
        sn.ptr[13 + k] = c->alt[k];
    }
    sn.bslot = c->bslot;
    sn.bounds_valid = c->bounds_valid;
    return BHX_OK;
}

int restore_snapshot(bhx_ctx* c, const bhx_ctx::Snapshot& sn, hipStream_t s) {
    double** a[13] = {&c->px, &c->py, &c->pz, &c->vx, &c->vy, &c->vz, &c->m, &c->cpx, &c->cpy, &c->cpz,
                      &c->cvx, &c->cvy, &c->cvz};
    for (int k = 0; k < 13; ++k) {
        *a[k] = sn.ptr[k];
        c->alt[k] = sn.ptr[13 + k];
    }
    c->bslot = sn.bslot;
    c->bounds_valid = sn.bounds_valid;
    double* cur[16];
    state_ptrs(c, cur);
    for (int k = 0; k < 16; ++k)
        HIPCHK(hipMemcpyAsync(cur[k], sn.buf + k * c->cap, c->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(c->bounds, sn.buf + 16 * c->cap, 12 * c->sys_cap * sizeof(double), hipMemcpyDeviceToDevice,
                          s));
    c->tree_valid = c->tree_built = false;
    return BHX_OK;
}

// One frame from its captured graph; the host mirrors the buffer exchange
// and bounds-slot flip the frame performs on the device.
int graph_frame(bhx_ctx* c, hipStream_t s, bool linear) {
    hipGraphExec_t g = nullptr;
    const int rc = step_graph_for(c, &g, linear);
    if (rc != BHX_OK) return rc;
    HIPCHK(hipGraphLaunch(g, s));
    c->hash_valid = false;   // the frame built a new tree
    swap_body_buffers(c);
    c->tree_slot = c->bslot;
    c->bslot ^= 1;
    return BHX_OK;
}

// Graph of the frame that starts from the current parity (capturing both
// parities when none matches).
int step_graph_for(bhx_ctx* c, hipGraphExec_t* out, bool linear) {
    const double* cur[32];
    buffer_set(c, cur);
    for (int k = 0; k < 4; ++k) {
        const auto& g = c->graphs[k];
        if (g.exec && g.linear == linear && g.bslot == c->bslot && std::equal(cur, cur + 32, g.bufs)) {
            c->graph_pair_next = (k >> 1) ^ 1;   // keep this pair, replace the other
            *out = g.exec;
            return BHX_OK;
        }
    }
    const int pair = c->graph_pair_next;
    auto drop_pair = [&]() {
        for (int k = 2 * pair; k < 2 * pair + 2; ++k) {
            if (c->graphs[k].exec) (void)hipGraphExecDestroy(c->graphs[k].exec);
            c->graphs[k] = bhx_ctx::StepGraph{};
        }
    };
    drop_pair();
    int rc = capture_step(c, c->graphs[2 * pair], linear);               // leaves the host state one frame on
    if (rc == BHX_OK) rc = capture_step(c, c->graphs[2 * pair + 1], linear);   // and back
    if (rc != BHX_OK) {
        drop_pair();
        return rc;
    }
    c->graph_pair_next = pair ^ 1;
    c->graph_valid = true;
    *out = c->graphs[2 * pair].exec;
    return BHX_OK;
}

// Temporary device buffers of one API call, freed when the call returns.
struct DevTmp {
    std::vector<void*> ptrs;
    ~DevTmp() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    hipError_t get(T** out, size_t count) {
        *out = nullptr;
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, (count ? count : 1) * sizeof(T));
        if (e == hipSuccess) {
            ptrs.push_back(p);
            *out = static_cast<T*>(p);
        }
        return e;
    }
};

// The per-body tree API needs the node table of a one-system build.
int need_tree_table(bhx_ctx* c) {
    if (!c->tree_built) return fail(c, BHX_E_STATE, "no tree built");
    if (c->nsys > 1) return fail(c, BHX_E_STATE, "several systems: the per-body tree API serves one system");
    return BHX_OK;
}

// Level of a node key (sentinel bit at 3 * level), or -1 when it is not a key.
int key_level(uint64_t k) {
    if (k == 0) return -1;
    const int top = 63 - __builtin_clzll(k);
    if (top % 3 != 0 || top / 3 > kMaxLevel) return -1;
    return top / 3;
}

}  // namespace bhx

extern "C" {

int bhx_config_default(bhx_config* cfg) {
    if (!cfg) return BHX_E_NULL;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->theta = 0.5;
    cfg->eps = 0.025;
    cfg->dt = 1e-3;
    cfg->leaf_capacity = 1;
    cfg->mode = BHX_MODE_AS_RUN;
    cfg->precision = BHX_FP64;
    cfg->quadrupole = 0;
    cfg->permute_acc = 1;
    cfg->compensated = 0;
    cfg->device = 0;
    cfg->geometry = 0;
    return BHX_OK;
}

int bhx_create(const bhx_config* cfg, size_t capacity, bhx_ctx** out) {
    if (!cfg || !out) return BHX_E_NULL;
    *out = nullptr;
    if (check_cfg(cfg) != BHX_OK) return BHX_E_ARG;
    if (capacity == 0 || capacity >= (1ull << 30)) return BHX_E_ARG;
    bhx_ctx* c = new bhx_ctx();
    c->cfg = *cfg;
    c->device = cfg->device;
    c->cap = capacity;
    c->walk_sub = walk_lists_for(cfg);
    int rc = BHX_OK;
    auto chk = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && rc == BHX_OK) rc = fail(c, e == hipErrorOutOfMemory ? BHX_E_NOMEM : BHX_E_HIP, what, e);
    };
    chk(hipSetDevice(c->device), "hipSetDevice");
    if (rc == BHX_OK) {
        chk(hipStreamCreateWithFlags(&c->s_main, hipStreamNonBlocking), "stream");
        chk(hipStreamCreateWithFlags(&c->s_side, hipStreamNonBlocking), "stream");
        chk(hipStreamCreateWithFlags(&c->s_adj, hipStreamNonBlocking), "stream");
        chk(hipStreamCreateWithFlags(&c->s_tree, hipStreamNonBlocking), "stream");
        chk(hipEventCreateWithFlags(&c->ev_drift, hipEventDisableTiming), "event");
        chk(hipEventCreateWithFlags(&c->ev_bounds, hipEventDisableTiming), "event");
        chk(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming), "event");
        chk(hipEventCreateWithFlags(&c->ev_tree, hipEventDisableTiming), "event");
        for (hipEvent_t& e : c->ev) chk(hipEventCreate(&e), "event");
        const size_t n = capacity;
        double** dp[] = {&c->px, &c->py, &c->pz, &c->vx, &c->vy, &c->vz, &c->m, &c->ax, &c->ay, &c->az,
                         &c->cpx, &c->cpy, &c->cpz, &c->cvx, &c->cvy, &c->cvz, &c->tx, &c->ty, &c->tz, &c->tm};
        for (double** p : dp) chk(dalloc(p, n), "alloc bodies");
        for (double*& p : c->alt) chk(dalloc(&p, n), "alloc bodies");
        chk(dalloc(&c->stage, 3 * n + 8 * 256), "alloc stage");
        chk(dalloc(&c->keys, n), "alloc keys");
        chk(dalloc(&c->keys_alt, n), "alloc keys");
        chk(dalloc(&c->vals, n), "alloc vals");
        chk(dalloc(&c->vals_alt, n), "alloc vals");
        chk(dalloc(&c->mkeys, n), "alloc keys");
        chk(dalloc(&c->tkeys, n), "alloc keys");
        chk(dalloc(&c->tperm, n), "alloc perm");
        chk(dalloc(&c->tb4, n), "alloc bodies");
        chk(dalloc(&c->tleaf, n), "alloc bodies");
        chk(dalloc(&c->nn, n + 1), "alloc build");
        chk(dalloc(&c->noff, n + 1), "alloc build");
        chk(dalloc(&c->nlo, n + 1), "alloc build");
        chk(dalloc(&c->scan_tmp, n / 2048 + 16), "alloc scan");
        chk(dalloc(&c->wcount, n), "alloc counts");
        chk(dalloc(&c->wq, 64), "alloc work counters");
        chk(dalloc(&c->digest, 4), "alloc digest");
        c->sort_tiles = sort_tiles_for(n);
        chk(dalloc(&c->sort_hist, 8 * 256), "alloc sort");
        chk(dalloc(&c->sort_tile_ctr, 8), "alloc sort");
        chk(dalloc(&c->sort_status, 8 * 256 * (c->sort_tiles + 1)), "alloc sort");
        chk(dalloc(&c->tperm_old, n), "alloc perm");
        chk(dalloc(&c->inv, n), "alloc perm");
        chk(dalloc(&c->sort2.keys_alt, n), "alloc keys");
        chk(dalloc(&c->sort2.vals_alt, n), "alloc vals");
        chk(dalloc(&c->sort2.hist, 8 * 256), "alloc sort");
        chk(dalloc(&c->sort2.tile_ctr, 8), "alloc sort");
        chk(dalloc(&c->sort2.status, 8 * 256 * (c->sort_tiles + 1)), "alloc sort");
        c->sum_tiles_cap = sum_tiles_for(n);
        chk(dalloc(&c->sum_tile, 6 * c->sum_tiles_cap), "alloc bounds");
        chk(dalloc(&c->sum_cnt, 6 * c->sum_tiles_cap + 8), "alloc bounds");
        chk(dalloc(&c->sum_piece, 6 * 64 * c->sum_tiles_cap), "alloc bounds");
        chk(dalloc(&c->ctr, 1), "alloc counters");
        chk(hipHostMalloc((void**)&c->ctr_host, sizeof(Counters), hipHostMallocDefault), "alloc pinned");
        if (rc == BHX_OK) chk(alloc_tree(c, 2 * n + 1024), "alloc tree");
        if (rc == BHX_OK) rc = set_layout(c, std::vector<uint32_t>{0u, 0u});
        if (rc == BHX_OK) chk(hipMemsetAsync(c->ctr, 0, sizeof(Counters), c->s_main), "memset");
        if (rc == BHX_OK) chk(hipMemsetAsync(c->digest, 0, 4 * sizeof(unsigned long long), c->s_main), "memset");
        if (rc == BHX_OK) chk(hipStreamSynchronize(c->s_main), "sync");
    }
    if (rc != BHX_OK) {
        free_all(c);
        delete c;
        return rc;
    }
    *out = c;
    return BHX_OK;
}

int bhx_destroy(bhx_ctx* c) {
    if (!c) return BHX_E_NULL;
    (void)hipSetDevice(c->device);
    if (c->s_main) (void)hipStreamSynchronize(c->s_main);
    if (c->s_side) (void)hipStreamSynchronize(c->s_side);
    free_all(c);
    delete c;
    return BHX_OK;
}

int bhx_set_config(bhx_ctx* c, const bhx_config* cfg) {
    if (!c || !cfg) return BHX_E_NULL;
    if (check_cfg(cfg) != BHX_OK) return fail(c, BHX_E_ARG, "invalid config");
    if (cfg->device != c->device) return fail(c, BHX_E_ARG, "device cannot change");
    const bool tree_shape = cfg->leaf_capacity != c->cfg.leaf_capacity || cfg->mode != c->cfg.mode ||
                            cfg->geometry != c->cfg.geometry;
    const bhx_config& o = c->cfg;
    const bool same = cfg->theta == o.theta && cfg->eps == o.eps && cfg->dt == o.dt &&
                      cfg->leaf_capacity == o.leaf_capacity && cfg->mode == o.mode && cfg->precision == o.precision &&
                      cfg->quadrupole == o.quadrupole && cfg->permute_acc == o.permute_acc &&
                      cfg->compensated == o.compensated && cfg->group_walk == o.group_walk &&
                      cfg->geometry == o.geometry && cfg->group_size == o.group_size;
    // a recording's backward pass re-runs its frames: they must be the recorded ones
    if (!same) free_grad(c);
    c->cfg = *cfg;
    c->walk_sub = walk_lists_for(cfg);
    if (tree_shape) c->tree_valid = false;
    drop_graphs(c);   // captured launches hold the old parameters
    return BHX_OK;
}

int bhx_get_config(const bhx_ctx* c, bhx_config* cfg) {
    if (!c || !cfg) return BHX_E_NULL;
    *cfg = c->cfg;
    return BHX_OK;
}

const char* bhx_last_error(const bhx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int bhx_sync(bhx_ctx* c) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->s_side));
    HIPCHK(hipStreamSynchronize(c->s_main));
    return BHX_OK;
}

int bhx_set_bodies(bhx_ctx* c, const double* pos, const double* vel, const double* mass, size_t n) {
    if (!c || !pos || !mass) return c ? fail(c, BHX_E_NULL, "bodies is a null pointer") : BHX_E_NULL;
    if (n > c->cap) return fail(c, BHX_E_CAPACITY, "more bodies than the context capacity");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    c->n = n;
    if (n) {
        HIPCHK(hipMemcpyAsync(c->stage, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(n)), dim3(kBlock), 0, s, c->stage, n, c->px, c->py, c->pz);
        if (vel) {
            HIPCHK(hipMemcpyAsync(c->stage, vel, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(n)), dim3(kBlock), 0, s, c->stage, n, c->vx, c->vy, c->vz);
        } else {
            double* z3[] = {c->vx, c->vy, c->vz};
            for (double* p : z3) hipLaunchKernelGGL(k_zero_d, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n);
        }
        HIPCHK(hipMemcpyAsync(c->m, mass, n * sizeof(double), hipMemcpyHostToDevice, s));
        double* zs[] = {c->ax, c->ay, c->az, c->cpx, c->cpy, c->cpz, c->cvx, c->cvy, c->cvz};
        for (double* p : zs) hipLaunchKernelGGL(k_zero_d, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    int rc = set_layout(c, std::vector<uint32_t>{0u, (uint32_t)n});
    if (rc != BHX_OK) return rc;
    c->bounds_valid = false;
    c->tree_valid = false;
    return BHX_OK;
}

int bhx_set_systems(bhx_ctx* c, const uint64_t* offsets, size_t nsys) {
    if (!c || !offsets) return c ? fail(c, BHX_E_NULL, "offsets is a null pointer") : BHX_E_NULL;
    if (nsys == 0 || nsys > kMaxSystems) return fail(c, BHX_E_ARG, "system count out of range");
    if (offsets[0] != 0 || offsets[nsys] != c->n) return fail(c, BHX_E_ARG, "offsets must run from 0 to the body count");
    std::vector<uint32_t> off(nsys + 1);
    for (size_t k = 0; k <= nsys; ++k) {
        if (k > 0 && offsets[k] <= offsets[k - 1]) return fail(c, BHX_E_ARG, "every system needs at least one body");
        off[k] = (uint32_t)offsets[k];
    }
    HIPCHK(hipSetDevice(c->device));
    return set_layout(c, off);
}

int bhx_num_systems(const bhx_ctx* c, size_t* nsys) {
    if (!c || !nsys) return BHX_E_NULL;
    *nsys = c->nsys;
    return BHX_OK;
}

int bhx_get_systems(const bhx_ctx* c, uint64_t* offsets) {
    if (!c || !offsets) return BHX_E_NULL;
    if (c->sys_off_h.size() != (size_t)c->nsys + 1) {
        offsets[0] = 0;
        offsets[1] = c->n;
        return BHX_OK;
    }
    for (size_t k = 0; k <= (size_t)c->nsys; ++k) offsets[k] = c->sys_off_h[k];
    return BHX_OK;
}

int bhx_system_bounds(bhx_ctx* c, double* centers, double* sizes) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    double* b4 = c->bounds_slot(c->bslot);
    if (c->n) HIPCHK(launch_bounds(c, b4, s));
    std::vector<double> h(4 * (size_t)c->nsys);
    HIPCHK(hipMemcpyAsync(h.data(), b4, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    c->bounds_valid = c->n > 0;
    for (uint32_t k = 0; k < c->nsys; ++k) {
        if (centers) for (int a = 0; a < 3; ++a) centers[3 * (size_t)k + a] = h[4 * (size_t)k + a];
        if (sizes) sizes[k] = h[4 * (size_t)k + 3];
    }
    return BHX_OK;
}

int bhx_system_roots(bhx_ctx* c, uint64_t* roots) {
    if (!c || !roots) return c ? fail(c, BHX_E_NULL, "roots is a null pointer") : BHX_E_NULL;
    if (!c->tree_built) return fail(c, BHX_E_STATE, "no tree built");
    HIPCHK(hipSetDevice(c->device));
    std::vector<uint32_t> h(c->nsys);
    HIPCHK(hipMemcpyAsync(h.data(), c->sys_root, h.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, c->s_main));
    HIPCHK(hipStreamSynchronize(c->s_main));
    for (uint32_t k = 0; k < c->nsys; ++k) roots[k] = h[k];
    return BHX_OK;
}

int bhx_set_accelerations(bhx_ctx* c, const double* acc) {
    if (!c || !acc) return c ? fail(c, BHX_E_NULL, "acc is a null pointer") : BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    if (c->n) {
        HIPCHK(hipMemcpyAsync(c->stage, acc, 3 * c->n * sizeof(double), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->stage, c->n, c->ax, c->ay, c->az);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_get_bodies(bhx_ctx* c, double* pos, double* vel, double* mass, double* acc) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    const size_t n = c->n;
    if (!n) return BHX_OK;
    struct { double* dst; double *x, *y, *z; } outs[] = {{pos, c->px, c->py, c->pz}, {vel, c->vx, c->vy, c->vz},
                                                         {acc, c->ax, c->ay, c->az}};
    for (auto& o : outs) {
        if (!o.dst) continue;
        hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(n)), dim3(kBlock), 0, s, o.x, o.y, o.z, n, c->stage);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(o.dst, c->stage, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if (mass) HIPCHK(hipMemcpyAsync(mass, c->m, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_get_compensation(bhx_ctx* c, double* pos_c, double* vel_c) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    const size_t n = c->n;
    if (!n) return BHX_OK;
    struct { double* dst; double *x, *y, *z; } outs[] = {{pos_c, c->cpx, c->cpy, c->cpz},
                                                         {vel_c, c->cvx, c->cvy, c->cvz}};
    for (auto& o : outs) {
        if (!o.dst) continue;
        hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(n)), dim3(kBlock), 0, s, o.x, o.y, o.z, n, c->stage);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(o.dst, c->stage, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    return BHX_OK;
}

int bhx_set_compensation(bhx_ctx* c, const double* pos_c, const double* vel_c) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    const size_t n = c->n;
    if (!n) return BHX_OK;
    struct { const double* src; double *x, *y, *z; } ins[] = {{pos_c, c->cpx, c->cpy, c->cpz},
                                                              {vel_c, c->cvx, c->cvy, c->cvz}};
    for (auto& o : ins) {
        if (o.src) {
            HIPCHK(hipMemcpyAsync(c->stage, o.src, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for(n)), dim3(kBlock), 0, s, c->stage, n, o.x, o.y, o.z);
        } else {
            double* z3[] = {o.x, o.y, o.z};
            for (double* p : z3) hipLaunchKernelGGL(k_zero_d, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    }
    return BHX_OK;
}

int bhx_num_bodies(const bhx_ctx* c, size_t* n) {
    if (!c || !n) return BHX_E_NULL;
    *n = c->n;
    return BHX_OK;
}

int bhx_bounds(bhx_ctx* c, double center[3], double* size) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    if (c->nsys > 1) return fail(c, BHX_E_STATE, "several systems: use bhx_system_bounds");
    double* b4 = c->bounds_slot(c->bslot);
    HIPCHK(launch_bounds(c, b4, s));
    double h[4];
    HIPCHK(hipMemcpyAsync(h, b4, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    c->bounds_valid = true;
    if (center) { center[0] = h[0]; center[1] = h[1]; center[2] = h[2]; }
    if (size) *size = h[3];
    return BHX_OK;
}

int bhx_bounds_path(bhx_ctx* c, uint32_t out[4]) {
    if (!c || !out) return c ? fail(c, BHX_E_NULL, "out is a null pointer") : BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->s_side));
    HIPCHK(hipStreamSynchronize(c->s_main));
    // the gate words follow the per-tile piece counts and their prefixes
    const uint32_t* gate = c->sum_cnt + 6 * c->sum_tiles_cap;
    HIPCHK(hipMemcpy(out, gate, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return BHX_OK;
}

int bhx_morton_keys(bhx_ctx* c, double size, uint64_t* keys_out) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    if (c->nsys > 1) return fail(c, BHX_E_STATE, "several systems: keys follow each system's bounds in bhx_step");
    double h[4] = {0.0, 0.0, 0.0, size};
    double* b4 = c->bounds_slot(2);      // scratch slot
    HIPCHK(hipMemcpyAsync(b4, h, sizeof h, hipMemcpyHostToDevice, s));
    HIPCHK(launch_morton(c->px, c->py, c->pz, c->n, nullptr, b4, c->mkeys, nullptr, s));
    if (keys_out && c->n) HIPCHK(hipMemcpyAsync(keys_out, c->mkeys, c->n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_sort_keys(bhx_ctx* c, uint64_t* keys, uint64_t* idx, size_t n) {
    if (!c || !keys || !idx) return c ? fail(c, BHX_E_NULL, "keys is a null pointer") : BHX_E_NULL;
    if (n > c->cap) return fail(c, BHX_E_CAPACITY, "more keys than the context capacity");
    if (n <= 1) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    HIPCHK(hipMemcpyAsync(c->keys, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    HIPCHK(launch_iota(c->vals, n, s));
    HIPCHK(radix_sort_pairs(c->keys, c->vals, c->keys_alt, c->vals_alt, n, c->sort_hist, c->sort_status,
                            c->sort_tile_ctr, c->ctr, s));
    std::vector<uint32_t> perm(n);
    HIPCHK(hipMemcpyAsync(keys, c->keys, n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(perm.data(), c->vals, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    int rc = read_counters(c, s);
    if (rc != BHX_OK) return rc;
    std::vector<uint64_t> old(idx, idx + n);
    for (size_t i = 0; i < n; ++i) idx[i] = old[perm[i]];
    return BHX_OK;
}

// ReorderBodiesByIndexes (Body.hpp:257-273) on the resident arrays: gather into
// the ping-pong buffers, then swap.  Accelerations follow only when
// permute_acc is set (the reference leaves them in place, ofApp.cpp:176).
static int reorder_by(bhx_ctx* c, const uint32_t* perm_dev, hipStream_t s) {
    double** all[16] = {&c->px, &c->py, &c->pz, &c->vx, &c->vy, &c->vz, &c->m, &c->cpx,
                        &c->cpy, &c->cpz, &c->cvx, &c->cvy, &c->cvz, &c->ax, &c->ay, &c->az};
    int slots[16];
    int na = 0;
    for (int k = 0; k < 7; ++k) slots[na++] = k;
    if (c->cfg.compensated)
        for (int k = 7; k < 13; ++k) slots[na++] = k;
    if (c->cfg.permute_acc)
        for (int k = 13; k < 16; ++k) slots[na++] = k;
    double* src[16];
    double* dst[16];
    for (int k = 0; k < na; ++k) {
        src[k] = *all[slots[k]];
        dst[k] = c->alt[slots[k]];
    }
    HIPCHK(launch_gather(src, dst, na, perm_dev, c->n, s));
    for (int k = 0; k < na; ++k) {
        const int sl = slots[k];
        double* t = *all[sl];
        *all[sl] = c->alt[sl];
        c->alt[sl] = t;
    }
    c->bounds_valid = false;
    c->tree_valid = false;
    return BHX_OK;
}

int bhx_sort_bodies(bhx_ctx* c, uint64_t* perm_out) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    if (c->n == 0) return BHX_OK;
    hipLaunchKernelGGL(k_copy_u64, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->mkeys, c->n, c->keys);
    HIPCHK(launch_iota(c->vals, c->n, s));
    HIPCHK(sort_pairs(c, c->keys, c->vals, s));
    int rc = reorder_by(c, c->vals, s);
    if (rc != BHX_OK) return rc;
    if (perm_out) {
        uint64_t* tmp = reinterpret_cast<uint64_t*>(c->stage);
        hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->vals, c->n, tmp);
        HIPCHK(hipMemcpyAsync(perm_out, tmp, c->n * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    }
    return read_counters(c, s);
}

int bhx_reorder_bodies(bhx_ctx* c, const uint64_t* perm) {
    if (!c || !perm) return c ? fail(c, BHX_E_NULL, "indexes is a null pointer") : BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    if (c->n == 0) return BHX_OK;
    for (size_t i = 0; i < c->n; ++i)
        if (perm[i] >= c->n) return fail(c, BHX_E_ARG, "index out of range");
    if (c->nsys > 1)   // bodies stay inside their system's range
        for (uint32_t k = 0; k < c->nsys; ++k)
            for (size_t i = c->sys_off_h[k]; i < c->sys_off_h[k + 1]; ++i)
                if (perm[i] < c->sys_off_h[k] || perm[i] >= c->sys_off_h[k + 1])
                    return fail(c, BHX_E_ARG, "permutation moves a body to another system");
    uint64_t* tmp = reinterpret_cast<uint64_t*>(c->stage);
    HIPCHK(hipMemcpyAsync(tmp, perm, c->n * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_u64_to_u32, dim3(grid_for(c->n)), dim3(kBlock), 0, s, tmp, c->n, c->vals_alt);
    int rc = reorder_by(c, c->vals_alt, s);
    if (rc != BHX_OK) return rc;
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_kick_drift(bhx_ctx* c, double dt) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->s_side));
    if (c->n) HIPCHK(launch_kick_drift(c, dt, c->s_main));
    HIPCHK(hipStreamSynchronize(c->s_main));
    c->bounds_valid = false;
    c->tree_valid = false;
    return BHX_OK;
}

int bhx_kick(bhx_ctx* c, double dt) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    if (c->n) HIPCHK(launch_kick(c, dt, c->s_main));
    HIPCHK(hipStreamSynchronize(c->s_main));
    return BHX_OK;
}

int bhx_build(bhx_ctx* c, const double center[3], double size, size_t* n_nodes) {
    if (!c || !center) return BHX_E_NULL;
    if (!std::isfinite(size) || !std::isfinite(center[0]) || !std::isfinite(center[1]) || !std::isfinite(center[2]))
        return fail(c, BHX_E_NONFINITE, "non-finite bounds");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    if (c->nsys > 1) return fail(c, BHX_E_STATE, "several systems: use bhx_build_systems");
    double h[4] = {center[0], center[1], center[2], size};
    HIPCHK(hipMemcpyAsync(c->bounds_slot(c->bslot), h, sizeof h, hipMemcpyHostToDevice, s));
    c->bounds_valid = false;
    int rc = build_on_device(c, s);
    if (rc != BHX_OK) return rc;
    rc = read_counters(c, s);
    if (rc != BHX_OK) return rc;
    c->tree_bounds_h.assign(h, h + 4);
    if (n_nodes) *n_nodes = c->n_nodes;
    return BHX_OK;
}

int bhx_build_systems(bhx_ctx* c, const double* centers, const double* sizes, size_t* n_nodes) {
    if (!c || !centers || !sizes) return c ? fail(c, BHX_E_NULL, "bounds is a null pointer") : BHX_E_NULL;
    std::vector<double> h(4 * (size_t)c->nsys);
    for (uint32_t k = 0; k < c->nsys; ++k) {
        for (int a = 0; a < 3; ++a) h[4 * (size_t)k + a] = centers[3 * (size_t)k + a];
        h[4 * (size_t)k + 3] = sizes[k];
    }
    for (double v : h)
        if (!std::isfinite(v)) return fail(c, BHX_E_NONFINITE, "non-finite bounds");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    HIPCHK(hipMemcpyAsync(c->bounds_slot(c->bslot), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
    c->bounds_valid = false;
    int rc = build_on_device(c, s);
    if (rc != BHX_OK) return rc;
    rc = read_counters(c, s);
    if (rc != BHX_OK) return rc;
    c->tree_bounds_h = h;
    if (n_nodes) *n_nodes = c->n_nodes;
    return BHX_OK;
}

int bhx_get_nodes(bhx_ctx* c, bhx_node* out, size_t cap, size_t* n_nodes) {
    if (!c) return BHX_E_NULL;
    if (!c->tree_built) return fail(c, BHX_E_STATE, "no tree built");
    if (n_nodes) *n_nodes = c->n_nodes;
    if (!out) return BHX_OK;
    const size_t m = c->n_nodes < cap ? c->n_nodes : cap;
    if (m == 0) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    const DeviceTree& t = c->tree;
    std::vector<uint64_t> key(m);
    std::vector<uint32_t> start(m), count(m), skip(m), cb(m), level(m), first(m);
    std::vector<double> center(3 * m), com(3 * m), mass(m), quad(6 * m);
    HIPCHK(hipMemcpyAsync(key.data(), t.key, m * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(start.data(), t.start, m * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(count.data(), t.count, m * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(skip.data(), t.skip, m * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(cb.data(), t.child_byte, m * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(level.data(), t.level, m * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(first.data(), t.first, m * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(center.data(), t.center, m * 24, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(com.data(), t.com, m * 24, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(mass.data(), t.mass, m * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(quad.data(), t.quad, m * 48, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t i = 0; i < m; ++i) {
        bhx_node& o = out[i];
        o.key = key[i];
        for (int a = 0; a < 3; ++a) { o.center[a] = center[3 * i + a]; o.com[a] = com[3 * i + a]; }
        // the node's system: last offset <= its first body
        const size_t k = (size_t)(std::upper_bound(c->sys_off_h.begin(), c->sys_off_h.end() - 1, start[i]) -
                                  c->sys_off_h.begin()) - 1;
        double sz = 4 * k + 3 < c->tree_bounds_h.size() ? c->tree_bounds_h[4 * k + 3] * c->tree_tscale : 0.0;
        for (uint32_t l = 0; l < level[i]; ++l) sz = sz * 0.5;
        o.size = sz;
        o.mass = mass[i];
        for (int q = 0; q < 6; ++q) o.quad[q] = quad[6 * i + q];
        o.n = count[i];
        o.first = first[i];
        o.level = (int32_t)level[i];
        o.child_byte = cb[i];
        o.skip = skip[i];
        o.body_start = start[i];
    }
    return BHX_OK;
}

int bhx_lookup(bhx_ctx* c, const uint64_t* keys, size_t n, int64_t* ids) { return bhx_lookup_in(c, 0, keys, n, ids); }

int bhx_lookup_in(bhx_ctx* c, size_t system, const uint64_t* keys, size_t n, int64_t* ids) {
    if (!c || !keys || !ids) return c ? fail(c, BHX_E_NULL, "null argument") : BHX_E_NULL;
    if (!c->tree_built) return fail(c, BHX_E_STATE, "no tree built");
    if (system >= c->nsys) return fail(c, BHX_E_ARG, "system out of range");
    if (n == 0) return BHX_OK;
    if (c->n_nodes == 0) {
        for (size_t i = 0; i < n; ++i) ids[i] = -1;
        return BHX_OK;
    }
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    uint64_t* dk = nullptr;
    int64_t* di = nullptr;
    HIPCHK(hipMalloc((void**)&dk, n * sizeof(uint64_t)));
    hipError_t e = hipMalloc((void**)&di, n * sizeof(int64_t));
    if (e == hipSuccess) e = hipMemcpyAsync(dk, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_lookup(c, (uint32_t)system, dk, n, di, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ids, di, n * sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dk);
    if (di) (void)hipFree(di);
    if (e != hipSuccess) return fail(c, BHX_E_HIP, "lookup", e);
    return BHX_OK;
}

int bhx_get_tree_order(bhx_ctx* c, uint64_t* tree_perm, uint64_t* tree_keys) {
    if (!c) return BHX_E_NULL;
    if (!c->tree_built) return fail(c, BHX_E_STATE, "no tree built");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    const size_t n = c->tree_n;
    if (n == 0) return BHX_OK;
    if (tree_perm) {
        uint64_t* tmp = reinterpret_cast<uint64_t*>(c->stage);
        hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(n)), dim3(kBlock), 0, s, c->tperm, n, tmp);
        HIPCHK(hipMemcpyAsync(tree_perm, tmp, n * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if (tree_keys) HIPCHK(hipMemcpyAsync(tree_keys, c->tkeys, n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_unsplit_leaves(bhx_ctx* c, size_t* count) {
    if (!c || !count) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    int rc = read_counters(c, c->s_main);
    if (rc != BHX_OK) return rc;
    *count = c->ctr_host->unsplit;
    return BHX_OK;
}

int bhx_walk(bhx_ctx* c, double* acc_out, uint64_t* counts) {
    if (!c) return BHX_E_NULL;
    if (c->tree_edited) return fail(c, BHX_E_STATE, "nodes were put or erased since the build: rebuild, or use "
                                                    "bhx_interaction_lists and bhx_force_lists");
    if (!c->tree_valid) return fail(c, BHX_E_STATE, "walk before build");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(launch_walk(c, s));
    if (acc_out && c->n) {
        hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->ax, c->ay, c->az, c->n, c->stage);
        HIPCHK(hipMemcpyAsync(acc_out, c->stage, 3 * c->n * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if (counts && c->n) {
        uint64_t* tmp = reinterpret_cast<uint64_t*>(c->stage);
        hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->wcount, c->n, tmp);
        HIPCHK(hipMemcpyAsync(counts, tmp, c->n * 8, hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_interaction_counts(bhx_ctx* c, uint64_t* counts) {
    if (!c || !counts) return c ? fail(c, BHX_E_NULL, "counts is a null pointer") : BHX_E_NULL;
    if (!c->tree_valid) return fail(c, BHX_E_STATE, "no tree built");
    if (c->n == 0) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    // per-body FP64 walk into scratch (the acceleration ping-pong buffers)
    HIPCHK(launch_walk_to(c, s, false, BHX_FP64, c->alt[13], c->alt[14], c->alt[15], c->wcount));
    uint64_t* tmp = reinterpret_cast<uint64_t*>(c->stage);
    hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->wcount, c->n, tmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(counts, tmp, c->n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_interaction_counts_quad(bhx_ctx* c, uint64_t* counts, uint64_t* quad_counts) {
    if (!c || !counts || !quad_counts)
        return c ? fail(c, BHX_E_NULL, "counts is a null pointer") : BHX_E_NULL;
    if (!c->tree_valid) return fail(c, BHX_E_STATE, "no tree built");
    if (c->n == 0) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    // the same per-body walk as bhx_interaction_counts; the quadrupole-term
    // counts go to a stale gather target (n doubles hold n u32), their u64
    // copy to another
    uint32_t* qc = reinterpret_cast<uint32_t*>(c->alt[12]);
    HIPCHK(launch_walk_to(c, s, false, BHX_FP64, c->alt[13], c->alt[14], c->alt[15], c->wcount, nullptr, qc));
    uint64_t* tmp = reinterpret_cast<uint64_t*>(c->stage);
    uint64_t* qtmp = reinterpret_cast<uint64_t*>(c->alt[11]);
    hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(c->n)), dim3(kBlock), 0, s, c->wcount, c->n, tmp);
    hipLaunchKernelGGL(k_u32_to_u64, dim3(grid_for(c->n)), dim3(kBlock), 0, s, (const uint32_t*)qc, c->n, qtmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(counts, tmp, c->n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(quad_counts, qtmp, c->n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_set_profiling(bhx_ctx* c, int on) {
    if (!c) return BHX_E_NULL;
    c->profiling = on != 0;
    return BHX_OK;
}

int bhx_last_timing(bhx_ctx* c, bhx_timing* t) {
    if (!c || !t) return BHX_E_NULL;
    *t = c->timing;
    return BHX_OK;
}

// ofApp::update() (bounds, keys, body sort on x_f — ofApp.cpp:169-177) then
// ofApp::draw() (kick + drift to x_{f+1}, tree on x_{f+1} with the bounds of
// x_f, walk, kick — ofApp.cpp:248-282).  The bounds of x_{f+1}, needed by the
// next frame, are computed on a side stream as soon as the drift is done, in
// parallel with this frame's tree build and walk.
int bhx_step(bhx_ctx* c, int nsteps) {
    if (!c) return BHX_E_NULL;
    if (nsteps < 0) return fail(c, BHX_E_ARG, "negative step count");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    const double dt = c->cfg.dt;
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(clear_nonfinite(c, s));
    for (int step = 0; step < nsteps && c->n; ++step) {
        const int cur = c->bslot;
        double* b_cur = c->bounds_slot(cur);
        double* b_next = c->bounds_slot(cur ^ 1);
        const SysView sv = c->sys();
        const bool prof = c->profiling;
        if (prof) HIPCHK(hipEventRecord(c->ev[0], s));
        if (!c->bounds_valid) {
            HIPCHK(launch_bounds(c, b_cur, s));
        } else {
            HIPCHK(hipStreamWaitEvent(s, c->ev_bounds, 0));
        }
        if (prof) HIPCHK(hipEventRecord(c->ev[1], s));
        // the tree on x_{f+1} with the bounds of x_f: beside the body sort when
        // the accelerations move with their bodies
        const bool branch = c->cfg.permute_acc != 0;
        if (branch) HIPCHK(fork_tree_branch(c, b_cur, dt, s));
        HIPCHK(launch_morton(c->px, c->py, c->pz, c->n, sv.id, b_cur, c->keys, c->vals, s, &c->ctr->nonfinite));
        if (prof) HIPCHK(hipEventRecord(c->ev[2], s));
        HIPCHK(sort_pairs(c, c->keys, c->vals, s));
        if (prof) HIPCHK(hipEventRecord(c->ev[3], s));
        HIPCHK(launch_permute_kick_drift(c, c->vals, dt, s, branch ? c->inv : nullptr));
        HIPCHK(hipEventRecord(c->ev_drift, s));
        if (prof) HIPCHK(hipEventRecord(c->ev[4], s));
        HIPCHK(hipStreamWaitEvent(c->s_side, c->ev_drift, 0));
        HIPCHK(launch_bounds(c, b_next, c->s_side));
        HIPCHK(hipEventRecord(c->ev_bounds, c->s_side));
        c->tree_valid = c->tree_built = false;
        HIPCHK(tree_order(c, b_cur, branch, s));
        if (prof) HIPCHK(hipEventRecord(c->ev[5], s));
        size_t total = 0;
        for (int attempt = 0; attempt < 2; ++attempt) {
            HIPCHK(build_tree(c, s, &total, true));
            if (total <= c->tree.cap) break;
            HIPCHK(hipStreamSynchronize(s));
            HIPCHK(alloc_tree(c, total + total / 4 + 1024));
        }
        if (total > c->tree.cap) return fail(c, BHX_E_NOMEM, "node capacity");
        c->n_nodes = total;
        c->tree_valid = c->tree_built = true;
        c->tree_n = c->n;
        if (prof) HIPCHK(hipEventRecord(c->ev[6], s));
        HIPCHK(launch_walk_kick(c, dt, s, &c->ctr->nonfinite));
        if (prof) HIPCHK(hipEventRecord(c->ev[7], s));
        if (prof) HIPCHK(hipEventRecord(c->ev[8], s));   // the kick is the walk's (group walks)
        c->bslot = cur ^ 1;
        c->bounds_valid = true;
        if (prof) {
            HIPCHK(hipEventSynchronize(c->ev[8]));
            float ms[8];
            for (int k = 0; k < 8; ++k) HIPCHK(hipEventElapsedTime(&ms[k], c->ev[k], c->ev[k + 1]));
            c->timing.bounds = ms[0];
            c->timing.keys = ms[1];
            c->timing.sort = ms[2];
            c->timing.kick_drift = ms[3];
            c->timing.treekeys = ms[4];
            c->timing.build = ms[5];
            c->timing.moments = 0.0;
            c->timing.walk = ms[6];
            c->timing.kick = ms[7];
        }
    }
    int rc = read_counters(c, s);
    HIPCHK(hipStreamSynchronize(c->s_side));
    if (rc != BHX_OK) return rc;
    // keep the build bounds of the last frame for node downloads
    if (c->tree_built) {
        c->tree_bounds_h.resize(4 * (size_t)c->nsys);
        HIPCHK(hipMemcpy(c->tree_bounds_h.data(), c->bounds_slot(c->tree_slot), c->tree_bounds_h.size() * sizeof(double),
                         hipMemcpyDeviceToHost));
    }
    const auto t1 = std::chrono::steady_clock::now();
    c->timing.total = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return BHX_OK;
}

int bhx_step_graph(bhx_ctx* c, int nsteps) {
    if (!c) return BHX_E_NULL;
    if (nsteps < 0) return fail(c, BHX_E_ARG, "negative step count");
    if (nsteps == 0 || c->n == 0) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipStreamSynchronize(c->s_side));
    if (!c->bounds_valid) {
        HIPCHK(launch_bounds(c, c->bounds_slot(c->bslot), s));
        c->bounds_valid = true;
    }
    HIPCHK(clear_nonfinite(c, s));
    int rc = take_snapshot(c, c->snaps[0], s);
    if (rc != BHX_OK) return rc;
    bool done = false;
    for (int attempt = 0; attempt < 3 && !done; ++attempt) {
        hipLaunchKernelGGL(k_fill_u32_api, dim3(1), dim3(64), 0, s, &c->ctr->overflow_seen, (size_t)1, 0u);
        HIPCHK(hipGetLastError());
        c->tree_valid = c->tree_built = false;
        for (int k = 0; k < nsteps; ++k) {
            rc = graph_frame(c, s);
            if (rc != BHX_OK) return rc;
        }
        rc = read_counters(c, s);
        if (rc != BHX_OK) return rc;
        if (!c->ctr_host->overflow_seen) {
            done = true;
            break;
        }
        // some frame needed more nodes than the tree holds: back to the start, grow, run again
        const size_t need = c->ctr_host->max_nodes;
        rc = restore_snapshot(c, c->snaps[0], s);
        if (rc != BHX_OK) return rc;
        HIPCHK(hipStreamSynchronize(s));
        HIPCHK(alloc_tree(c, need + need / 4 + 1024));
    }
    if (!done) return fail(c, BHX_E_NOMEM, "node capacity");
    HIPCHK(hipStreamSynchronize(c->s_side));
    c->n_nodes = c->ctr_host->total_nodes;
    c->tree_valid = c->tree_built = true;
    c->tree_n = c->n;
    c->bounds_valid = true;
    c->tree_bounds_h.resize(4 * (size_t)c->nsys);
    HIPCHK(hipMemcpy(c->tree_bounds_h.data(), c->bounds_slot(c->tree_slot), c->tree_bounds_h.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
    const auto t1 = std::chrono::steady_clock::now();
    c->timing.total = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return BHX_OK;
}

int bhx_autotune(bhx_ctx* c, const int32_t* caps, size_t ncaps, const int32_t* group_sizes, size_t ngroups,
                 int steps, double max_err, size_t nsample, uint64_t seed, bhx_tune_entry* out, int32_t* chosen,
                 int32_t* chosen_group) {
    if (!c || !caps) return c ? fail(c, BHX_E_NULL, "caps is a null pointer") : BHX_E_NULL;
    if (ncaps == 0 || steps < 1 || nsample == 0) return fail(c, BHX_E_ARG, "empty tuning set");
    for (size_t k = 0; k < ncaps; ++k)
        if (caps[k] < 1 || caps[k] > kMaxLeafCap) return fail(c, BHX_E_ARG, "leaf capacity out of range");
    const int32_t own_group = c->cfg.group_size;
    if (!group_sizes || ngroups == 0) {
        group_sizes = &own_group;
        ngroups = 1;
    }
    for (size_t g = 0; g < ngroups; ++g)
        if (group_sizes[g] != 0 && group_sizes[g] != 16 && group_sizes[g] != 32 && group_sizes[g] != 64)
            return fail(c, BHX_E_ARG, "group size must be 16, 32 or 64");
    if (c->n == 0) return fail(c, BHX_E_STATE, "no bodies");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    if (!c->bounds_valid) {
        HIPCHK(launch_bounds(c, c->bounds_slot(c->bslot), s));
        c->bounds_valid = true;
    }
    int rc = take_snapshot(c, c->snaps[1], s);
    if (rc != BHX_OK) return rc;
    // seeded sample of body positions (sorted: the direct sum streams each target's system once per block)
    const size_t ns = nsample < c->n ? nsample : c->n;
    std::vector<int64_t> tg(ns);
    {
        uint64_t st = seed * 0x9e3779b97f4a7c15ull + 1;
        std::vector<char> used(c->n, 0);
        for (size_t i = 0; i < ns;) {
            st += 0x9e3779b97f4a7c15ull;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            z ^= z >> 31;
            const size_t j = (size_t)(z % c->n);
            if (!used[j]) { used[j] = 1; tg[i++] = (int64_t)j; }
        }
        std::sort(tg.begin(), tg.end());
    }
    int64_t* dt = nullptr;
    double *dref = nullptr, *dwalk = nullptr;
    HIPCHK(dalloc(&dt, ns));
    HIPCHK(dalloc(&dref, 3 * ns));
    HIPCHK(dalloc(&dwalk, 3 * ns));
    HIPCHK(hipMemcpyAsync(dt, tg.data(), ns * sizeof(int64_t), hipMemcpyHostToDevice, s));
    const bhx_config base = c->cfg;
    int32_t best = base.leaf_capacity, best_group = base.group_size;
    double best_ms = 0.0;
    bool found = false;
    std::vector<double> ref(3 * ns), walk(3 * ns), err(ns);
    for (size_t kk = 0; kk < ncaps * ngroups && rc == BHX_OK; ++kk) {
        const size_t k = kk / ngroups, g = kk % ngroups;
        bhx_config cfg = base;
        cfg.leaf_capacity = caps[k];
        cfg.group_size = group_sizes[g];
        rc = restore_snapshot(c, c->snaps[1], s);
        if (rc != BHX_OK) break;
        c->cfg = cfg;
        c->walk_sub = walk_lists_for(&cfg);
        drop_graphs(c);
        rc = bhx_step_graph(c, 1);                     // capture + first frame
        if (rc != BHX_OK) break;
        const auto t0 = std::chrono::steady_clock::now();
        rc = bhx_step_graph(c, steps);
        const auto t1 = std::chrono::steady_clock::now();
        if (rc != BHX_OK) break;
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / steps;
        // the last frame's walk is at the current positions: compare with the FP64 direct sum
        hipError_t e = launch_direct(c, dt, ns, c->cfg.eps, dref, nullptr, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_pick3, dim3(grid_for(ns)), dim3(kBlock), 0, s, c->ax, c->ay, c->az, dt, ns, dwalk);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(ref.data(), dref, 3 * ns * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(walk.data(), dwalk, 3 * ns * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { rc = fail(c, BHX_E_HIP, "autotune error sample", e); break; }
        for (size_t i = 0; i < ns; ++i) {
            const double dx = walk[3 * i] - ref[3 * i], dy = walk[3 * i + 1] - ref[3 * i + 1],
                         dz = walk[3 * i + 2] - ref[3 * i + 2];
            const double r = std::sqrt(ref[3 * i] * ref[3 * i] + ref[3 * i + 1] * ref[3 * i + 1] +
                                       ref[3 * i + 2] * ref[3 * i + 2]);
            err[i] = r > 0.0 ? std::sqrt(dx * dx + dy * dy + dz * dz) / r : 0.0;
        }
        std::nth_element(err.begin(), err.begin() + ns / 2, err.end());
        const double med = err[ns / 2];
        if (out) {
            out[kk].leaf_capacity = caps[k];
            out[kk].group_size = group_sizes[g];
            out[kk].ms_per_step = ms;
            out[kk].median_err = med;
        }
        if (med <= max_err && (!found || ms < best_ms)) {
            found = true;
            best = caps[k];
            best_group = group_sizes[g];
            best_ms = ms;
        }
    }
    (void)hipFree(dt);
    (void)hipFree(dref);
    (void)hipFree(dwalk);
    // back to the starting state, with the chosen capacity and group size
    c->cfg = base;
    c->cfg.leaf_capacity = best;
    c->cfg.group_size = best_group;
    c->walk_sub = walk_lists_for(&c->cfg);
    drop_graphs(c);
    const int rc2 = restore_snapshot(c, c->snaps[1], s);
    HIPCHK(hipStreamSynchronize(s));
    if (rc != BHX_OK) return rc;
    if (rc2 != BHX_OK) return rc2;
    if (chosen) *chosen = best;
    if (chosen_group) *chosen_group = best_group;
    return BHX_OK;
}

int bhx_direct(bhx_ctx* c, const int64_t* targets, size_t nt, double eps, double* acc_out, double* pot_out) {
    if (!c || !targets || !acc_out) return c ? fail(c, BHX_E_NULL, "null argument") : BHX_E_NULL;
    for (size_t k = 0; k < nt; ++k)
        if (targets[k] < 0 || (size_t)targets[k] >= c->n) return fail(c, BHX_E_ARG, "target out of range");
    if (nt == 0) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    int64_t* dt = nullptr;
    double *da = nullptr, *dp = nullptr;
    hipError_t e = hipMalloc((void**)&dt, nt * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&da, nt * 24);
    if (e == hipSuccess) e = hipMalloc((void**)&dp, nt * 8);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, targets, nt * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_direct(c, dt, nt, eps, da, dp, s);
    if (e == hipSuccess) e = hipMemcpyAsync(acc_out, da, nt * 24, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && pot_out) e = hipMemcpyAsync(pot_out, dp, nt * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (dt) (void)hipFree(dt);
    if (da) (void)hipFree(da);
    if (dp) (void)hipFree(dp);
    if (e != hipSuccess) return fail(c, BHX_E_HIP, "direct sum", e);
    return BHX_OK;
}

int bhx_energy(bhx_ctx* c, double* ekin, double* epot, double* virial, double* pot_out) {
    if (!c) return BHX_E_NULL;
    double K = 0.0;
    int rc = bhx_diagnostics(c, &K, nullptr, nullptr);
    if (rc != BHX_OK) return rc;
    double W = 0.0;
    const size_t n = c->n;
    if (n) {
        hipStream_t s = c->s_main;
        // the tree of the current positions in their own bounds
        HIPCHK(launch_bounds(c, c->bounds_slot(c->bslot), s));
        c->bounds_valid = true;
        rc = build_on_device(c, s);
        if (rc != BHX_OK) return rc;
        rc = read_counters(c, s);
        if (rc != BHX_OK) return rc;
        double* tmp = nullptr;
        if (dalloc(&tmp, 4 * n) != hipSuccess) return fail(c, BHX_E_NOMEM, "potential scratch");
        std::vector<double> phi(n), m(n);
        hipError_t e = launch_walk_potential(c, s, tmp, tmp + n, tmp + 2 * n, tmp + 3 * n);
        if (e == hipSuccess) e = hipMemcpyAsync(phi.data(), tmp + 3 * n, n * 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(m.data(), c->m, n * 8, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(tmp);
        HIPCHK(e);
        for (size_t i = 0; i < n; ++i) W += 0.5 * m[i] * phi[i];
        if (pot_out) std::copy(phi.begin(), phi.end(), pot_out);
    }
    if (ekin) *ekin = K;
    if (epot) *epot = W;
    if (virial) *virial = W != 0.0 ? 2.0 * K / std::fabs(W) : 0.0;
    return BHX_OK;
}

int bhx_diagnostics(bhx_ctx* c, double* ekin, double mom[3], double angmom[3]) {
    if (!c) return BHX_E_NULL;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    double* out = c->stage + 7 * 256;
    HIPCHK(launch_diag(c, out, s));
    double h[7];
    HIPCHK(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (ekin) *ekin = h[0];
    if (mom) { mom[0] = h[1]; mom[1] = h[2]; mom[2] = h[3]; }
    if (angmom) { angmom[0] = h[4]; angmom[1] = h[5]; angmom[2] = h[6]; }
    return BHX_OK;
}


// ---- the reference's per-body tree API -----------------------------------

int bhx_interaction_lists(bhx_ctx* c, const double* points, size_t nq, double theta, uint64_t* offsets,
                          int64_t* entries, size_t cap, size_t* total) {
    if (!c || !offsets || (nq && !points)) return c ? fail(c, BHX_E_NULL, "null argument") : BHX_E_NULL;
    int rc = need_tree_table(c);
    if (rc != BHX_OK) return rc;
    if (!(theta >= 0.0)) return fail(c, BHX_E_ARG, "theta must be >= 0");
    offsets[0] = 0;
    if (total) *total = 0;
    if (nq == 0 || c->n_nodes == 0) {
        for (size_t q = 0; q <= nq; ++q) offsets[q] = 0;
        return BHX_OK;
    }
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    DevTmp tmp;
    double* dp = nullptr;
    uint64_t* dcnt = nullptr;
    uint32_t* derr = nullptr;
    HIPCHK(tmp.get(&dp, 3 * nq));
    HIPCHK(tmp.get(&dcnt, nq + 1));
    HIPCHK(tmp.get(&derr, 1));
    HIPCHK(hipMemsetAsync(derr, 0, sizeof(uint32_t), s));
    HIPCHK(hipMemcpyAsync(dp, points, 3 * nq * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(launch_ref_lists(c, dp, nq, theta, nullptr, dcnt, nullptr, derr, s));
    std::vector<uint64_t> cnt(nq);
    uint32_t herr = 0;
    HIPCHK(hipMemcpyAsync(cnt.data(), dcnt, nq * sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&herr, derr, sizeof herr, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return fail(c, BHX_E_DEPTH, "traversal stack bound exceeded");
    for (size_t q = 0; q < nq; ++q) offsets[q + 1] = offsets[q] + cnt[q];
    const size_t tot = offsets[nq];
    if (total) *total = tot;
    if (!entries || tot == 0) return BHX_OK;
    if (cap < tot) return fail(c, BHX_E_CAPACITY, "entries hold fewer values than the lists");
    uint64_t* doff = nullptr;
    int64_t* dent = nullptr;
    HIPCHK(tmp.get(&doff, nq + 1));
    HIPCHK(tmp.get(&dent, tot));
    HIPCHK(hipMemcpyAsync(doff, offsets, (nq + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    HIPCHK(launch_ref_lists(c, dp, nq, theta, doff, dcnt, dent, derr, s));
    HIPCHK(hipMemcpyAsync(entries, dent, tot * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_force_lists(bhx_ctx* c, const double* points, size_t nq, const uint64_t* offsets, const int64_t* entries,
                    double* acc) {
    if (!c || (nq && (!points || !offsets || !acc))) return c ? fail(c, BHX_E_NULL, "null argument") : BHX_E_NULL;
    int rc = need_tree_table(c);
    if (rc != BHX_OK) return rc;
    if (nq == 0) return BHX_OK;
    for (size_t q = 0; q < nq; ++q)
        if (offsets[q + 1] < offsets[q]) return fail(c, BHX_E_ARG, "offsets must not decrease");
    const size_t tot = offsets[nq] - offsets[0];
    if (tot && !entries) return fail(c, BHX_E_NULL, "entries is a null pointer");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    DevTmp tmp;
    double *dp = nullptr, *dacc = nullptr;
    uint64_t* doff = nullptr;
    int64_t* dent = nullptr;
    uint32_t* derr = nullptr;
    HIPCHK(tmp.get(&dp, 3 * nq));
    HIPCHK(tmp.get(&dacc, 3 * nq));
    HIPCHK(tmp.get(&doff, nq + 1));
    HIPCHK(tmp.get(&dent, tot));
    HIPCHK(tmp.get(&derr, 1));
    std::vector<uint64_t> off(offsets, offsets + nq + 1);
    for (auto& o : off) o -= offsets[0];
    HIPCHK(hipMemsetAsync(derr, 0, sizeof(uint32_t), s));
    HIPCHK(hipMemcpyAsync(dp, points, 3 * nq * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dacc, acc, 3 * nq * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(doff, off.data(), (nq + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    if (tot) HIPCHK(hipMemcpyAsync(dent, entries + offsets[0], tot * sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCHK(launch_force_lists(c, dp, nq, doff, dent, dacc, derr, s));
    uint32_t herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, derr, sizeof herr, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return fail(c, BHX_E_ARG, "a list entry names no node or body of the tree");
    HIPCHK(hipMemcpyAsync(acc, dacc, 3 * nq * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

int bhx_mac(bhx_ctx* c, const int64_t* node_ids, const double* points, size_t n, double theta, uint8_t* accept) {
    if (!c || (n && (!node_ids || !points || !accept))) return c ? fail(c, BHX_E_NULL, "null argument") : BHX_E_NULL;
    int rc = need_tree_table(c);
    if (rc != BHX_OK) return rc;
    if (n == 0) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    DevTmp tmp;
    int64_t* did = nullptr;
    double* dp = nullptr;
    uint8_t* dout = nullptr;
    uint32_t* derr = nullptr;
    HIPCHK(tmp.get(&did, n));
    HIPCHK(tmp.get(&dp, 3 * n));
    HIPCHK(tmp.get(&dout, n));
    HIPCHK(tmp.get(&derr, 1));
    HIPCHK(hipMemsetAsync(derr, 0, sizeof(uint32_t), s));
    HIPCHK(hipMemcpyAsync(did, node_ids, n * sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dp, points, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(launch_mac(c, did, dp, n, theta, dout, derr, s));
    uint32_t herr = 0;
    HIPCHK(hipMemcpyAsync(accept, dout, n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&herr, derr, sizeof herr, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return fail(c, BHX_E_ARG, "a node index names no node of the tree");
    return BHX_OK;
}

int bhx_node_put(bhx_ctx* c, const bhx_node* rec, int64_t* id_out) {
    if (!c || !rec) return c ? fail(c, BHX_E_NULL, "node is a null pointer") : BHX_E_NULL;
    int rc = need_tree_table(c);
    if (rc != BHX_OK) return rc;
    const int lvl = key_level(rec->key);
    if (lvl < 0 || lvl != rec->level) return fail(c, BHX_E_ARG, "key and level do not match");
    if (rec->n < 0 || rec->n > (int64_t)UINT32_MAX || rec->child_byte > 0xFFu)
        return fail(c, BHX_E_ARG, "count or child bits out of range");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    DevTmp tmp;
    uint64_t* dk = nullptr;
    int64_t* di = nullptr;
    uint32_t* derr = nullptr;
    HIPCHK(tmp.get(&dk, 1));
    HIPCHK(tmp.get(&di, 1));
    HIPCHK(tmp.get(&derr, 1));
    HIPCHK(hipMemsetAsync(derr, 0, sizeof(uint32_t), s));
    HIPCHK(hipMemcpyAsync(dk, &rec->key, sizeof(uint64_t), hipMemcpyHostToDevice, s));
    int64_t id = -1;
    if (c->n_nodes) {
        HIPCHK(launch_lookup(c, 0, dk, 1, di, s));
        HIPCHK(hipMemcpyAsync(&id, di, sizeof id, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    const bool insert = id < 0;
    if (insert) {
        if (c->n_nodes >= c->tree.cap) return fail(c, BHX_E_CAPACITY, "the node table is full");
        id = (int64_t)c->n_nodes;
    }
    HIPCHK(launch_node_put(c, *rec, (uint32_t)id, insert, derr, s));
    uint32_t herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, derr, sizeof herr, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (herr) return fail(c, BHX_E_HASH_FULL, "node hash table full");
    if (insert) ++c->n_nodes;
    c->tree_valid = false;
    c->tree_edited = true;
    if (id_out) *id_out = id;
    return BHX_OK;
}

int bhx_node_erase(bhx_ctx* c, uint64_t key, int* erased) {
    if (!c) return BHX_E_NULL;
    int rc = need_tree_table(c);
    if (rc != BHX_OK) return rc;
    if (erased) *erased = 0;
    if (c->n_nodes == 0 || key_level(key) < 0) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    DevTmp tmp;
    uint32_t* ddone = nullptr;
    HIPCHK(tmp.get(&ddone, 1));
    HIPCHK(launch_node_erase(c, key, ddone, s));
    uint32_t done = 0;
    HIPCHK(hipMemcpyAsync(&done, ddone, sizeof done, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (done) {
        c->tree_valid = false;
        c->tree_edited = true;
    }
    if (erased) *erased = done ? 1 : 0;
    return BHX_OK;
}

int bhx_prune_empty(bhx_ctx* c, uint64_t key, size_t* removed) {
    if (!c) return BHX_E_NULL;
    int rc = need_tree_table(c);
    if (rc != BHX_OK) return rc;
    if (removed) *removed = 0;
    const int lvl = key_level(key);
    if (lvl < 0) return fail(c, BHX_E_ARG, "not a node key");
    if (c->n_nodes == 0) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    HIPCHK(hipStreamSynchronize(c->s_side));
    DevTmp tmp;
    uint32_t* dn = nullptr;
    HIPCHK(tmp.get(&dn, 1));
    HIPCHK(hipMemsetAsync(dn, 0, sizeof(uint32_t), s));
    HIPCHK(launch_prune(c, key, (uint32_t)lvl, dn, s));
    uint32_t n = 0;
    HIPCHK(hipMemcpyAsync(&n, dn, sizeof n, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (n) {
        c->tree_valid = false;
        c->tree_edited = true;
    }
    if (removed) *removed = n;
    return BHX_OK;
}

int bhx_decode_morton_keys(bhx_ctx* c, const uint64_t* keys, size_t n, double size, double* pos) {
    if (!c || (n && (!keys || !pos))) return c ? fail(c, BHX_E_NULL, "null argument") : BHX_E_NULL;
    if (n == 0) return BHX_OK;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->s_main;
    DevTmp tmp;
    uint64_t* dk = nullptr;
    double* dp = nullptr;
    HIPCHK(tmp.get(&dk, n));
    HIPCHK(tmp.get(&dp, 3 * n));
    HIPCHK(hipMemcpyAsync(dk, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice, s));
    HIPCHK(launch_decode(dk, n, size, dp, s));
    HIPCHK(hipMemcpyAsync(pos, dp, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BHX_OK;
}

}  // extern "C"
