/*
 * bhx_oracle.c — serial CPU restatement of the reference Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Built with: gcc -O2 -std=c11 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 * No FMA contraction, no -ffast-math, no AVX-512: the double->uint64 key
 * conversion is spelled out with x86-64 SSE2 semantics (orc_x86_cvt_u64).
 *
 * Citations are file:line into the reference's src/ directory.
 *
 * Pinning (what checks each function independently of its own formula):
 *   orc_bounds            closed form on power-of-two N; hand-summed cases
 *   orc_morton_key        survey KATs incl. the x86 negative case, and a bit-by-bit
 *                         interleave of the effective 21-bit value for random t
 *   orc_radix_sort        numpy stable argsort of the same keys
 *   orc_tree_build_insert orc_tree_build_recursive (independent definition of the
 *                         node set) and invariants (N_parent = sum N_child ...)
 *   orc_tree_moments_*    brute-force mass / COM / quadrupole from member lists
 *   orc_walk              theta = 0 reduces to orc_direct; two-body closed form
 *   orc_direct            two-body closed form; Newton's third law (sum m a = 0)
 *   orc_kick_drift/kick   closed form on constant acceleration
 *   orc_traj_direct_adjoint  central finite differences of orc_traj_direct (no tree,
 *                         no decisions: differentiable everywhere); two-body closed form
 */
#define _POSIX_C_SOURCE 199309L
#include "bhx_oracle.h"

/* Planted mistakes, for tests only.  tests/test_oracle.py compiles this file
 * a second time with one of these defined and checks that the known-answer
 * tests for the reference's rounding order FAIL on the result, so each KAT
 * provably distinguishes the reference's form from the plausible alternative:
 *   ORC_MUTANT_MEAN_DIV   centre = sum / N        (reference: sum * fl(1/N))
 *   ORC_MUTANT_HALF_DT    kick  += a * (dt * 0.5) (reference: (a * dt) * 0.5)
 *   ORC_MUTANT_FMA        kick and drift contracted into fused multiply-adds
 *   ORC_MUTANT_MAC_MUL    s^2 < theta^2 d^2       (reference: s^2 / d^2 < theta^2)
 * None is defined in the oracle the tests and the bench load. */

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------- */
/* bounds                                                                     */
/* ------------------------------------------------------------------------- */

/* OctantBounds(Body*, size_t) — OctantUtils.cpp:27-59 with Vec3D::operator/=
 * (MathUtils.cpp:81-90): the centre is a left-to-right FP64 sum in array
 * order, multiplied by fl(1/N); the size is max(0, max_{i,a} |p_ia| - c_a). */
void orc_bounds(const double *pos, size_t n, double center[3], double *size)
{
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (size_t i = 0; i < n; ++i) {
        cx = cx + pos[3 * i + 0];
        cy = cy + pos[3 * i + 1];
        cz = cz + pos[3 * i + 2];
    }
    double dn = (double)n;
    if (dn != 0.0) {
#ifdef ORC_MUTANT_MEAN_DIV
        cx = cx / dn;
        cy = cy / dn;
        cz = cz / dn;
#else
        double inv = 1.0 / dn;
        cx = cx * inv;
        cy = cy * inv;
        cz = cz * inv;
#endif
    }
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double t;
        t = fabs(pos[3 * i + 0]) - cx; if (t > s) s = t;
        t = fabs(pos[3 * i + 1]) - cy; if (t > s) s = t;
        t = fabs(pos[3 * i + 2]) - cz; if (t > s) s = t;
    }
    center[0] = cx; center[1] = cy; center[2] = cz;
    *size = s;
}

/* ------------------------------------------------------------------------- */
/* Morton keys                                                                */
/* ------------------------------------------------------------------------- */

/* The reference writes (spatialKey)(xScaled) (MortonKeys.hpp:162-164).  For
 * negative or huge values that is undefined in C++; the x86-64 SSE2 lowering
 * is cvttsd2si (truncate to int64, 0x8000000000000000 when out of range or
 * NaN) followed by a reinterpretation as uint64.  That is the semantics fixed
 * here and on the GPU. */
uint64_t orc_x86_cvt_u64(double u)
{
    if (!(u >= -9223372036854775808.0 && u < 9223372036854775808.0))
        return 0x8000000000000000ull;
    return (uint64_t)(int64_t)u;
}

/* interleaveBits' per-axis swizzle (MortonKeys.hpp:100-116).  The five
 * shift/mask steps and their masks (MortonKeys.hpp:51-58) are the published
 * "magic bits" constants for 21-bit spreading. */
uint64_t orc_spread21(uint64_t t)
{
    t = (t | (t << 32)) & 0x001f00000000ffffull;
    t = (t | (t << 16)) & 0x001f0000ff0000ffull;
    t = (t | (t << 8))  & 0x100f00f00f00f00full;
    t = (t | (t << 4))  & 0x10c30c30c30c30c3ull;
    t = (t | (t << 2))  & 0x1249249249249249ull;
    return t;
}

/* computeMortonKey — MortonKeys.hpp:137-181 (+ interleaveBits :95-124):
 * origin-centred domain [-size, size], frac = 2^21 / (2 size), key =
 * 1<<63 | X<<2 | Y<<1 | Z. */
uint64_t orc_morton_key(const double p[3], double size)
{
    double xs = p[0] + size;
    double ys = p[1] + size;
    double zs = p[2] + size;
    double frac = 2097152.0 / (2.0 * size);
    xs = xs * frac;
    ys = ys * frac;
    zs = zs * frac;
    uint64_t X = orc_spread21(orc_x86_cvt_u64(xs));
    uint64_t Y = orc_spread21(orc_x86_cvt_u64(ys));
    uint64_t Z = orc_spread21(orc_x86_cvt_u64(zs));
    return (1ull << 63) | (X << 2) | (Y << 1) | Z;
}

/* decodeMortonKey — MortonKeys.hpp:204-228: de-interleave the three 21-bit
 * coordinates, then c / 2^21 * 2 * size - size, left to right. */
void orc_decode_morton_key(uint64_t key, double size, double out[3])
{
    uint64_t c[3] = {0, 0, 0};
    for (int b = 0; b < 21; ++b) {
        c[0] |= ((key >> (3 * b + 2)) & 1ull) << b;
        c[1] |= ((key >> (3 * b + 1)) & 1ull) << b;
        c[2] |= ((key >> (3 * b)) & 1ull) << b;
    }
    for (int a = 0; a < 3; ++a) out[a] = (double)c[a] / 2097152.0 * 2.0 * size - size;
}

/* ofApp::update() key loop — ofApp.cpp:172-175. */
void orc_morton_keys(const double *pos, size_t n, double size, uint64_t *keys)
{
    for (size_t i = 0; i < n; ++i)
        keys[i] = orc_morton_key(pos + 3 * i, size);
}

/* ThreePassRadixSortMortonKeys + RadixSortPass — MortonKeys.hpp:373-470:
 * nine stable LSD byte passes at bit offsets 0,8,16 / 21,29,37 / 42,50,58,
 * carrying the index array along. */
void orc_radix_sort(uint64_t *keys, uint64_t *idx, size_t n)
{
    static const int shifts[9] = {0, 8, 16, 21, 29, 37, 42, 50, 58};
    uint64_t *tk = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint64_t *ti = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    size_t hist[256];
    for (int p = 0; p < 9; ++p) {
        int sh = shifts[p];
        memset(hist, 0, sizeof hist);
        for (size_t i = 0; i < n; ++i) hist[(keys[i] >> sh) & 0xFF]++;
        size_t run = 0;
        for (int b = 0; b < 256; ++b) { size_t c = hist[b]; hist[b] = run; run += c; }
        for (size_t i = 0; i < n; ++i) {
            size_t b = (keys[i] >> sh) & 0xFF;
            tk[hist[b]] = keys[i];
            ti[hist[b]] = idx[i];
            hist[b]++;
        }
        memcpy(keys, tk, n * sizeof(uint64_t));
        memcpy(idx, ti, n * sizeof(uint64_t));
    }
    free(tk);
    free(ti);
}

/* ReorderBodiesByIndexes — Body.hpp:257-273: new[i] = old[idx[i]]. */
void orc_gather(double *arr, size_t width, const uint64_t *idx, size_t n)
{
    double *tmp = (double *)malloc((n ? n : 1) * width * sizeof(double));
    for (size_t i = 0; i < n; ++i)
        for (size_t w = 0; w < width; ++w) tmp[i * width + w] = arr[idx[i] * width + w];
    memcpy(arr, tmp, n * width * sizeof(double));
    free(tmp);
}

/* The half kick v + (a*dt)*0.5 of Body.hpp:310,322, two roundings then an
 * addition (no fused multiply-add). */
static inline double orc_half_kick(double v, double a, double dt)
{
#if defined(ORC_MUTANT_HALF_DT)
    return v + a * (dt * 0.5);
#elif defined(ORC_MUTANT_FMA)
    return fma(a, dt * 0.5, v);
#else
    return v + (a * dt) * 0.5;
#endif
}

/* The drift x + v*dt of Body.hpp:311, rounded product then addition. */
static inline double orc_drift(double x, double v, double dt)
{
#ifdef ORC_MUTANT_FMA
    return fma(v, dt, x);
#else
    return x + v * dt;
#endif
}

/* ComputePositionAtHalfTimeStep — Body.hpp:306-313: v += (a*dt)*0.5; x += v*dt. */
void orc_kick_drift(double dt, double *pos, double *vel, const double *acc, size_t n)
{
    for (size_t i = 0; i < 3 * n; ++i) {
        vel[i] = orc_half_kick(vel[i], acc[i], dt);
        pos[i] = orc_drift(pos[i], vel[i], dt);
    }
}

/* ComputeVelocityAndPosition — Body.hpp:317-324: v += (a*dt)*0.5. */
void orc_kick(double dt, double *vel, const double *acc, size_t n)
{
    for (size_t i = 0; i < 3 * n; ++i) vel[i] = orc_half_kick(vel[i], acc[i], dt);
}

/* ------------------------------------------------------------------------- */
/* octree container: node array + open-addressed key -> index map            */
/* ------------------------------------------------------------------------- */

struct orc_tree {
    orc_node *nodes;
    size_t    count, cap;
    uint64_t *hkeys;      /* 0 = empty slot (a valid key always has its sentinel bit) */
    int64_t  *hvals;
    size_t    hcap;       /* power of two */
    int64_t  *members;    /* leaf member lists, array order within each leaf */
    size_t    nmembers;
    size_t    unsplit;    /* leaves at the depth cap holding more than one body */
};

static uint64_t orc_hash64(uint64_t k)
{
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

void orc_tree_path_keys(const double *pos, size_t n, const double c[3], double s, uint64_t *keys)
{
    for (size_t i = 0; i < n; ++i) keys[i] = orc_tree_path_key(pos + 3 * i, c, s);
}

static void orc_map_put(orc_tree *t, uint64_t key, int64_t val);

static void orc_map_grow(orc_tree *t)
{
    size_t oc = t->hcap;
    uint64_t *ok = t->hkeys;
    int64_t *ov = t->hvals;
    t->hcap = oc ? oc * 2 : 1024;
    t->hkeys = (uint64_t *)calloc(t->hcap, sizeof(uint64_t));
    t->hvals = (int64_t *)malloc(t->hcap * sizeof(int64_t));
    for (size_t i = 0; i < oc; ++i)
        if (ok[i]) orc_map_put(t, ok[i], ov[i]);
    free(ok);
    free(ov);
}

static void orc_map_put(orc_tree *t, uint64_t key, int64_t val)
{
    size_t m = t->hcap - 1, h = orc_hash64(key) & m;
    while (t->hkeys[h] && t->hkeys[h] != key) h = (h + 1) & m;
    t->hkeys[h] = key;
    t->hvals[h] = val;
}

/* LinearHashedOctree::lookUpNode — HashedOctree.cpp:111-115 (-1 for a missing key). */
int64_t orc_tree_lookup(const orc_tree *t, uint64_t key)
{
    if (!t->hcap || !key) return -1;
    size_t m = t->hcap - 1, h = orc_hash64(key) & m;
    while (t->hkeys[h]) {
        if (t->hkeys[h] == key) return t->hvals[h];
        h = (h + 1) & m;
    }
    return -1;
}

static orc_tree *orc_tree_new(void)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    orc_map_grow(t);
    return t;
}

static int64_t orc_tree_add(orc_tree *t, const orc_node *nd)
{
    if (t->count == t->cap) {
        t->cap = t->cap ? t->cap * 2 : 1024;
        t->nodes = (orc_node *)realloc(t->nodes, t->cap * sizeof(orc_node));
    }
    if (2 * (t->count + 1) > t->hcap) orc_map_grow(t);
    int64_t id = (int64_t)t->count++;
    t->nodes[id] = *nd;
    orc_map_put(t, nd->key, id);
    return id;
}

void orc_tree_free(orc_tree *t)
{
    if (!t) return;
    free(t->nodes);
    free(t->hkeys);
    free(t->hvals);
    free(t->members);
    free(t);
}

size_t orc_tree_size(const orc_tree *t) { return t->count; }
size_t orc_tree_unsplit(const orc_tree *t) { return t->unsplit; }

void orc_tree_nodes(const orc_tree *t, orc_node *out)
{
    memcpy(out, t->nodes, t->count * sizeof(orc_node));
}

/* DetermineOctant — OctantUtils.hpp:64-70 (strict >, x is the high bit). */
static inline unsigned orc_octant(const double c[3], const double p[3])
{
    return ((unsigned)(p[0] > c[0]) << 2) | ((unsigned)(p[1] > c[1]) << 1) | (unsigned)(p[2] > c[2]);
}

/* Child geometry — HOTNode(parent, octant, ...) HashedNode.cpp:48-61 with
 * OctantDir (OctantUtils.hpp:36-46): size' = size*0.5,
 * centre'_a = centre_a + (dir_a*0.5)*size'; key' = key<<3 | o
 * (GetChildKey, OctantUtils.hpp:97-101). */
static inline void orc_child_geometry(const double c[3], double s, unsigned o,
                                      double cc[3], double *cs)
{
    double h = s * 0.5;
    double dx = (o & 4) ? 1.0 : -1.0;
    double dy = (o & 2) ? 1.0 : -1.0;
    double dz = (o & 1) ? 1.0 : -1.0;
    cc[0] = c[0] + (dx * 0.5) * h;
    cc[1] = c[1] + (dy * 0.5) * h;
    cc[2] = c[2] + (dz * 0.5) * h;
    *cs = h;
}

/* The 21-level descent a body takes during insertion, written as one key:
 * sentinel 1, then the octant of every level (OctantUtils.hpp:64-70,97-101,
 * HashedNode.cpp:50-56).  Level-L node of the body = key >> (63 - 3L). */
uint64_t orc_tree_path_key(const double p[3], const double c[3], double s)
{
    uint64_t k = 1;
    double cc[3] = {c[0], c[1], c[2]};
    double cs = s;
    for (int L = 1; L <= ORC_MAX_LEVEL; ++L) {
        unsigned o = orc_octant(cc, p);
        k = (k << 3) | o;
        double nc[3], ns;
        orc_child_geometry(cc, cs, o, nc, &ns);
        cc[0] = nc[0]; cc[1] = nc[1]; cc[2] = nc[2];
        cs = ns;
    }
    return k;
}

static void orc_node_init(orc_node *nd, uint64_t key, const double c[3], double s, int level)
{
    memset(nd, 0, sizeof *nd);
    nd->key = key;
    nd->center[0] = c[0]; nd->center[1] = c[1]; nd->center[2] = c[2];
    nd->size = s;
    nd->level = level;
    nd->first = -1;
    nd->member_begin = -1;
}

/* createChildNode — HashedOctree.cpp:697-711: new child holding one body. */
static int64_t orc_create_child(orc_tree *t, int64_t parent, unsigned o, const double com[3],
                                double m, int64_t first)
{
    orc_node nd;
    const orc_node *pn = &t->nodes[parent];
    double cc[3], cs;
    orc_child_geometry(pn->center, pn->size, o, cc, &cs);
    orc_node_init(&nd, (pn->key << 3) | o, cc, cs, pn->level + 1);
    nd.com[0] = com[0]; nd.com[1] = com[1]; nd.com[2] = com[2];
    nd.mass = m;
    nd.n = 1;
    nd.first = first;
    t->nodes[parent].child_byte |= 1u << o;
    return orc_tree_add(t, &nd);
}

/* Leaf member lists: every body descends from the root to the leaf holding
 * it; lists keep array order. */
static void orc_fill_members(orc_tree *t, const double *pos, size_t n)
{
    free(t->members);
    t->members = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
    t->nmembers = n;
    int64_t *leaf_of = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
    for (size_t k = 0; k < t->count; ++k) t->nodes[k].member_count = 0;
    for (size_t i = 0; i < n; ++i) {
        int64_t id = t->count ? 0 : -1;
        while (id >= 0 && t->nodes[id].child_byte) {
            unsigned o = orc_octant(t->nodes[id].center, pos + 3 * i);
            id = orc_tree_lookup(t, (t->nodes[id].key << 3) | o);
        }
        leaf_of[i] = id;
        if (id >= 0) t->nodes[id].member_count++;
    }
    int64_t run = 0;
    for (size_t k = 0; k < t->count; ++k) {
        orc_node *nd = &t->nodes[k];
        if (nd->child_byte == 0) {
            nd->member_begin = run;
            run += nd->member_count;
            nd->member_count = 0;
        }
    }
    for (size_t i = 0; i < n; ++i) {
        if (leaf_of[i] < 0) continue;
        orc_node *nd = &t->nodes[leaf_of[i]];
        t->members[nd->member_begin + nd->member_count++] = (int64_t)i;
    }
    free(leaf_of);
}

/* buildLinearHashedOctreeInPlace + insertBody / pushHOTNodeBodyToChild /
 * processNodeInsertion / createChildNode — HashedOctree.hpp:211-244,
 * HashedOctree.cpp:450-711.  Leaf capacity 1, sequential insertion in array
 * order.  Payload as executed (the moments pass never runs, HashedOctree.hpp:
 * 220-234 with HashedOctree.cpp:829-834): N = count, mass = running sum in
 * insertion order, baryCenter = position of the first body inserted into the
 * cell, quadrupole = 0.  Depth is capped at 21 levels (a deeper child key
 * would overflow the 64-bit key); a capped cell keeps all its bodies as one
 * leaf and is counted in `unsplit`.  An empty input yields an empty tree
 * (the root is pruned, HashedOctree.cpp:633-638). */
orc_tree *orc_tree_build_insert(const double *pos, const double *mass, size_t n,
                                const double c[3], double s)
{
    orc_tree *t = orc_tree_new();
    if (n == 0) return t;
    orc_node root;
    orc_node_init(&root, 1, c, s, 0);
    orc_tree_add(t, &root);
    for (size_t i = 0; i < n; ++i) {
        const double *p = pos + 3 * i;
        int64_t id = 0;
        for (;;) {
            orc_node *nd = &t->nodes[id];
            if (nd->n == 0) {                                   /* case 3 */
                nd->mass = mass[i];
                nd->com[0] = p[0]; nd->com[1] = p[1]; nd->com[2] = p[2];
                nd->n = 1;
                nd->first = (int64_t)i;
                break;
            }
            if (nd->level == ORC_MAX_LEVEL) {                   /* depth cap */
                nd->n = nd->n + 1;
                nd->mass = nd->mass + mass[i];
                break;
            }
            if (nd->n == 1) {                                   /* case 5: push resident */
                unsigned ro = orc_octant(nd->center, nd->com);
                double rc[3] = {nd->com[0], nd->com[1], nd->com[2]};
                orc_create_child(t, id, ro, rc, nd->mass, nd->first);
                nd = &t->nodes[id];
            }
            unsigned o = orc_octant(nd->center, p);            /* cases 4/5 */
            nd->n = nd->n + 1;
            nd->mass = nd->mass + mass[i];
            if (nd->child_byte & (1u << o)) {
                id = orc_tree_lookup(t, (nd->key << 3) | o);
                continue;
            }
            orc_create_child(t, id, o, p, mass[i], (int64_t)i);
            break;
        }
    }
    orc_fill_members(t, pos, n);
    for (size_t k = 0; k < t->count; ++k)
        if (t->nodes[k].child_byte == 0 && t->nodes[k].n > 1) t->unsplit++;
    return t;
}

/* Independent definition of the same node set by recursive subdivision:
 * a cell is split iff it holds more than `leaf_cap` bodies and is above the
 * depth cap; a child exists iff it receives at least one body.  With
 * leaf_cap = 1 this is the reference's insertion result; larger capacities
 * are the bucket generalisation.  Payload as in orc_tree_build_insert. */
static void orc_rec(orc_tree *t, const double *pos, const double *mass, int64_t *idx,
                    int64_t *scratch, size_t cnt, int64_t self, int leaf_cap)
{
    orc_node *nd = &t->nodes[self];
    nd->n = (int64_t)cnt;
    double msum = 0.0;
    for (size_t k = 0; k < cnt; ++k) msum = k ? msum + mass[idx[k]] : mass[idx[k]];
    nd->mass = cnt ? msum : 0.0;
    if (cnt) {
        nd->first = idx[0];
        nd->com[0] = pos[3 * idx[0]];
        nd->com[1] = pos[3 * idx[0] + 1];
        nd->com[2] = pos[3 * idx[0] + 2];
    }
    if ((int64_t)cnt <= leaf_cap || nd->level == ORC_MAX_LEVEL) return;
    size_t bucket[9] = {0};
    for (size_t k = 0; k < cnt; ++k) bucket[orc_octant(nd->center, pos + 3 * idx[k]) + 1]++;
    for (int o = 0; o < 8; ++o) bucket[o + 1] += bucket[o];
    size_t fill[8];
    for (int o = 0; o < 8; ++o) fill[o] = bucket[o];
    for (size_t k = 0; k < cnt; ++k) scratch[fill[orc_octant(nd->center, pos + 3 * idx[k])]++] = idx[k];
    memcpy(idx, scratch, cnt * sizeof(int64_t));
    for (unsigned o = 0; o < 8; ++o) {
        size_t b = bucket[o], e = bucket[o + 1];
        if (e == b) continue;
        orc_node ch;
        double cc[3], cs;
        orc_child_geometry(t->nodes[self].center, t->nodes[self].size, o, cc, &cs);
        orc_node_init(&ch, (t->nodes[self].key << 3) | o, cc, cs, t->nodes[self].level + 1);
        t->nodes[self].child_byte |= 1u << o;
        int64_t cid = orc_tree_add(t, &ch);
        orc_rec(t, pos, mass, idx + b, scratch + b, e - b, cid, leaf_cap);
    }
}

orc_tree *orc_tree_build_recursive(const double *pos, const double *mass, size_t n,
                                   const double c[3], double s, int leaf_cap)
{
    orc_tree *t = orc_tree_new();
    if (n == 0) return t;
    if (leaf_cap < 1) leaf_cap = 1;
    orc_node root;
    orc_node_init(&root, 1, c, s, 0);
    orc_tree_add(t, &root);
    int64_t *idx = (int64_t *)malloc(n * sizeof(int64_t));
    int64_t *scr = (int64_t *)malloc(n * sizeof(int64_t));
    for (size_t i = 0; i < n; ++i) idx[i] = (int64_t)i;
    orc_rec(t, pos, mass, idx, scr, n, 0, leaf_cap);
    free(idx);
    free(scr);
    orc_fill_members(t, pos, n);
    for (size_t k = 0; k < t->count; ++k)
        if (t->nodes[k].child_byte == 0 && t->nodes[k].n > leaf_cap) t->unsplit++;
    return t;
}

/* Intended moments (what computeNodeBaryCenters, HashedOctree.cpp:721-827,
 * sets out to compute; written from the physics, not from its loop, which
 * resets leaves and divides inside the child loop):
 *   leaf:     M = sum m_j, c = sum m_j x_j / M, Q = sum m_j (3 d d^T - |d|^2 I), d = x_j - c
 *   internal: M = sum M_k, c = sum M_k c_k / M,
 *             Q = sum [M_k (3 d d^T - |d|^2 I) + Q_k], d = c_k - c
 * (the Q accumulation follows HashedOctree.cpp:785-820).  A one-body cell
 * keeps the body position as its centre of mass exactly. */
static void orc_quad_add(double q[6], double m, const double d[3])
{
    double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    q[0] += m * (3.0 * d[0] * d[0] - d2);
    q[1] += m * (3.0 * d[0] * d[1]);
    q[2] += m * (3.0 * d[0] * d[2]);
    q[3] += m * (3.0 * d[1] * d[1] - d2);
    q[4] += m * (3.0 * d[1] * d[2]);
    q[5] += m * (3.0 * d[2] * d[2] - d2);
}

static void orc_moments_rec(orc_tree *t, int64_t id, const double *pos, const double *mass)
{
    orc_node *nd = &t->nodes[id];
    double M = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = 0; k < 6; ++k) nd->quad[k] = 0.0;
    if (nd->child_byte == 0) {
        const int64_t *mem = t->members + nd->member_begin;
        for (int64_t k = 0; k < nd->member_count; ++k) {
            int64_t j = mem[k];
            M += mass[j];
            sx += mass[j] * pos[3 * j];
            sy += mass[j] * pos[3 * j + 1];
            sz += mass[j] * pos[3 * j + 2];
        }
        nd->mass = M;
        if (nd->member_count == 1) {
            int64_t j = mem[0];
            nd->com[0] = pos[3 * j]; nd->com[1] = pos[3 * j + 1]; nd->com[2] = pos[3 * j + 2];
        } else if (M > 0.0) {
            nd->com[0] = sx / M; nd->com[1] = sy / M; nd->com[2] = sz / M;
        }
        for (int64_t k = 0; k < nd->member_count; ++k) {
            int64_t j = mem[k];
            double d[3] = {pos[3 * j] - nd->com[0], pos[3 * j + 1] - nd->com[1], pos[3 * j + 2] - nd->com[2]};
            orc_quad_add(nd->quad, mass[j], d);
        }
        return;
    }
    int64_t kids[8];
    int nk = 0;
    for (unsigned o = 0; o < 8; ++o) {
        if (!(nd->child_byte & (1u << o))) continue;
        int64_t cid = orc_tree_lookup(t, (nd->key << 3) | o);
        orc_moments_rec(t, cid, pos, mass);
        nd = &t->nodes[id];
        kids[nk++] = cid;
    }
    for (int k = 0; k < nk; ++k) {
        const orc_node *ch = &t->nodes[kids[k]];
        M += ch->mass;
        sx += ch->mass * ch->com[0];
        sy += ch->mass * ch->com[1];
        sz += ch->mass * ch->com[2];
    }
    nd->mass = M;
    if (nd->n == 1) {
        const orc_node *ch = &t->nodes[kids[0]];
        nd->com[0] = ch->com[0]; nd->com[1] = ch->com[1]; nd->com[2] = ch->com[2];
    } else if (M > 0.0) {
        nd->com[0] = sx / M; nd->com[1] = sy / M; nd->com[2] = sz / M;
    }
    for (int k = 0; k < nk; ++k) {
        const orc_node *ch = &t->nodes[kids[k]];
        double d[3] = {ch->com[0] - nd->com[0], ch->com[1] - nd->com[1], ch->com[2] - nd->com[2]};
        orc_quad_add(nd->quad, ch->mass, d);
        for (int q = 0; q < 6; ++q) nd->quad[q] += ch->quad[q];
    }
}

void orc_tree_moments_intended(orc_tree *t, const double *pos, const double *mass)
{
    if (t->count) orc_moments_rec(t, 0, pos, mass);
}

/* ------------------------------------------------------------------------- */
/* walk and forces                                                            */
/* ------------------------------------------------------------------------- */

/* LinearHashedOctree::MAC — HashedOctree.cpp:872-885: size^2/d^2 < theta^2,
 * d measured from the point r to the node's baryCenter. */
int orc_mac_raw(double size, const double com[3], const double r[3], double theta)
{
    double dx = com[0] - r[0];
    double dy = com[1] - r[1];
    double dz = com[2] - r[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    double s2 = size * size;
#ifdef ORC_MUTANT_MAC_MUL
    return s2 < (theta * theta) * d2;
#else
    return (s2 / d2) < (theta * theta);
#endif
}

static inline int orc_mac(const orc_node *nd, const double p[3], double theta)
{
    return orc_mac_raw(nd->size, nd->com, p, theta);
}

/* One term of ComputeHOTForceInteractionList — HashedOctree.hpp:404-455:
 * softened monopole D1 = (1/sqrt(D2)) * (1/D2), plus the quadrupole
 * correction when the entry holds more than one body. */
static inline void orc_term(double a[3], const double p[3], const double c[3], double m,
                            int64_t nb, const double *q, double eps)
{
    double dx = c[0] - p[0], dy = c[1] - p[1], dz = c[2] - p[2];
    double D2 = dx * dx + dy * dy + dz * dz;
    D2 += eps * eps;
    double D1 = 1.0 / sqrt(D2);
    D2 = 1.0 / D2;
    D1 = D1 * D2;
    a[0] += m * dx * D1;
    a[1] += m * dy * D1;
    a[2] += m * dz * D1;
    if (q && nb > 1) {
        double qx = q[0] * dx + q[1] * dy + q[2] * dz;
        double qy = q[1] * dx + q[3] * dy + q[4] * dz;
        double qz = q[2] * dx + q[4] * dy + q[5] * dz;
        D1 *= D2;
        a[0] -= qx * D1;
        a[1] -= qy * D1;
        a[2] -= qz * D1;
        qx = dx * qx + dy * qy + dz * qz;
        qx *= 2.5;
        D1 *= D2;
        a[0] += qx * dx * D1;
        a[1] += qx * dy * D1;
        a[2] += qx * dz * D1;
    }
}

typedef struct { int64_t id; int body; } orc_entry; /* body=1: member body, id = body index */

/* TraverseHOTInteractionList — HashedOctree.hpp:336-385: explicit-stack DFS
 * from the root; children scanned in octant order 0..7 and either appended
 * (MAC passed) or pushed (LIFO).  A popped leaf is appended unless its
 * barycentre equals the body position exactly.  Bucket leaves (more than
 * one member) append each member except positions equal to the body's. */
static int64_t orc_list(const orc_tree *t, const double *pos, const double p[3], double theta,
                        int64_t *stack, orc_entry *list)
{
    if (t->count == 0) return 0;
    int64_t sp = 0, nl = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const orc_node *nd = &t->nodes[stack[--sp]];
        if (nd->child_byte) {
            for (unsigned o = 0; o < 8; ++o) {
                if (!(nd->child_byte & (1u << o))) continue;
                int64_t cid = orc_tree_lookup(t, (nd->key << 3) | o);
                if (orc_mac(&t->nodes[cid], p, theta)) { list[nl].id = cid; list[nl].body = 0; nl++; }
                else stack[sp++] = cid;
            }
        } else if (nd->member_count <= 1) {
            if (nd->com[0] != p[0] || nd->com[1] != p[1] || nd->com[2] != p[2]) {
                list[nl].id = nd - t->nodes; list[nl].body = 0; nl++;
            }
        } else {
            const int64_t *mem = t->members + nd->member_begin;
            for (int64_t k = 0; k < nd->member_count; ++k) {
                const double *q = pos + 3 * mem[k];
                if (q[0] != p[0] || q[1] != p[1] || q[2] != p[2]) {
                    list[nl].id = mem[k]; list[nl].body = 1; nl++;
                }
            }
        }
    }
    return nl;
}

/* One point's list in the reference's order: node index and kind (0: node,
 * 1: member body of a multi-body leaf, index = body). */
int64_t orc_tree_list(const orc_tree *t, const double *pos, const double p[3], double theta,
                      int64_t *ids, int32_t *is_body)
{
    size_t cap = t->count + t->nmembers + 8;
    int64_t *stack = (int64_t *)malloc(cap * sizeof(int64_t));
    orc_entry *list = (orc_entry *)malloc(cap * sizeof(orc_entry));
    int64_t nl = orc_list(t, pos, p, theta, stack, list);
    for (int64_t e = 0; e < nl; ++e) {
        if (ids) ids[e] = list[e].id;
        if (is_body) is_body[e] = list[e].body;
    }
    free(stack);
    free(list);
    return nl;
}

void orc_walk_subset(const orc_tree *t, const double *pos, const double *mass,
                     const int64_t *targets, size_t nt, double theta, double eps,
                     int use_quad, double *acc, int64_t *counts)
{
    size_t cap = t->count + t->nmembers + 8;
    int64_t *stack = (int64_t *)malloc(cap * sizeof(int64_t));
    orc_entry *list = (orc_entry *)malloc(cap * sizeof(orc_entry));
    for (size_t k = 0; k < nt; ++k) {
        int64_t i = targets ? targets[k] : (int64_t)k;
        const double *p = pos + 3 * i;
        int64_t nl = orc_list(t, pos, p, theta, stack, list);
        double a[3] = {0.0, 0.0, 0.0};
        for (int64_t e = 0; e < nl; ++e) {
            if (list[e].body) {
                orc_term(a, p, pos + 3 * list[e].id, mass[list[e].id], 1, NULL, eps);
            } else {
                const orc_node *nd = &t->nodes[list[e].id];
                orc_term(a, p, nd->com, nd->mass, nd->n, use_quad ? nd->quad : NULL, eps);
            }
        }
        acc[3 * k] = a[0]; acc[3 * k + 1] = a[1]; acc[3 * k + 2] = a[2];
        if (counts) counts[k] = nl;
    }
    free(stack);
    free(list);
}

/* ComputeHOTOctreeForce — HashedOctree.hpp:262-272 (acc reset first, ofApp.cpp:270). */
void orc_walk(const orc_tree *t, const double *pos, const double *mass, size_t n,
              double theta, double eps, int use_quad, double *acc, int64_t *counts)
{
    orc_walk_subset(t, pos, mass, NULL, n, theta, eps, use_quad, acc, counts);
}

/* FP64 direct sum (no reference counterpart): the same softened monopole
 * term over every other body, j != i by index.  pot = -sum m_j / sqrt(r^2+eps^2). */
void orc_direct(const double *pos, const double *mass, size_t n,
                const int64_t *targets, size_t nt, double eps, double *acc, double *pot)
{
    for (size_t k = 0; k < nt; ++k) {
        int64_t i = targets ? targets[k] : (int64_t)k;
        const double *p = pos + 3 * i;
        double a[3] = {0.0, 0.0, 0.0}, ph = 0.0;
        for (size_t j = 0; j < n; ++j) {
            if ((int64_t)j == i) continue;
            orc_term(a, p, pos + 3 * j, mass[j], 1, NULL, eps);
            double dx = pos[3 * j] - p[0], dy = pos[3 * j + 1] - p[1], dz = pos[3 * j + 2] - p[2];
            ph -= mass[j] / sqrt(dx * dx + dy * dy + dz * dz + eps * eps);
        }
        acc[3 * k] = a[0]; acc[3 * k + 1] = a[1]; acc[3 * k + 2] = a[2];
        if (pot) pot[k] = ph;
    }
}

/* ------------------------------------------------------------------------- */
/* one frame                                                                  */
/* ------------------------------------------------------------------------- */

static double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ofApp::update() then ofApp::draw() — ofApp.cpp:169-177, 248-282:
 *   bounds, keys and body sort on x_f; kick+drift to x_{f+1}; build the tree
 *   on x_{f+1} with the bounds of x_f; walk; kick.
 * permute_acc = 0 keeps the as-run behaviour where the acceleration array is
 * not reordered with the bodies (ofApp.cpp:176,248); 1 reorders it. */
void orc_frame(double *pos, double *vel, double *mass, double *acc, uint64_t *keys,
               size_t n, double dt, double theta, double eps, int leaf_cap, int mode,
               int use_quad, int permute_acc, orc_frame_stats *st)
{
    orc_frame_stats s;
    memset(&s, 0, sizeof s);
    double t0 = orc_now();
    double c[3], size;
    orc_bounds(pos, n, c, &size);
    double t1 = orc_now();
    orc_morton_keys(pos, n, size, keys);
    double t2 = orc_now();
    uint64_t *idx = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    for (size_t i = 0; i < n; ++i) idx[i] = i;
    orc_radix_sort(keys, idx, n);
    orc_gather(pos, 3, idx, n);
    orc_gather(vel, 3, idx, n);
    orc_gather(mass, 1, idx, n);
    if (permute_acc) orc_gather(acc, 3, idx, n);
    free(idx);
    double t3 = orc_now();
    orc_kick_drift(dt, pos, vel, acc, n);
    double t4 = orc_now();
    orc_tree *t = (leaf_cap <= 1) ? orc_tree_build_insert(pos, mass, n, c, size)
                                  : orc_tree_build_recursive(pos, mass, n, c, size, leaf_cap);
    if (mode == ORC_INTENDED) orc_tree_moments_intended(t, pos, mass);
    double t5 = orc_now();
    int64_t *cnt = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
    orc_walk(t, pos, mass, n, theta, eps, use_quad, acc, cnt);
    for (size_t i = 0; i < n; ++i) s.interactions += cnt[i];
    free(cnt);
    double t6 = orc_now();
    orc_kick(dt, vel, acc, n);
    double t7 = orc_now();
    s.nodes = (int64_t)orc_tree_size(t);
    orc_tree_free(t);
    s.t_bounds = t1 - t0; s.t_keys = t2 - t1; s.t_sort = t3 - t2; s.t_kick_drift = t4 - t3;
    s.t_build = t5 - t4; s.t_walk = t6 - t5; s.t_kick = t7 - t6;
    if (st) *st = s;
}

/* ------------------------------------------------------------------------- */
/* trajectory gradient (no reference counterpart: the differentiable form)   */
/* ------------------------------------------------------------------------- */

/* Softened direct-sum accelerations a_i = sum_{j != i} m_j d (|d|^2+eps^2)^-3/2,
 * d = x_j - x_i, in plain FP64 (the order of orc_direct). */
static void orc_acc_all(const double *pos, const double *mass, size_t n, double eps, double *acc)
{
    orc_direct(pos, mass, n, NULL, n, eps, acc, NULL);
}

/* Vector-Jacobian product of orc_acc_all: for weights w (3 per body),
 * gx += (da/dx)^T w and gm += (da/dm)^T w, pair by pair:
 *   J = m_j (r^-3 I - 3 r^-5 d d^T),  gx_i -= J w_i,  gx_j += J w_i,
 *   gm_j += r^-3 (d . w_i). */
static void orc_acc_vjp(const double *pos, const double *mass, size_t n, double eps, const double *w,
                        double *gx, double *gm)
{
    for (size_t i = 0; i < n; ++i) {
        const double *wi = w + 3 * i;
        for (size_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double d[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1], pos[3 * j + 2] - pos[3 * i + 2]};
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + eps * eps;
            double inv = 1.0 / sqrt(r2);
            double inv3 = inv * inv * inv;
            double inv5 = inv3 / r2;
            double dw = d[0] * wi[0] + d[1] * wi[1] + d[2] * wi[2];
            for (int a = 0; a < 3; ++a) {
                double jw = mass[j] * (inv3 * wi[a] - 3.0 * inv5 * dw * d[a]);
                gx[3 * i + a] -= jw;
                gx[3 * j + a] += jw;
            }
            gm[j] += inv3 * dw;
        }
    }
}

/* Leapfrog (kick-drift-kick) with direct-sum forces for nsteps frames:
 *   h = v + a dt/2; x += h dt; a = A(x); v = h + a dt/2.
 * a_0 is A(x_0) when init_acc, else the given acc (NULL = 0).  pos, vel, acc
 * are updated in place (acc must hold 3n doubles). */
void orc_traj_direct(double *pos, double *vel, const double *mass, double *acc, size_t n, double eps,
                     double dt, int nsteps, int init_acc)
{
    if (init_acc) orc_acc_all(pos, mass, n, eps, acc);
    for (int f = 0; f < nsteps; ++f) {
        for (size_t k = 0; k < 3 * n; ++k) {
            vel[k] = vel[k] + acc[k] * (0.5 * dt);
            pos[k] = pos[k] + vel[k] * dt;
        }
        orc_acc_all(pos, mass, n, eps, acc);
        for (size_t k = 0; k < 3 * n; ++k) vel[k] = vel[k] + acc[k] * (0.5 * dt);
    }
}

/* Reverse pass of orc_traj_direct from (pos, vel, mass, acc) at frame 0:
 * (gx, gv, ga) = dL/d(x_T, v_T, a_T) in, (gx0, gv0, ga0, gm) out.  States are
 * recomputed forward and kept for every frame (small n only). */
void orc_traj_direct_adjoint(const double *pos0, const double *vel0, const double *mass, const double *acc0,
                             size_t n, double eps, double dt, int nsteps, int init_acc,
                             const double *gx, const double *gv, const double *ga,
                             double *gx0, double *gv0, double *ga0, double *gm)
{
    const size_t m3 = 3 * n;
    double *xs = malloc(sizeof(double) * m3 * (size_t)(nsteps + 1));
    double *v = malloc(sizeof(double) * m3), *a = malloc(sizeof(double) * m3);
    double *X = malloc(sizeof(double) * m3), *V = malloc(sizeof(double) * m3), *A = malloc(sizeof(double) * m3);
    double *w = malloc(sizeof(double) * m3);
    memcpy(xs, pos0, sizeof(double) * m3);
    memcpy(v, vel0, sizeof(double) * m3);
    if (acc0) memcpy(a, acc0, sizeof(double) * m3); else memset(a, 0, sizeof(double) * m3);
    if (init_acc) orc_acc_all(xs, mass, n, eps, a);
    for (int f = 0; f < nsteps; ++f) {
        double *x = xs + m3 * (size_t)(f + 1);
        memcpy(x, xs + m3 * (size_t)f, sizeof(double) * m3);
        orc_traj_direct(x, v, mass, a, n, eps, dt, 1, 0);
    }
    for (size_t k = 0; k < m3; ++k) {
        X[k] = gx ? gx[k] : 0.0;
        V[k] = gv ? gv[k] : 0.0;
        A[k] = ga ? ga[k] : 0.0;
    }
    memset(gm, 0, sizeof(double) * n);
    for (int f = nsteps - 1; f >= 0; --f) {
        const double *y = xs + m3 * (size_t)(f + 1);
        for (size_t k = 0; k < m3; ++k) w[k] = A[k] + V[k] * (0.5 * dt);
        orc_acc_vjp(y, mass, n, eps, w, X, gm);          /* X += (dA/dy)^T w */
        for (size_t k = 0; k < m3; ++k) {
            double h = V[k] + X[k] * dt;
            V[k] = h;
            A[k] = h * (0.5 * dt);
        }
    }
    if (init_acc) {
        orc_acc_vjp(xs, mass, n, eps, A, X, gm);
        memset(A, 0, sizeof(double) * m3);
    }
    memcpy(gx0, X, sizeof(double) * m3);
    memcpy(gv0, V, sizeof(double) * m3);
    if (ga0) memcpy(ga0, A, sizeof(double) * m3);
    free(xs); free(v); free(a); free(X); free(V); free(A); free(w);
}
