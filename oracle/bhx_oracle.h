/*
 * bhx_oracle.h — serial CPU restatement of the reference Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product (libbhx) never links, calls or falls back to it.
 *
 * Every function restates one piece of the reference
 * (DavidRichardson02/Serial-Hashed-Octree-Barnes-Hut-N-Body-Simulator-0_3,
 * files under src/) in plain C with the same FP64 operation order, compiled
 * with -ffp-contract=off so that no multiply-add is fused.  File:line
 * citations are given at each definition in bhx_oracle.c.
 */
#ifndef BHX_ORACLE_H
#define BHX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Node payload, one record per octree cell (HOTNode, HashedNode.hpp:60-95). */
typedef struct orc_node {
    uint64_t key;          /* path key: 1 sentinel bit, then 3 bits per level      */
    double   center[3];    /* nodeBounds.center                                     */
    double   size;         /* nodeBounds.size (used as the full edge)               */
    double   com[3];       /* baryCenter                                            */
    double   mass;         /* mass                                                  */
    double   quad[6];      /* Qxx Qxy Qxz Qyy Qyz Qzz                               */
    int64_t  n;            /* N: bodies at or below the node                        */
    int64_t  first;        /* smallest array index in the cell                      */
    int32_t  level;        /* depth, root = 0                                       */
    uint32_t child_byte;   /* bit o set iff child in octant o exists                */
    int64_t  member_begin; /* leaves: range into the tree's member list             */
    int64_t  member_count;
} orc_node;

typedef struct orc_tree orc_tree;

enum { ORC_AS_RUN = 0, ORC_INTENDED = 1 };
enum { ORC_MAX_LEVEL = 21 };

/* ---- bounds, keys, sort, reorder, KDK ---- */
void     orc_bounds(const double *pos, size_t n, double center[3], double *size);
uint64_t orc_x86_cvt_u64(double u);
uint64_t orc_spread21(uint64_t t);
uint64_t orc_morton_key(const double p[3], double size);
void     orc_morton_keys(const double *pos, size_t n, double size, uint64_t *keys);
void     orc_decode_morton_key(uint64_t key, double size, double out[3]);
void     orc_radix_sort(uint64_t *keys, uint64_t *idx, size_t n);
void     orc_gather(double *arr, size_t width, const uint64_t *idx, size_t n);
void     orc_kick_drift(double dt, double *pos, double *vel, const double *acc, size_t n);
void     orc_kick(double dt, double *vel, const double *acc, size_t n);

/* ---- octree ---- */
uint64_t  orc_tree_path_key(const double p[3], const double c[3], double s);
void      orc_tree_path_keys(const double *pos, size_t n, const double c[3], double s, uint64_t *keys);
orc_tree *orc_tree_build_insert(const double *pos, const double *mass, size_t n,
                                const double c[3], double s);
orc_tree *orc_tree_build_recursive(const double *pos, const double *mass, size_t n,
                                   const double c[3], double s, int leaf_cap);
void      orc_tree_moments_intended(orc_tree *t, const double *pos, const double *mass);
size_t    orc_tree_size(const orc_tree *t);
size_t    orc_tree_unsplit(const orc_tree *t);
void      orc_tree_nodes(const orc_tree *t, orc_node *out);
int64_t   orc_tree_lookup(const orc_tree *t, uint64_t key);
void      orc_tree_free(orc_tree *t);

/* ---- walk and forces ---- */
int  orc_mac_raw(double size, const double com[3], const double r[3], double theta);
int64_t orc_tree_list(const orc_tree *t, const double *pos, const double p[3], double theta,
                      int64_t *ids, int32_t *is_body);
void orc_walk(const orc_tree *t, const double *pos, const double *mass, size_t n,
              double theta, double eps, int use_quad, double *acc, int64_t *counts);
void orc_walk_subset(const orc_tree *t, const double *pos, const double *mass,
                     const int64_t *targets, size_t nt, double theta, double eps,
                     int use_quad, double *acc, int64_t *counts);
void orc_direct(const double *pos, const double *mass, size_t n,
                const int64_t *targets, size_t nt, double eps, double *acc, double *pot);

/* ---- one frame in the reference order (update() then draw()) ---- */
typedef struct orc_frame_stats {
    double t_bounds, t_keys, t_sort, t_kick_drift, t_build, t_walk, t_kick;
    int64_t nodes;
    int64_t interactions;
} orc_frame_stats;

void orc_frame(double *pos, double *vel, double *mass, double *acc, uint64_t *keys,
               size_t n, double dt, double theta, double eps, int leaf_cap, int mode,
               int use_quad, int permute_acc, orc_frame_stats *st);

/* ---- trajectory gradient with direct-sum forces (differentiable form) ---- */
void orc_traj_direct(double *pos, double *vel, const double *mass, double *acc, size_t n, double eps,
                     double dt, int nsteps, int init_acc);
void orc_traj_direct_adjoint(const double *pos0, const double *vel0, const double *mass, const double *acc0,
                             size_t n, double eps, double dt, int nsteps, int init_acc,
                             const double *gx, const double *gv, const double *ga,
                             double *gx0, double *gv0, double *ga0, double *gm);

#ifdef __cplusplus
}
#endif
#endif
