/*
 * bhx.h — C ABI of the MI355X Barnes-Hut engine (libbhx.so).
 *
 * Plain pointers and sizes only; every entry point returns a bhx_status.
 * Host arrays of 3-vectors are interleaved xyz doubles (the memory layout
 * of the reference's Vec3D[]), masses are doubles, keys are uint64.
 *
 * Each entry point replaces one function of the reference's per-step path
 * (DavidRichardson02/Serial-Hashed-Octree-Barnes-Hut-N-Body-Simulator-0_3,
 * src/); the replaced interface is cited at each declaration.  State lives
 * in a context that owns all device buffers; bodies stay resident in HBM
 * between calls, in structure-of-arrays FP64 form.
 */
#ifndef BHX_H
#define BHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bhx_status {
    BHX_OK = 0,
    BHX_E_NULL = 1,        /* a required pointer was NULL                         */
    BHX_E_ARG = 2,         /* an argument is out of range                         */
    BHX_E_DEPTH = 3,       /* cells at the 21-level cap hold more bodies than fit  */
    BHX_E_HASH_FULL = 4,   /* node hash table could not take every node            */
    BHX_E_NONFINITE = 5,   /* non-finite bounds, or a step met a non-finite state  */
    BHX_E_HIP = 6,         /* a HIP runtime call failed                            */
    BHX_E_NOMEM = 7,       /* device allocation failed                             */
    BHX_E_STATE = 8,       /* call out of order (e.g. walk before build)           */
    BHX_E_CAPACITY = 9,    /* more bodies than the context was created for         */
    BHX_E_SORT = 10,       /* radix-sort look-back did not complete                */
    BHX_E_IO = 11          /* a file could not be opened, read, written or parsed   */
} bhx_status;

/* Node payload semantics. */
typedef enum bhx_mode {
    BHX_MODE_AS_RUN = 0,   /* as the reference executes: barycentre = first body in
                              the cell (array order), quadrupole = 0               */
    BHX_MODE_INTENDED = 1  /* true mass-weighted centre of mass and quadrupole     */
} bhx_mode;

typedef enum bhx_precision {
    BHX_FP64 = 0,          /* FP64 walk with the reference's operation sequence     */
    BHX_FP32 = 1           /* FP32 walk in group-relative coordinates               */
} bhx_precision;

typedef struct bhx_config {
    double theta;          /* opening angle: node opens unless size^2/d^2 < theta^2 */
    double eps;            /* Plummer softening                                     */
    double dt;             /* leapfrog step                                         */
    int32_t leaf_capacity; /* bodies per leaf before a cell splits (1..64); 1 = reference */
    int32_t mode;          /* bhx_mode                                              */
    int32_t precision;     /* bhx_precision                                         */
    int32_t quadrupole;    /* 1: add the quadrupole term for cells of > 1 body      */
    int32_t permute_acc;   /* 1: reorder accelerations with the bodies; 0: as run   */
    int32_t compensated;   /* 1: compensated (Kahan) FP64 kick/drift accumulation   */
    int32_t device;        /* HIP device ordinal                                    */
    int32_t group_walk;    /* 0: per-body opening test, the reference's lists exactly;
                              1: one conservative test per 64-body group (a cell is
                              accepted only if every body of the group accepts it),
                              shared interaction lists evaluated densely            */
    int32_t geometry;      /* cell geometry of the tree.  0: the reference's — the
                              root edge is the bounds size (a half-extent), child
                              edge = edge/2, child centre = centre +- edge/4
                              (HashedNode.cpp:50-56), so outer bodies can lie
                              outside every cell.  1: enclosing cube — the same
                              recurrence from a root edge of twice the size, so
                              every cell contains its bodies and the opening test
                              uses the true edge.  Morton keys are unchanged.       */
    int32_t group_size;    /* FP32 group walk: bodies that share one interaction list,
                              64, 32 or 16 (one 64-body wave walks 1, 2 or 4 lists);
                              smaller groups evaluate fewer pairs but keep more
                              lists (bhx_autotune measures which is faster).  0: 64 */
} bhx_config;

typedef struct bhx_ctx bhx_ctx;

/* Downloaded node record (HOTNode, HashedNode.hpp:60-95). */
typedef struct bhx_node {
    uint64_t key;          /* nodeKey: sentinel 1 then 3 bits per level             */
    double   center[3];    /* nodeBounds.center                                     */
    double   size;         /* nodeBounds.size                                       */
    double   com[3];       /* baryCenter                                            */
    double   mass;         /* mass                                                  */
    double   quad[6];      /* quadrupoleMoment                                      */
    int64_t  n;            /* N                                                     */
    int64_t  first;        /* smallest array index among the node's bodies          */
    int32_t  level;        /* depth (root 0)                                        */
    uint32_t child_byte;   /* childByte                                             */
    int64_t  skip;         /* pre-order index of the next node outside the subtree  */
    int64_t  body_start;   /* first body of the node in tree order                  */
} bhx_node;

typedef struct bhx_timing {   /* milliseconds of the last bhx_step, per phase; with the group walks
                                 the closing kick is applied by the walk kernel (walk includes it,
                                 kick reads 0) */
    double bounds, keys, sort, kick_drift, treekeys, build, moments, walk, kick, total;
} bhx_timing;

/* ---- context ---------------------------------------------------------- */

/* Defaults: theta 0.5, eps 0.025 (SOFTENING, HashedOctree.hpp:151), dt 1e-3
 * (ofApp.h:62), leaf capacity 1, AS_RUN, FP64, no quadrupole, permute_acc 1. */
int bhx_config_default(bhx_config *cfg);

/* LinearHashedOctree() (HashedOctree.cpp:24-27) plus the caller-owned Body[] /
 * Vec3D[] buffers of ofApp::setup (ofApp.cpp:83-84,136): one context holds up
 * to `capacity` bodies on the configured device. */
int bhx_create(const bhx_config *cfg, size_t capacity, bhx_ctx **out);
int bhx_destroy(bhx_ctx *ctx);
int bhx_set_config(bhx_ctx *ctx, const bhx_config *cfg);
int bhx_get_config(const bhx_ctx *ctx, bhx_config *cfg);
const char *bhx_last_error(const bhx_ctx *ctx);
int bhx_sync(bhx_ctx *ctx);

/* Body[] contents (Body.hpp:30-65): positions, velocities (xyz interleaved),
 * masses; accelerations start at zero as `new Vec3D[N]` does (ofApp.cpp:136).
 * vel and acc may be NULL (zero). */
int bhx_set_bodies(bhx_ctx *ctx, const double *pos, const double *vel, const double *mass, size_t n);
int bhx_set_accelerations(bhx_ctx *ctx, const double *acc);
int bhx_get_bodies(bhx_ctx *ctx, double *pos, double *vel, double *mass, double *acc);
int bhx_num_bodies(const bhx_ctx *ctx, size_t *n);

/* Compensation terms of the compensated leapfrog (config `compensated`): the
 * rounding error carried per body for position and velocity (3 per body,
 * xyz interleaved); part of the state a frame continues from.  NULL skips
 * (get) or zeroes (set).  bhx_set_bodies zeroes them.  (No reference
 * counterpart: the reference's leapfrog is uncompensated.) */
int bhx_get_compensation(bhx_ctx *ctx, double *pos_c, double *vel_c);
int bhx_set_compensation(bhx_ctx *ctx, const double *pos_c, const double *vel_c);

/* ---- the per-step functions ---------------------------------------------- */

/* OctantBounds(Body*, size_t) — OctantUtils.cpp:27-59: centre = sequential
 * FP64 mean in array order, size = max(0, max |p| - centre). */
int bhx_bounds(bhx_ctx *ctx, double center[3], double *size);

/* computeMortonKey over all bodies — MortonKeys.hpp:137-181, ofApp.cpp:172-175.
 * keys_out may be NULL (keys stay on the device). */
int bhx_morton_keys(bhx_ctx *ctx, double size, uint64_t *keys_out);

/* ThreePassRadixSortMortonKeys(keys, indexes, n) — MortonKeys.hpp:439-470:
 * stable ascending sort of host arrays; idx is carried along (u64, as size_t). */
int bhx_sort_keys(bhx_ctx *ctx, uint64_t *keys, uint64_t *idx, size_t n);

/* RadixSortBodies(Body*, n) — Body.hpp:275-302 (+ ReorderBodiesByIndexes
 * :257-273): sorts the resident bodies by the keys of the last
 * bhx_morton_keys call.  perm_out (may be NULL) receives new[i] = old[perm[i]]. */
int bhx_sort_bodies(bhx_ctx *ctx, uint64_t *perm_out);

/* bhx_sort_keys on DEVICE arrays with no allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  keys
 * and idx are device pointers to n values on `device` (a HIP device ordinal),
 * sorted in place on `stream` (a hipStream_t; NULL = the null stream) into the
 * same stable ascending order as bhx_sort_keys, idx carried along.  All
 * working memory is a static device array of about 25 MiB per device, held
 * until the process ends; it fixes the maximum n at BHX_NOALLOC_MAX_KEYS
 * (BHX_E_CAPACITY beyond).  Calls on one device share that array: they are
 * serialised and each returns when its sort is done (the stream is
 * synchronised).  BHX_E_SORT: the look-back did not complete. */
#define BHX_NOALLOC_MAX_KEYS ((size_t)1 << 20)
int bhx_sort_keys_noalloc(int device, void *stream, uint64_t *keys, uint64_t *idx, size_t n);

/* ReorderBodiesByIndexes(Body*, size_t*, n) — Body.hpp:257-273 with a caller
 * permutation. */
int bhx_reorder_bodies(bhx_ctx *ctx, const uint64_t *perm);

/* ComputePositionAtHalfTimeStep — Body.hpp:306-313: v += (a dt) 0.5; x += v dt. */
int bhx_kick_drift(bhx_ctx *ctx, double dt);

/* ComputeVelocityAndPosition — Body.hpp:317-324: v += (a dt) 0.5. */
int bhx_kick(bhx_ctx *ctx, double dt);

/* buildLinearHashedOctreeInPlace(tree, bodies, n, bounds) — HashedOctree.hpp:
 * 211-244 and HashedOctree.cpp:450-711: node set, payload and hash table for
 * the resident bodies in the given root bounds.  n_nodes may be NULL. */
int bhx_build(bhx_ctx *ctx, const double center[3], double size, size_t *n_nodes);

/* Node records of the last build, in pre-order (depth first, children in
 * octant order).  `cap` is the capacity of out in records. */
int bhx_get_nodes(bhx_ctx *ctx, bhx_node *out, size_t cap, size_t *n_nodes);

/* LinearHashedOctree::lookUpNode(key) — HashedOctree.cpp:111-115, batched
 * through the device hash table: ids[i] = pre-order index or -1. */
int bhx_lookup(bhx_ctx *ctx, const uint64_t *keys, size_t n, int64_t *ids);

/* Bodies in the order the tree holds them: tree_perm[i] = array index of the
 * i-th body in tree order. */
int bhx_get_tree_order(bhx_ctx *ctx, uint64_t *tree_perm, uint64_t *tree_keys);

/* Number of leaves at the depth cap that hold more than leaf_capacity bodies. */
int bhx_unsplit_leaves(bhx_ctx *ctx, size_t *count);

/* ComputeHOTOctreeForce — HashedOctree.hpp:262-272, TraverseHOTInteractionList
 * :336-385, MAC (HashedOctree.cpp:872-885), ComputeHOTForceInteractionList
 * HashedOctree.hpp:393-459: accelerations of all resident bodies from the last
 * build (array order).  acc_out and counts (interaction-list length per body,
 * the value TraverseHOTInteractionList returns) may be NULL. */
int bhx_walk(bhx_ctx *ctx, double *acc_out, uint64_t *counts);

/* TraverseHOTInteractionList's return value (HashedOctree.hpp:336-385) for
 * every body (array order): the length of the reference's own per-body
 * interaction list on the last tree, whichever walk the configuration
 * selects.  The resident accelerations are left unchanged. */
int bhx_interaction_counts(bhx_ctx *ctx, uint64_t *counts);

/* The same per-body reference lists, split for ComputeHOTForceInteractionList
 * (HashedOctree.hpp:393-459): counts as above, and quad_counts[i] = how many of
 * body i's entries take the quadrupole branch (an accepted node of more than
 * one body, quadrupoles on; 0 with them off).  Accelerations unchanged. */
int bhx_interaction_counts_quad(bhx_ctx *ctx, uint64_t *counts, uint64_t *quad_counts);

/* ---- the reference's per-body tree API -------------------------------------
 * ComputeHOTOctreeForce (HashedOctree.hpp:262-272) is bhx_walk, batched.  Code
 * that calls the per-body pair itself, the MAC, or edits the node table uses
 * these, on the node table of the last build (one system).  Node indexes are
 * the pre-order indexes of bhx_get_nodes; a record whose key reads 0 there was
 * erased.  Nodes put or erased since the build leave bhx_walk refusing
 * (BHX_E_STATE) until the next build; these calls see the edits. */

/* TraverseHOTInteractionList(tree, bodyPosition, theta, walkList, interactList)
 * → long (HashedOctree.hpp:336-385) with LinearHashedOctree::MAC
 * (HashedOctree.cpp:872-885), for nq points (xyz interleaved): the explicit-
 * stack depth-first search from the root in the reference's order (children
 * scanned in octant order through the hash table, accepted children
 * appended, opened ones pushed, the last pushed popped first; a popped leaf
 * appended unless its barycentre equals the point).  Point q's list is
 * entries[offsets[q] .. offsets[q+1]) (offsets: nq + 1 values, offsets[0] =
 * 0; offsets[q+1] - offsets[q] is the reference's return value).  An entry
 * >= 0 is a node index; an entry < 0 is -(1 + i) for body i (array order) of
 * a leaf that holds several bodies (leaf capacity > 1, or a cell at the
 * 21-level cap), each such body except one at the point itself, in tree
 * order.  entries NULL: only offsets and *total; else cap must be >= total
 * (BHX_E_CAPACITY).  The walkList / interactList scratch of the reference is
 * device memory here. */
int bhx_interaction_lists(bhx_ctx *ctx, const double *points, size_t nq, double theta, uint64_t *offsets,
                          int64_t *entries, size_t cap, size_t *total);

/* ComputeHOTForceInteractionList(bodyPosition, bodyMass, acceleration,
 * interactList, listLength) — HashedOctree.hpp:393-459: acc (3 per point, in
 * and out) += the softened terms of each point's entries (as returned by
 * bhx_interaction_lists) in list order, with the reference's FP64 operation
 * sequence and the configured eps (the reference's SOFTENING); the
 * quadrupole correction for nodes of more than one body when the
 * configuration enables it.  Offsets need not start at 0 (nq + 1 values). */
int bhx_force_lists(bhx_ctx *ctx, const double *points, size_t nq, const uint64_t *offsets,
                    const int64_t *entries, double *acc);

/* LinearHashedOctree::MAC(node, r, theta) — HashedOctree.cpp:872-885:
 * accept[k] = (size^2 / |baryCenter - r|^2 < theta^2) for node node_ids[k] and
 * point k (size: the cell edge of the node's level). */
int bhx_mac(bhx_ctx *ctx, const int64_t *node_ids, const double *points, size_t n, double theta,
            uint8_t *accept);

/* LinearHashedOctree::insertHOTNode(node) — HashedOctree.cpp:152-205: the
 * record (key, center, com, mass, quad, n, first, level, child_byte,
 * body_start; level must match the key) replaces the node of that key or is
 * appended as a new node and entered into the hash table; *id_out receives its
 * index.  A node's size follows from its level (the reference's recurrence);
 * the record's size field is not stored. */
int bhx_node_put(bhx_ctx *ctx, const bhx_node *node, int64_t *id_out);

/* LinearHashedOctree::deleteNode(key) — HashedOctree.cpp:647-659: the node
 * leaves the hash table (lookups miss it; parents keep their child bits, and
 * traversals skip the missing child).  *erased = 1 if it was there. */
int bhx_node_erase(bhx_ctx *ctx, uint64_t key, int *erased);

/* pruneEmptyNodes(lookUpNode(key)) — HashedOctree.cpp:610-639, and
 * PruneEmptyNodesFromTree (HashedOctree.hpp:248-252) with key = ROOT_KEY (1):
 * every node at or below `key` with N = 0 and no children is erased. */
int bhx_prune_empty(bhx_ctx *ctx, uint64_t key, size_t *removed);

/* decodeMortonKey(key, size) — MortonKeys.hpp:204-228 for n keys: the three
 * 21-bit coordinates scaled back as c / 2^21 * 2 * size - size (xyz
 * interleaved). */
int bhx_decode_morton_keys(bhx_ctx *ctx, const uint64_t *keys, size_t n, double size, double *pos);

/* One frame in the reference order (ofApp::update then ofApp::draw,
 * ofApp.cpp:169-177,248-282), repeated nsteps times on the resident bodies.
 * Returns BHX_E_NONFINITE when some frame met a non-finite position (at its
 * keys) or made a non-finite velocity (at its closing kick); the bodies are
 * left as the frames made them.  (The reference checks neither.) */
int bhx_step(bhx_ctx *ctx, int nsteps);

/* bhx_step with every frame replayed from a captured HIP graph: no host wait
 * between frames (the node count is checked once at the end; when a frame
 * needed more nodes than the tree holds, the call restarts from a device copy
 * of its starting state with larger arrays).  Results are identical to
 * bhx_step.  (No reference counterpart.) */
int bhx_step_graph(bhx_ctx *ctx, int nsteps);
int bhx_last_timing(bhx_ctx *ctx, bhx_timing *t);
/* 1: time each phase with events in bhx_step (adds synchronisation). */
int bhx_set_profiling(bhx_ctx *ctx, int on);

/* ---- ensembles of independent systems (no reference counterpart) -------- */

/* Splits the resident bodies into nsys independent systems: system k is the
 * contiguous range [offsets[k], offsets[k+1]) (offsets[0] = 0,
 * offsets[nsys] = n, every system non-empty, nsys <= 65536).  From then on
 * bounds, keys, sorts, trees and walks are per system — each system steps
 * exactly as it would alone in its own context — and bodies never leave
 * their system's range.  bhx_set_bodies resets the context to one system. */
int bhx_set_systems(bhx_ctx *ctx, const uint64_t *offsets, size_t nsys);
int bhx_num_systems(const bhx_ctx *ctx, size_t *nsys);

/* OctantBounds of every system (centers: 3 per system, sizes: 1 per system;
 * either may be NULL). */
int bhx_system_bounds(bhx_ctx *ctx, double *centers, double *sizes);

/* bhx_build with one root box per system. */
int bhx_build_systems(bhx_ctx *ctx, const double *centers, const double *sizes, size_t *n_nodes);

/* The current system offsets (nsys + 1 values, offsets[0] = 0). */
int bhx_get_systems(const bhx_ctx *ctx, uint64_t *offsets);

/* Pre-order index of every system's root in the last tree. */
int bhx_system_roots(bhx_ctx *ctx, uint64_t *roots);

/* bhx_lookup in one system's tree (bhx_lookup is system 0). */
int bhx_lookup_in(bhx_ctx *ctx, size_t system, const uint64_t *keys, size_t n, int64_t *ids);

/* One candidate of bhx_autotune. */
typedef struct bhx_tune_entry {
    int32_t leaf_capacity;
    int32_t group_size;     /* bodies per interaction list of the FP32 group walk */
    double ms_per_step;     /* wall time per bhx_step_graph frame */
    double median_err;      /* median |a - a*| / |a*| on the sample, a* = FP64 direct sum */
} bhx_tune_entry;

/* Picks the walk tiling: the leaf capacity (bodies per leaf bucket, which
 * sets how much of each list is direct body terms) and the group size
 * (bodies sharing one interaction list: 64, 32 or 16, see bhx_config).  From
 * the current state, for each of the ncaps x ngroups candidates runs `steps`
 * graph frames, times them and measures the median force error of the last
 * frame on `nsample` seeded bodies; keeps the fastest candidate whose median
 * error is <= max_err (the configured pair if none is).  group_sizes NULL
 * (or ngroups 0): the configured group size only.  The bodies are left in
 * their starting state; out (may be NULL) receives one entry per candidate
 * (capacity-major); chosen_group may be NULL. */
int bhx_autotune(bhx_ctx *ctx, const int32_t *caps, size_t ncaps, const int32_t *group_sizes, size_t ngroups,
                 int steps, double max_err, size_t nsample, uint64_t seed, bhx_tune_entry *out, int32_t *chosen,
                 int32_t *chosen_group);

/* ---- differentiation (no reference counterpart) -------------------------- */

/* Runs nsteps frames from the resident state (as bhx_step_graph) and records
 * them for bhx_adjoint: a device checkpoint of the state every
 * checkpoint_every frames, and, when it fits the memory set by
 * bhx_set_record_memory, every frame's walk data (tree, body orders,
 * payloads: about 60 bytes per body and 130 per node), so that bhx_adjoint
 * re-runs no frames.  init_acc != 0 first sets the accelerations to the
 * walk of the starting positions (a_0 = A(x_0)); otherwise the resident
 * accelerations are the frames' a_0.  Needs group_walk = 1, the intended
 * payload, permute_acc = 1 and one list per 64-body group (group_size 0 or
 * 64); monopole or quadrupole cells, FP64 or FP32 walk.  Changing the
 * configuration drops the recording. */
int bhx_record(bhx_ctx *ctx, int nsteps, int checkpoint_every, int init_acc);

/* Device memory (bytes) bhx_record may use to keep every frame's walk data;
 * 0: keep none (the reverse pass re-runs each checkpoint segment);
 * UINT64_MAX (the default): half of the device memory free at recording. */
int bhx_set_record_memory(bhx_ctx *ctx, uint64_t bytes);

/* Entries of each walk group's interaction list the reverse pass stores for
 * its evaluation kernel (0: the default, 8,192, fewer when there are many
 * groups; else a multiple of 64, at least 64).  A longer list is evaluated
 * by the traversal itself; the gradients agree to rounding. */
int bhx_set_adjoint_list_capacity(bhx_ctx *ctx, uint32_t entries);

/* Whether the last recording kept every frame's walk data (*kept = 1) or
 * its reverse pass re-runs the checkpoint segments (*kept = 0). */
int bhx_record_keeps_frames(bhx_ctx *ctx, int *kept);

/* State after the recorded frames, in the body order at the start of the
 * recording (any pointer may be NULL). */
int bhx_recorded_bodies(bhx_ctx *ctx, double *pos, double *vel, double *mass, double *acc);

/* Decision digest of the last recording (FP64 walk only): a wrapping sum of
 * 64-bit hashes over every recorded tree (each body's starting identity,
 * tree slot, leaf node and leaf key) and every group's interaction list
 * (each accepted cell and opened leaf with its group), each mixed with its
 * frame index, the a_0 walk included.  Two recordings whose digests are
 * equal walked the same trees and lists in every frame (up to a 64-bit hash
 * collision): the same frozen decisions, so finite differences between them
 * see the derivative bhx_adjoint computes. */
int bhx_record_digest(bhx_ctx *ctx, uint64_t *digest);

/* Reverse pass of the recorded frames: from dL/dx_T, dL/dv_T, dL/da_T (each
 * 3 per body in the starting order, NULL = 0) to dL/dx_0, dL/dv_0, dL/da_0
 * (3 per body) and dL/dm (1 per body), all in the starting order; NULL
 * outputs are skipped.  Every walk is differentiated with the tree, the
 * interaction lists and the body orders of its recorded frame (frozen
 * decisions), read from the recording's stored frames or, when the recording
 * did not keep them, from the frames of each checkpoint segment re-run
 * bitwise as recorded.  dL/da_0 is 0 when bhx_record computed a_0 (its
 * adjoint flows into dL/dx_0 and dL/dm).  The bodies are left at the end of
 * the recording.  Every sum has a fixed order (no atomics), so results repeat
 * bit for bit. */
int bhx_adjoint(bhx_ctx *ctx, const double *gx, const double *gv, const double *ga,
                double *gx0, double *gv0, double *gm, double *ga0);

/* Wall time of the last bhx_record and bhx_adjoint calls. */
int bhx_grad_timing(bhx_ctx *ctx, double *record_ms, double *adjoint_ms);

/* ---- diagnostics (no reference counterpart) ------------------------------ */

/* FP64 direct sum over the target's own system for each target (array
 * indices). */
int bhx_direct(bhx_ctx *ctx, const int64_t *targets, size_t nt, double eps,
               double *acc_out, double *pot_out);

/* Which evaluation the last one-system OctantBounds centre took (bhx_bounds,
 * bhx_system_bounds with one system, or a frame's bounds): out[0..2] = tiles
 * of x, y, z whose partial sums had to be added element by element, out[3] =
 * 1 when the plain left-to-right chain ran (more than three quarters of some
 * axis's tiles were such tiles), 0 when the chain over exact runs ran.  Both
 * give the same bits; this reports the path for tests and profiling.  (No
 * reference counterpart.) */
int bhx_bounds_path(bhx_ctx *ctx, uint32_t out[4]);

/* FP64 kinetic energy, linear and angular momentum of the resident bodies. */
int bhx_diagnostics(bhx_ctx *ctx, double *ekin, double mom[3], double angmom[3]);

/* Energies of the resident bodies: kinetic K = sum m v^2 / 2 (FP64),
 * potential W = sum m phi / 2 summed in FP64 over phi from the group walk of
 * the configured precision on a tree built for the current positions
 * (softened; cell quadrupoles when configured; a body at zero separation —
 * the body itself — adds nothing), and the virial ratio 2K/|W|.  pot_out
 * (n doubles, array order; may be NULL) receives phi.  The accelerations are
 * left as they are; the last tree is replaced. */
int bhx_energy(bhx_ctx *ctx, double *ekin, double *epot, double *virial, double *pot_out);

/* ---- initial conditions (no reference counterpart) ----------------------- */

/* Plummer sphere (Aarseth, Henon & Wielen 1974) with counter-based Philox
 * random numbers: n equal masses summing to total_mass, scale radius a,
 * isotropic velocities in virial equilibrium (G = 1), mass fraction cut
 * 0.999, shifted to the centre-of-mass frame. */
int bhx_ic_plummer(size_t n, uint64_t seed, double total_mass, double a,
                   double *pos, double *vel, double *mass);

/* Ensemble of Plummer systems: N_k = round(exp(U[ln nmin, ln nmax])), total
 * mass and scale radius 1 + jitter * U[-1, 1] (masses, radii may be NULL). */
int bhx_ic_ensemble_sizes(size_t nsys, uint64_t seed, uint64_t nmin, uint64_t nmax, double jitter,
                          uint64_t *sizes, double *masses, double *radii);

/* Bodies of that ensemble, system after system (system k from
 * bhx_ic_plummer(N_k, seed + 1 + k, M_k, a_k)), on up to `threads` host
 * threads (0: all). */
int bhx_ic_ensemble(size_t nsys, uint64_t seed, const uint64_t *sizes, const double *masses,
                    const double *radii, int threads, double *pos, double *vel, double *mass);

/* Two disk galaxies on a parabolic orbit (G = 1).  Each galaxy: an
 * exponential disk (surface density ~ exp(-R/R_d), vertical sech^2(z/z_d),
 * cut at disk_cut R_d) holding disk_fraction of its mass and bodies, and a
 * Hernquist halo (scale a, cut at halo_cut a) holding the rest.  Halo
 * velocities are drawn from the isotropic distribution function of the
 * untruncated Hernquist halo (Hernquist 1990, eq. 17) with the halo's mass;
 * disk velocities are the circular speed of halo + exponential disk (Freeman
 * 1970) plus Gaussian dispersion disk_sigma * v_c per component.  The
 * secondary has mass total / (1 + mass_ratio) and lengths scaled by
 * (M_2 / M_1)^(1/3).  Bodies are equal-mass, so the body counts split as the
 * masses.  The galaxies start `separation` apart on the parabolic orbit with
 * the given pericentre, in the x-y plane, each shifted to its own centre of
 * mass first; disk k is inclined by incl_k about the x axis.  The pair is
 * not in exact equilibrium (each component's velocities ignore the other's
 * truncation). */
typedef struct bhx_galaxy_params {
    double total_mass;     /* both galaxies                              (1)   */
    double mass_ratio;     /* primary : secondary                        (3)   */
    double disk_fraction;  /* of each galaxy's mass and bodies           (0.2) */
    double disk_scale;     /* R_d of the primary                         (1)   */
    double disk_height;    /* z_d of the primary                         (0.1) */
    double disk_cut;       /* disk truncation radius / R_d               (10)  */
    double halo_scale;     /* Hernquist a of the primary                 (3)   */
    double halo_cut;       /* halo truncation radius / a                 (10)  */
    double disk_sigma;     /* disk velocity dispersion / v_c             (0.1) */
    double separation;     /* initial distance of the centres            (30)  */
    double pericentre;     /* pericentre of the parabolic orbit          (5)   */
    double incl1, incl2;   /* disk inclinations to the orbital plane  (0, pi/3) */
} bhx_galaxy_params;

int bhx_galaxy_defaults(bhx_galaxy_params *p);

/* n bodies: the primary's first (halo then disk), then the secondary's;
 * counter-based draws (body i depends only on seed and i), on up to
 * `threads` host threads (0: all).  Centre-of-mass frame. */
int bhx_ic_galaxy_pair(size_t n, uint64_t seed, const bhx_galaxy_params *p, int threads,
                       double *pos, double *vel, double *mass);

/* ---- snapshots (no reference counterpart) --------------------------------
 * File layout, little-endian: a 256-byte header — magic "BHXSNAP1", uint32
 * header size (256), uint32 version (1), uint64 n, uint64 nsys, double time,
 * uint64 step, uint32 flags (bit 0: accelerations stored, bit 1:
 * compensation terms stored), uint32 0, then the configuration (theta, eps,
 * dt as doubles; leaf_capacity, mode, precision, quadrupole, permute_acc,
 * compensated, device, group_walk, geometry, group_size as int32), zero padding — followed by
 * pos[3n], vel[3n], mass[n] (doubles, xyz interleaved), acc[3n] if stored,
 * the position then velocity compensation terms [6n] if stored, and
 * offsets[nsys + 1] (uint64). */
typedef struct bhx_snapshot_info {
    uint64_t n;            /* bodies                                            */
    uint64_t nsys;         /* systems (1: one system)                           */
    double time;           /* simulation time, as given to the writer           */
    uint64_t step;         /* frame counter, as given to the writer             */
    int32_t has_acc;       /* accelerations stored                              */
    int32_t has_comp;      /* compensation terms stored                         */
    bhx_config cfg;        /* configuration stored with the bodies              */
} bhx_snapshot_info;

/* Writes host arrays (acc, comp NULL: not stored; comp = 6n doubles, the
 * position terms then the velocity terms; offsets NULL: one system, the
 * info's nsys is then taken as 1). */
int bhx_snapshot_write(const char *path, const bhx_snapshot_info *info, const double *pos, const double *vel,
                       const double *mass, const double *acc, const double *comp, const uint64_t *offsets);
int bhx_snapshot_read_info(const char *path, bhx_snapshot_info *info);
/* Arrays sized from the header; any pointer may be NULL (skipped); acc and
 * comp are left untouched when the file holds none. */
int bhx_snapshot_read(const char *path, double *pos, double *vel, double *mass, double *acc, double *comp,
                      uint64_t *offsets);

/* The context's frame state — bodies, accelerations, compensation terms and
 * systems — with its configuration to a file; and back: the state is
 * restored (a frame stepped after bhx_load equals the frame the saving
 * context would have stepped, bit for bit, under the same configuration),
 * the context's configuration is left as it is (the stored one is in the
 * header).  time / step may be NULL on load. */
int bhx_save(bhx_ctx *ctx, const char *path, double time, uint64_t step);
int bhx_load(bhx_ctx *ctx, const char *path, double *time, uint64_t *step);

/* ---- visualisation export (VisualizeBodies, Body.hpp:326-336, and
 * visualizeTree / visualizeNode, HashedOctree.cpp:939-974, as offline files)
 * Legacy VTK, binary (big-endian as the format requires).  Bodies: a point
 * per body with point data "mass" and "speed".  Nodes: an axis-aligned box
 * per node, centred on the node centre with edge = node size (the box the
 * reference draws with ofDrawBox(center, size)), cell data "level", "count"
 * and "mass"; only levels min_level..max_level. */
int bhx_vtk_write_bodies(const char *path, const double *pos, const double *vel, const double *mass, size_t n);
int bhx_vtk_write_nodes(const char *path, const bhx_node *nodes, size_t count, int min_level, int max_level);
/* Both files from the context: its bodies, and the nodes of its last build
 * (either path may be NULL). */
int bhx_export_vtk(bhx_ctx *ctx, const char *bodies_path, const char *nodes_path, int min_level, int max_level);

#ifdef __cplusplus
}
#endif
#endif
