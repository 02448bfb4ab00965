// bhx_dropin.hpp — the reference's C++ API for the Barnes-Hut step, backed by
// the MI355X engine (libbhx.so) through the C ABI in bhx.h.
//
// Same class and function names, signatures and argument meaning as the
// reference (DavidRichardson02/Serial-Hashed-Octree-Barnes-Hut-N-Body-
// Simulator-0_3, src/), so ofApp-style code switches by including this header
// instead of Body.hpp / HashedOctree.hpp / MortonKeys.hpp / OctantUtils.hpp
// and linking libbhx.so.  Every computation runs on the GPU; the host code
// here only moves the caller's arrays to and from the device (the reference's
// API hands over host arrays) and rethrows status codes the way the
// reference reports errors (std::invalid_argument for null inputs).
//
// All calls share one engine context per process (created on first use with
// the reference's defaults: SOFTENING 0.025, leaf capacity 1, as-run node
// payload, reference per-body walk in FP64); `bhx::dropin::configure` changes
// it.  Not thread-safe, like the reference's LinearHashedOctree.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "bhx.h"

// ------------------------------------------------------------------ types

typedef uint64_t spatialKey;                                  // MortonKeys.hpp:41
static const spatialKey ROOT_KEY = 1;                         // OctantUtils.hpp:20

struct ofColorRGBA {                                          // stands in for ofColor (4 x u8)
    unsigned char r = 0, g = 255, b = 0, a = 255;
};

// Vec3D (MathUtils.hpp:23-73): three FP64 components; the operators used on
// the path.  Arrays of Vec3D are interleaved xyz doubles.
class Vec3D {
public:
    Vec3D() : x(0), y(0), z(0) {}
    Vec3D(double _x, double _y, double _z) : x(_x), y(_y), z(_z) {}
    bool operator!=(const Vec3D& o) const { return x != o.x || y != o.y || z != o.z; }
    Vec3D operator+(const Vec3D& o) const { return Vec3D(x + o.x, y + o.y, z + o.z); }
    Vec3D operator-(const Vec3D& o) const { return Vec3D(x - o.x, y - o.y, z - o.z); }
    Vec3D operator*(double s) const { return Vec3D(x * s, y * s, z * s); }
    void operator+=(const Vec3D& o) { x = x + o.x; y = y + o.y; z = z + o.z; }
    void reset() { x = y = z = 0; }
    double x, y, z;
};
static_assert(sizeof(Vec3D) == 3 * sizeof(double), "Vec3D is three packed doubles");

// ResetVec3D (MathUtils.hpp:81-93): zero an array of vectors; null throws.
static inline void ResetVec3D(Vec3D*& vec3d, int numVecs) {
    if (!vec3d) throw std::invalid_argument("vec3d is a null pointer in 'ResetVec3D'.");
    for (int i = 0; i < numVecs; ++i) vec3d[i].reset();
}

// Body (Body.hpp:30-65).
class Body {
public:
    Body() : mass(0.0), bodyKey(0), displayRadius(0.0) {}
    Body(Vec3D p, Vec3D v, double m) : position(p), velocity(v), mass(m), bodyKey(0), displayRadius(m) {}
    Body(Vec3D p, Vec3D v, double m, ofColorRGBA c) : position(p), velocity(v), mass(m), bodyKey(0), bodyColor(c),
                                                      displayRadius(m) {}
    void setBodyKey(spatialKey k) { bodyKey = k; }
    spatialKey getBodyKey() const { return bodyKey; }
    Vec3D position;
    Vec3D velocity;
    double mass;
    spatialKey bodyKey;
    ofColorRGBA bodyColor;
    double displayRadius;
};

// OctantBounds (OctantUtils.hpp:112-123).
class OctantBounds;
namespace bhx { namespace dropin { inline void bounds_of(Body* b, size_t n, OctantBounds& out); } }

class OctantBounds {
public:
    OctantBounds() : center(0, 0, 0), size(0.0) {}
    OctantBounds(Vec3D c, double s) : center(c), size(s) {}
    // Centre at the sequential mean of the positions, size max(0, max |p| - c)
    // (OctantUtils.cpp:27-59), computed on the GPU.
    OctantBounds(Body* bodies, size_t numBodies) : center(0, 0, 0), size(0.0) {
        bhx::dropin::bounds_of(bodies, numBodies, *this);
    }
    Vec3D center;
    double size;
};

// HOTNode (HashedNode.hpp:60-95).
class HOTNode {
public:
    double quadrupoleMoment[6] = {0, 0, 0, 0, 0, 0};
    Vec3D baryCenter;
    double mass = 0.0;
    long N = 0;
    uint8_t childByte = 0;
    OctantBounds nodeBounds;
    spatialKey nodeKey = 0;
};

namespace bhx {
namespace dropin {

class Engine {
public:
    static Engine& get() {
        static Engine e;
        return e;
    }
    bhx_ctx* ctx(size_t need) {
        if (!ctx_ || need > cap_) {
            if (ctx_) bhx_destroy(ctx_);
            ctx_ = nullptr;
            cap_ = need < 1024 ? 1024 : need;
            check(bhx_create(&cfg_, cap_, &ctx_), "bhx_create");
        }
        return ctx_;
    }
    bhx_config& config() { return cfg_; }
    void apply() {
        if (ctx_) check(bhx_set_config(ctx_, &cfg_), "bhx_set_config");
    }
    void check(int rc, const char* what) const {
        if (rc == BHX_OK) return;
        std::string msg = std::string(what) + ": " + (ctx_ ? bhx_last_error(ctx_) : "") + " (status " +
                          std::to_string(rc) + ")";
        if (rc == BHX_E_NULL) throw std::invalid_argument(msg);
        throw std::runtime_error(msg);
    }
    ~Engine() {
        if (ctx_) bhx_destroy(ctx_);
    }
    // positions (xyz) and masses of the bodies last handed to the engine
    std::vector<double> last_pos, last_mass;

private:
    Engine() {
        bhx_config_default(&cfg_);
        cfg_.eps = 0.025;        // SOFTENING, HashedOctree.hpp:151
        cfg_.leaf_capacity = 1;
        cfg_.mode = BHX_MODE_AS_RUN;
        cfg_.precision = BHX_FP64;
        cfg_.group_walk = 0;
        cfg_.permute_acc = 1;
    }
    bhx_config cfg_{};
    bhx_ctx* ctx_ = nullptr;
    size_t cap_ = 0;
};

// Change the engine settings used by every drop-in call (softening, leaf
// capacity, payload mode, precision, walk kind ...).
inline void configure(const bhx_config& cfg) {
    Engine& e = Engine::get();
    e.config() = cfg;
    e.apply();
}

// Caller's Body[] -> device (positions, velocities, masses).
inline bhx_ctx* upload(Body* b, size_t n) {
    Engine& e = Engine::get();
    bhx_ctx* c = e.ctx(n);
    std::vector<double> pos(3 * n), vel(3 * n), mass(n);
    for (size_t i = 0; i < n; ++i) {
        pos[3 * i] = b[i].position.x; pos[3 * i + 1] = b[i].position.y; pos[3 * i + 2] = b[i].position.z;
        vel[3 * i] = b[i].velocity.x; vel[3 * i + 1] = b[i].velocity.y; vel[3 * i + 2] = b[i].velocity.z;
        mass[i] = b[i].mass;
    }
    e.check(bhx_set_bodies(c, pos.data(), vel.data(), mass.data(), n), "bhx_set_bodies");
    e.last_pos.swap(pos);
    e.last_mass.swap(mass);
    return c;
}

// Device -> caller's Body[] (positions and velocities).
inline void download(bhx_ctx* c, Body* b, size_t n) {
    std::vector<double> pos(3 * n), vel(3 * n);
    Engine::get().check(bhx_get_bodies(c, pos.data(), vel.data(), nullptr, nullptr), "bhx_get_bodies");
    for (size_t i = 0; i < n; ++i) {
        b[i].position = Vec3D(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        b[i].velocity = Vec3D(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    }
}

inline void bounds_of(Body* b, size_t n, OctantBounds& out) {
    if (!b && n) throw std::invalid_argument("bodies is a null pointer");
    bhx_ctx* c = upload(b, n);
    double ctr[3], s;
    Engine::get().check(bhx_bounds(c, ctr, &s), "bhx_bounds");
    out.center = Vec3D(ctr[0], ctr[1], ctr[2]);
    out.size = s;
}

}  // namespace dropin
}  // namespace bhx

// ------------------------------------------------------------------ MortonKeys

// computeMortonKey over a whole Body array (the loop of ofApp.cpp:172-175 in
// one call): bodyKey = key of position in the origin-centred domain +-size.
inline void ComputeMortonKeys(Body* bodies, size_t numBodies, double size) {
    if (!bodies && numBodies) throw std::invalid_argument("bodies is a null pointer");
    bhx_ctx* c = bhx::dropin::upload(bodies, numBodies);
    std::vector<uint64_t> keys(numBodies);
    bhx::dropin::Engine::get().check(bhx_morton_keys(c, size, keys.data()), "bhx_morton_keys");
    for (size_t i = 0; i < numBodies; ++i) bodies[i].bodyKey = keys[i];
}

// computeMortonKey (MortonKeys.hpp:137-181) for one position.  One device
// round trip per call: use ComputeMortonKeys for arrays.
inline spatialKey computeMortonKey(const Vec3D& p, const double size) {
    Body b(p, Vec3D(), 0.0);
    ComputeMortonKeys(&b, 1, size);
    return b.bodyKey;
}

// ThreePassRadixSortMortonKeys (MortonKeys.hpp:439-470): stable ascending
// sort of keys, carrying indexes along.
inline void ThreePassRadixSortMortonKeys(spatialKey* keys, size_t* indexes, size_t numKeys) {
    if (!keys || !indexes) throw std::invalid_argument("keys is a null pointer");
    if (numKeys <= 1) return;
    static_assert(sizeof(size_t) == sizeof(uint64_t), "size_t indexes are 64-bit");
    bhx_ctx* c = bhx::dropin::Engine::get().ctx(numKeys);
    bhx::dropin::Engine::get().check(
        bhx_sort_keys(c, reinterpret_cast<uint64_t*>(keys), reinterpret_cast<uint64_t*>(indexes), numKeys),
        "bhx_sort_keys");
}

// ------------------------------------------------------------------ Body

// ReorderBodiesByIndexes (Body.hpp:257-273): new[i] = old[indexes[i]].
inline void ReorderBodiesByIndexes(Body* bodies, size_t* indexes, size_t numBodies) {
    if (!bodies || !indexes) throw std::invalid_argument("bodies is a null pointer");
    std::vector<Body> tmp(bodies, bodies + numBodies);
    for (size_t i = 0; i < numBodies; ++i) bodies[i] = tmp[indexes[i]];
}

// RadixSortBodies (Body.hpp:275-302): sort the bodies by their bodyKey.
inline void RadixSortBodies(Body* bodies, size_t numBodies) {
    if (!bodies) throw std::invalid_argument("bodies is a null pointer");
    std::vector<spatialKey> keys(numBodies);
    std::vector<size_t> idx(numBodies);
    for (size_t i = 0; i < numBodies; ++i) {
        keys[i] = bodies[i].bodyKey;
        idx[i] = i;
    }
    ThreePassRadixSortMortonKeys(keys.data(), idx.data(), numBodies);
    ReorderBodiesByIndexes(bodies, idx.data(), numBodies);
}

// ComputePositionAtHalfTimeStep (Body.hpp:306-313): v += (a dt) 0.5; x += v dt.
inline void ComputePositionAtHalfTimeStep(double dt, Body*& bodies, size_t numBodies, Vec3D*& acc) {
    if (!bodies || !acc) throw std::invalid_argument("bodies is a null pointer");
    bhx_ctx* c = bhx::dropin::upload(bodies, numBodies);
    auto& e = bhx::dropin::Engine::get();
    e.check(bhx_set_accelerations(c, &acc[0].x), "bhx_set_accelerations");
    e.check(bhx_kick_drift(c, dt), "bhx_kick_drift");
    bhx::dropin::download(c, bodies, numBodies);
}

// ComputeVelocityAndPosition (Body.hpp:317-324): v += (a dt) 0.5.
inline void ComputeVelocityAndPosition(double dt, Body*& bodies, size_t numBodies, Vec3D*& acc) {
    if (!bodies || !acc) throw std::invalid_argument("bodies is a null pointer");
    bhx_ctx* c = bhx::dropin::upload(bodies, numBodies);
    auto& e = bhx::dropin::Engine::get();
    e.check(bhx_set_accelerations(c, &acc[0].x), "bhx_set_accelerations");
    e.check(bhx_kick(c, dt), "bhx_kick");
    bhx::dropin::download(c, bodies, numBodies);
}

// ------------------------------------------------------------------ octree

// LinearHashedOctree (HashedOctree.hpp:77-135).  The tree lives on the GPU:
// lookups go through the device hash table (bhx_lookup) and return a HOTNode
// record from a host copy of the node table, valid until the next build
// (HashedOctree.cpp:111-115: nullptr if absent).  Edits (insertHOTNode,
// deleteNode, pruneEmptyNodes) change the device table; the records already
// handed out stay where they are.
class LinearHashedOctree {
public:
    // lookUpNode — HashedOctree.cpp:111-115, through the device hash table.
    HOTNode* lookUpNode(const spatialKey code) {
        if (!built_) return nullptr;
        std::vector<int64_t> ids = lookUp(std::vector<spatialKey>{code});
        return ids[0] < 0 ? nullptr : node_at(ids[0]);
    }
    // ids of keys from the device hash table (-1 for absent keys)
    std::vector<int64_t> lookUp(const std::vector<spatialKey>& keys) {
        std::vector<int64_t> ids(keys.size(), -1);
        if (!keys.empty() && built_)
            bhx::dropin::Engine::get().check(
                bhx_lookup(bhx::dropin::Engine::get().ctx(1), keys.data(), keys.size(), ids.data()), "bhx_lookup");
        return ids;
    }
    // nodes of the table (records deleted since the build are not counted)
    size_t size() {
        load();
        size_t k = 0;
        for (const HOTNode& h : cache_) k += h.nodeKey != 0;
        return k;
    }
    // GetNodeTreeDepth — HashedOctree.cpp:64-73: levels below the root.
    size_t GetNodeTreeDepth(const spatialKey nodeKey) const {
        size_t depth = 0;
        for (spatialKey k = nodeKey; k > 1; k >>= 3) ++depth;
        return depth;
    }
    // getParentNode / getChildNode — HashedOctree.cpp:83-100 (the child of
    // octant 0, as the reference computes it).
    HOTNode* getParentNode(const HOTNode* childNode) { return lookUpNode(childNode->nodeKey >> 3); }
    HOTNode* getChildNode(const HOTNode* parentNode) { return lookUpNode(parentNode->nodeKey << 3); }
    // descendantOf — HashedOctree.cpp:219-237.
    bool descendantOf(spatialKey nodeKey, spatialKey parentKey) const {
        if (parentKey == 1) return true;
        if (nodeKey <= parentKey) return false;
        nodeKey >>= (GetNodeTreeDepth(nodeKey) - GetNodeTreeDepth(parentKey)) * 3;
        return nodeKey == parentKey;
    }
    // MAC — HashedOctree.cpp:872-885, evaluated on the device for this node.
    bool MAC(HOTNode* node, const Vec3D& r, double theta) {
        const int64_t id = id_of(node);
        if (id < 0) throw std::invalid_argument("MAC: node is not a node of this tree");
        uint8_t acc = 0;
        bhx::dropin::Engine::get().check(
            bhx_mac(bhx::dropin::Engine::get().ctx(1), &id, &r.x, 1, theta, &acc), "bhx_mac");
        return acc != 0;
    }
    // insertHOTNode — HashedOctree.cpp:152-205: the node replaces the node of
    // its key or is added (the reference deletes the caller's object when the
    // key exists; here the caller keeps it and the table holds a copy).
    void insertHOTNode(HOTNode*& node) {
        if (!node) throw std::invalid_argument("insertHOTNode: node is a null pointer");
        bhx_node r{};
        r.key = node->nodeKey;
        r.center[0] = node->nodeBounds.center.x;
        r.center[1] = node->nodeBounds.center.y;
        r.center[2] = node->nodeBounds.center.z;
        r.size = node->nodeBounds.size;
        r.com[0] = node->baryCenter.x;
        r.com[1] = node->baryCenter.y;
        r.com[2] = node->baryCenter.z;
        r.mass = node->mass;
        for (int q = 0; q < 6; ++q) r.quad[q] = node->quadrupoleMoment[q];
        r.n = node->N;
        r.level = (int32_t)GetNodeTreeDepth(node->nodeKey);
        r.child_byte = node->childByte;
        int64_t id = -1;
        bhx::dropin::Engine::get().check(bhx_node_put(bhx::dropin::Engine::get().ctx(1), &r, &id), "bhx_node_put");
        load();
        if ((size_t)id < cache_.size()) cache_[(size_t)id] = *node;
        else if ((size_t)id == cache_.size() && cache_.size() < cache_.capacity()) cache_.push_back(*node);
        else loaded_ = false;    // reloaded (and moved) at the next access
        edited_ = true;
    }
    // deleteNode — HashedOctree.cpp:647-659.
    void deleteNode(spatialKey nodeCode) {
        int erased = 0;
        bhx::dropin::Engine::get().check(bhx_node_erase(bhx::dropin::Engine::get().ctx(1), nodeCode, &erased),
                                         "bhx_node_erase");
        if (erased) forget(nodeCode);
    }
    // pruneEmptyNodes — HashedOctree.cpp:610-639: erases the nodes at or below
    // `node` that hold no body and have no children.
    void pruneEmptyNodes(HOTNode* node) {
        if (!node) return;
        size_t removed = 0;
        bhx::dropin::Engine::get().check(
            bhx_prune_empty(bhx::dropin::Engine::get().ctx(1), node->nodeKey, &removed), "bhx_prune_empty");
        if (removed) {
            loaded_ = false;
            edited_ = true;
        }
    }
    void clear() {
        cache_.clear();
        body_nodes_.clear();
        built_ = false;
        loaded_ = false;
        edited_ = false;
    }
    void deleteTree() { clear(); }
    void mark_built(double size) {
        clear();
        built_ = true;
        size_ = size;
        current() = this;
    }
    ~LinearHashedOctree() {
        if (current() == this) current() = nullptr;
    }
    // the tree of the last build (its records are the valid list entries)
    static LinearHashedOctree*& current() {
        static LinearHashedOctree* t = nullptr;
        return t;
    }
    bool edited() const { return edited_; }

    // the list entry (bhx_interaction_lists) a HOTNode* stands for, or a
    // negative value below -(1 + bodies) when it is not one of this tree's
    int64_t entry_of(const HOTNode* p) {
        const int64_t id = id_of(p);
        if (id >= 0) return id;
        if (!body_nodes_.empty() && p >= body_nodes_.data() && p < body_nodes_.data() + body_nodes_.size())
            return -1 - (int64_t)(p - body_nodes_.data());
        return INT64_MIN;
    }
    // the HOTNode* of a list entry: a node record, or body i as a one-body node
    HOTNode* node_of_entry(int64_t e) {
        if (e >= 0) return node_at(e);
        const size_t i = (size_t)(-1 - e);
        if (body_nodes_.empty()) {
            const auto& eng = bhx::dropin::Engine::get();
            const size_t n = eng.last_mass.size();
            body_nodes_.resize(n);
            for (size_t k = 0; k < n; ++k) {
                HOTNode& h = body_nodes_[k];
                h.baryCenter = Vec3D(eng.last_pos[3 * k], eng.last_pos[3 * k + 1], eng.last_pos[3 * k + 2]);
                h.mass = eng.last_mass[k];
                h.N = 1;
            }
        }
        if (i >= body_nodes_.size()) throw std::runtime_error("list entry names no body");
        return &body_nodes_[i];
    }

private:
    HOTNode* node_at(int64_t id) {
        load();
        if (id < 0 || (size_t)id >= cache_.size()) return nullptr;
        return &cache_[(size_t)id];
    }
    int64_t id_of(const HOTNode* p) {
        load();
        if (cache_.empty() || p < cache_.data() || p >= cache_.data() + cache_.size()) return -1;
        return (int64_t)(p - cache_.data());
    }
    void forget(spatialKey key) {
        for (HOTNode& h : cache_)
            if (h.nodeKey == key) h.nodeKey = 0;
        edited_ = true;
    }
    void load() {
        if (!built_ || loaded_) return;
        bhx_ctx* c = bhx::dropin::Engine::get().ctx(1);
        size_t nn = 0;
        bhx::dropin::Engine::get().check(bhx_get_nodes(c, nullptr, 0, &nn), "bhx_get_nodes");
        std::vector<bhx_node> raw(nn);
        bhx::dropin::Engine::get().check(bhx_get_nodes(c, raw.data(), nn, &nn), "bhx_get_nodes");
        std::vector<HOTNode> fresh;
        fresh.reserve(2 * nn + 64);    // room for inserted nodes without moving the records
        fresh.resize(nn);
        for (size_t i = 0; i < nn; ++i) {
            HOTNode& h = fresh[i];
            for (int q = 0; q < 6; ++q) h.quadrupoleMoment[q] = raw[i].quad[q];
            h.baryCenter = Vec3D(raw[i].com[0], raw[i].com[1], raw[i].com[2]);
            h.mass = raw[i].mass;
            h.N = (long)raw[i].n;
            h.childByte = (uint8_t)raw[i].child_byte;
            h.nodeBounds = OctantBounds(Vec3D(raw[i].center[0], raw[i].center[1], raw[i].center[2]), raw[i].size);
            h.nodeKey = raw[i].key;    // 0: a deleted record
        }
        cache_.swap(fresh);
        loaded_ = true;
    }
    std::vector<HOTNode> cache_;
    std::vector<HOTNode> body_nodes_;   // bodies of multi-body leaves as list entries
    bool built_ = false, loaded_ = false, edited_ = false;
    double size_ = 0.0;
};

// LookUpNode (HashedOctree.hpp:103) — the free-function form of lookUpNode.
static inline HOTNode* LookUpNode(LinearHashedOctree& HTree, spatialKey code) { return HTree.lookUpNode(code); }

// buildLinearHashedOctreeInPlace (HashedOctree.hpp:211-244).
inline void buildLinearHashedOctreeInPlace(LinearHashedOctree& tree, Body*& bodies, const size_t& numBodies,
                                           OctantBounds& bounds) {
    if (!bodies && numBodies) throw std::invalid_argument("bodies is a null pointer");
    bhx_ctx* c = bhx::dropin::upload(bodies, numBodies);
    const double ctr[3] = {bounds.center.x, bounds.center.y, bounds.center.z};
    size_t nn = 0;
    bhx::dropin::Engine::get().check(bhx_build(c, ctr, bounds.size, &nn), "bhx_build");
    tree.mark_built(bounds.size);
}

// PruneEmptyNodesFromTree (HashedOctree.hpp:248-252).
inline void PruneEmptyNodesFromTree(LinearHashedOctree& HTree) { HTree.pruneEmptyNodes(HTree.lookUpNode(ROOT_KEY)); }

// ComputeHOTOctreeBaryCenters (HashedOctree.hpp:254-257).  The engine's build
// already writes the configured payload (mode AS_RUN: what the reference's
// insertion leaves, where this pass does nothing — its root argument is null
// by then; mode INTENDED: mass, centre of mass and quadrupole), so nothing
// is left to do here.
inline void ComputeHOTOctreeBaryCenters(LinearHashedOctree& HTree, HOTNode* rootNode) {
    (void)HTree;
    (void)rootNode;
}

// TraverseHOTInteractionList (HashedOctree.hpp:336-385): the interaction list
// of one position on the GPU in the reference's order; interactList receives
// HOTNode records of the tree (a body of a multi-body leaf comes as a
// one-body record).  walkList is not used (the stack is device memory).
// Returns the list length.
static inline long TraverseHOTInteractionList(LinearHashedOctree& LHTree, Vec3D& bodyPosition, double theta,
                                              HOTNode**& walkList, HOTNode**& interactList) {
    (void)walkList;
    if (!interactList) throw std::invalid_argument("interactList is a null pointer");
    if (LHTree.size() == 0) return 0;
    auto& e = bhx::dropin::Engine::get();
    uint64_t off[2] = {0, 0};
    size_t total = 0;
    bhx_ctx* c = e.ctx(1);
    e.check(bhx_interaction_lists(c, &bodyPosition.x, 1, theta, off, nullptr, 0, &total), "bhx_interaction_lists");
    std::vector<int64_t> ent(total ? total : 1);
    e.check(bhx_interaction_lists(c, &bodyPosition.x, 1, theta, off, ent.data(), ent.size(), &total),
            "bhx_interaction_lists");
    for (size_t k = 0; k < total; ++k) interactList[k] = LHTree.node_of_entry(ent[k]);
    return (long)total;
}

// ComputeHOTForceInteractionList (HashedOctree.hpp:393-459): acceleration +=
// the softened terms of the list's nodes in list order, on the GPU with the
// reference's FP64 operation sequence.  bodyMass is not used (as in the
// reference).  The entries must be records handed out by the tree of the
// last build (the drop-in shares one engine, hence one tree, per process).
inline void ComputeHOTForceInteractionList(Vec3D& bodyPosition, double& bodyMass, Vec3D& acceleration,
                                           HOTNode**& interactList, long listLength) {
    (void)bodyMass;
    if (listLength <= 0) return;
    if (!interactList) throw std::invalid_argument("interactList is a null pointer");
    LinearHashedOctree* tree = LinearHashedOctree::current();
    if (!tree) throw std::invalid_argument("no tree has been built");
    std::vector<int64_t> ent((size_t)listLength);
    for (long k = 0; k < listLength; ++k) {
        ent[(size_t)k] = tree->entry_of(interactList[k]);
        if (ent[(size_t)k] == INT64_MIN) throw std::invalid_argument("list entry is not a node of the tree");
    }
    const uint64_t off[2] = {0, (uint64_t)listLength};
    auto& e = bhx::dropin::Engine::get();
    e.check(bhx_force_lists(e.ctx(1), &bodyPosition.x, 1, off, ent.data(), &acceleration.x), "bhx_force_lists");
}

// ComputeHOTOctreeForce (HashedOctree.hpp:262-272): accelerations of all
// bodies from the tree built on them, in one batched walk (bhx_walk).  After
// edits to the node table the per-body lists run instead (all bodies in one
// batch), which see the edits.  walkList and interactList are accepted and
// not used.
inline void ComputeHOTOctreeForce(LinearHashedOctree& tree, Body*& bodies, Vec3D*& acc, const size_t& numBodies,
                                  HOTNode**& walkList, HOTNode**& interactList, double thetaMAC) {
    (void)walkList;
    (void)interactList;
    if (!bodies || !acc) throw std::invalid_argument("bodies is a null pointer");
    auto& e = bhx::dropin::Engine::get();
    bhx_ctx* c = e.ctx(numBodies);
    size_t n = 0;
    e.check(bhx_num_bodies(c, &n), "bhx_num_bodies");
    if (n != numBodies) throw std::invalid_argument("tree was built on a different body count");
    if (e.config().theta != thetaMAC) {
        e.config().theta = thetaMAC;
        e.apply();
    }
    if (!tree.edited()) {
        e.check(bhx_walk(c, &acc[0].x, nullptr), "bhx_walk");
        return;
    }
    std::vector<double> pts(3 * n);
    for (size_t i = 0; i < n; ++i) {
        pts[3 * i] = bodies[i].position.x;
        pts[3 * i + 1] = bodies[i].position.y;
        pts[3 * i + 2] = bodies[i].position.z;
    }
    std::vector<uint64_t> off(n + 1);
    size_t total = 0;
    e.check(bhx_interaction_lists(c, pts.data(), n, thetaMAC, off.data(), nullptr, 0, &total), "bhx_interaction_lists");
    std::vector<int64_t> ent(total ? total : 1);
    e.check(bhx_interaction_lists(c, pts.data(), n, thetaMAC, off.data(), ent.data(), ent.size(), &total),
            "bhx_interaction_lists");
    e.check(bhx_force_lists(c, pts.data(), n, off.data(), ent.data(), &acc[0].x), "bhx_force_lists");
}

// decodeMortonKey (MortonKeys.hpp:204-228), on the GPU.
inline Vec3D decodeMortonKey(const spatialKey _mortonKey, const double _size) {
    Vec3D p;
    auto& e = bhx::dropin::Engine::get();
    e.check(bhx_decode_morton_keys(e.ctx(1), &_mortonKey, 1, _size, &p.x), "bhx_decode_morton_keys");
    return p;
}
