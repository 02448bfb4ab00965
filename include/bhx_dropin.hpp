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

// LinearHashedOctree (HashedOctree.hpp:77-135): the tree lives on the GPU;
// lookUpNode answers from the device hash table and returns a HOTNode record
// valid until the next build (HashedOctree.cpp:111-115: nullptr if absent).
class LinearHashedOctree {
public:
    HOTNode* lookUpNode(const spatialKey code) {
        load();
        auto it = index_.find(code);
        return it == index_.end() ? nullptr : &cache_[it->second];
    }
    size_t size() {
        load();
        return cache_.size();
    }
    // ids of keys from the device hash table (-1 for absent keys)
    std::vector<int64_t> lookUp(const std::vector<spatialKey>& keys) {
        std::vector<int64_t> ids(keys.size());
        if (!keys.empty())
            bhx::dropin::Engine::get().check(
                bhx_lookup(bhx::dropin::Engine::get().ctx(1), keys.data(), keys.size(), ids.data()), "bhx_lookup");
        return ids;
    }
    void clear() {
        cache_.clear();
        index_.clear();
        built_ = false;
        loaded_ = false;
    }
    void deleteTree() { clear(); }
    void mark_built(double size) {
        built_ = true;
        loaded_ = false;
        size_ = size;
    }

private:
    void load() {
        if (!built_ || loaded_) return;
        bhx_ctx* c = bhx::dropin::Engine::get().ctx(1);
        size_t nn = 0;
        bhx::dropin::Engine::get().check(bhx_get_nodes(c, nullptr, 0, &nn), "bhx_get_nodes");
        std::vector<bhx_node> raw(nn);
        bhx::dropin::Engine::get().check(bhx_get_nodes(c, raw.data(), nn, &nn), "bhx_get_nodes");
        cache_.resize(nn);
        index_.clear();
        for (size_t i = 0; i < nn; ++i) {
            HOTNode& h = cache_[i];
            for (int q = 0; q < 6; ++q) h.quadrupoleMoment[q] = raw[i].quad[q];
            h.baryCenter = Vec3D(raw[i].com[0], raw[i].com[1], raw[i].com[2]);
            h.mass = raw[i].mass;
            h.N = (long)raw[i].n;
            h.childByte = (uint8_t)raw[i].child_byte;
            h.nodeBounds = OctantBounds(Vec3D(raw[i].center[0], raw[i].center[1], raw[i].center[2]), raw[i].size);
            h.nodeKey = raw[i].key;
            index_[raw[i].key] = i;
        }
        loaded_ = true;
    }
    std::vector<HOTNode> cache_;
    std::unordered_map<spatialKey, size_t> index_;
    bool built_ = false, loaded_ = false;
    double size_ = 0.0;
};

// buildLinearHashedOctreeInPlace (HashedOctree.hpp:211-244).
inline void buildLinearHashedOctreeInPlace(LinearHashedOctree& tree, Body*& bodies, const size_t& numBodies,
                                           OctantBounds& bounds) {
    if (!bodies && numBodies) throw std::invalid_argument("bodies is a null pointer");
    bhx_ctx* c = bhx::dropin::upload(bodies, numBodies);
    const double ctr[3] = {bounds.center.x, bounds.center.y, bounds.center.z};
    size_t nn = 0;
    bhx::dropin::Engine::get().check(bhx_build(c, ctr, bounds.size, &nn), "bhx_build");
    tree.clear();
    tree.mark_built(bounds.size);
}

// ComputeHOTOctreeForce (HashedOctree.hpp:262-272): accelerations of all
// bodies from the tree built on them.  The walk and interaction lists live in
// LDS; walkList and interactList are accepted and not used.
inline void ComputeHOTOctreeForce(LinearHashedOctree& tree, Body*& bodies, Vec3D*& acc, const size_t& numBodies,
                                  HOTNode**& walkList, HOTNode**& interactList, double thetaMAC) {
    (void)tree;
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
    e.check(bhx_walk(c, &acc[0].x, nullptr), "bhx_walk");
}
