// Drives the Barnes-Hut step through the drop-in C++ API the way the
// reference's ofApp::update() / ofApp::draw() do (ofApp.cpp:169-177,
// 248-282), then writes the final bodies, accelerations and a few tree
// lookups as raw doubles for the test to compare with the oracle.
//
// usage: dropin_frame N FRAMES SEED THETA DT OUTFILE [perbody]
// perbody: the forces come from the loop of ComputeHOTOctreeForce
// (HashedOctree.hpp:266-271) written out, TraverseHOTInteractionList then
// ComputeHOTForceInteractionList for each body.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bhx_dropin.hpp"

int main(int argc, char** argv) {
    if (argc != 7 && argc != 8) {
        std::fprintf(stderr, "usage: %s N FRAMES SEED THETA DT OUTFILE [perbody]\n", argv[0]);
        return 2;
    }
    const bool perbody = argc == 8 && std::strcmp(argv[7], "perbody") == 0;
    const size_t n = std::strtoull(argv[1], nullptr, 10);
    const int frames = std::atoi(argv[2]);
    const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
    const double theta = std::atof(argv[4]);
    const double dt = std::atof(argv[5]);

    std::vector<double> pos(3 * n), vel(3 * n), mass(n);
    if (bhx_ic_plummer(n, seed, 1.0, 1.0, pos.data(), vel.data(), mass.data()) != BHX_OK) return 1;
    Body* bodies = new Body[n];
    for (size_t i = 0; i < n; ++i)
        bodies[i] = Body(Vec3D(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                         Vec3D(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]), mass[i]);
    Vec3D* acc = new Vec3D[n];          // zero, as `new Vec3D[N]` in ofApp::setup
    LinearHashedOctree tree;
    bhx_config cfg;
    bhx_config_default(&cfg);
    cfg.eps = 0.01;
    cfg.theta = theta;
    bhx::dropin::configure(cfg);

    for (int f = 0; f < frames; ++f) {
        // ofApp::update()
        OctantBounds rootNodeBounds(bodies, n);
        ComputeMortonKeys(bodies, n, rootNodeBounds.size);
        RadixSortBodies(bodies, n);
        // ofApp::draw()
        ComputePositionAtHalfTimeStep(dt, bodies, n, acc);
        buildLinearHashedOctreeInPlace(tree, bodies, n, rootNodeBounds);
        HOTNode** walkList = new HOTNode*[n];
        HOTNode** interactList = new HOTNode*[n];
        ResetVec3D(acc, (int)n);
        if (perbody) {
            for (size_t i = 0; i < n; ++i) {
                const long len = TraverseHOTInteractionList(tree, bodies[i].position, theta, walkList, interactList);
                ComputeHOTForceInteractionList(bodies[i].position, bodies[i].mass, acc[i], interactList, len);
            }
        } else {
            ComputeHOTOctreeForce(tree, bodies, acc, n, walkList, interactList, theta);
        }
        ComputeVelocityAndPosition(dt, bodies, n, acc);
        delete[] walkList;
        delete[] interactList;
        if (f + 1 < frames) tree.deleteTree();
    }
    // tree checks: the root and one of its children resolve, a missing key does not
    const HOTNode* root = tree.lookUpNode(ROOT_KEY);
    spatialKey child = 0;
    if (root)
        for (int o = 0; o < 8; ++o)
            if (root->childByte & (1 << o)) {
                child = (ROOT_KEY << 3) | (spatialKey)o;
                break;
            }
    const HOTNode* ch = child ? tree.lookUpNode(child) : nullptr;
    const HOTNode* missing = tree.lookUpNode((spatialKey)2 << 60);
    const double ch_key = ch ? (double)ch->nodeKey : -1.0;   // before the edits below
    // the MAC of the root's first child seen from the origin, on the device
    HOTNode* chn = child ? tree.lookUpNode(child) : nullptr;
    const double mac_child = chn ? (tree.MAC(chn, Vec3D(0, 0, 0), theta) ? 1.0 : 0.0) : -1.0;
    // decodeMortonKey of the survey's known key: (0, 0, 0) for size 1
    const Vec3D dec = decodeMortonKey(0xF000000000000000ull, 1.0);
    // table edits: pruning removes nothing from a built tree; deleting a node
    // makes its lookups miss
    const size_t before = tree.size();
    PruneEmptyNodesFromTree(tree);
    const size_t after_prune = tree.size();
    if (child) tree.deleteNode(child);
    const HOTNode* gone = child ? tree.lookUpNode(child) : nullptr;
    const size_t after_delete = tree.size();
    // errors come back as the reference reports them
    bool threw = false;
    try {
        RadixSortBodies(nullptr, 4);
    } catch (const std::invalid_argument&) {
        threw = true;
    }

    FILE* out = std::fopen(argv[6], "wb");
    if (!out) return 1;
    for (size_t i = 0; i < n; ++i) {
        const double rec[10] = {bodies[i].position.x, bodies[i].position.y, bodies[i].position.z,
                                bodies[i].velocity.x, bodies[i].velocity.y, bodies[i].velocity.z,
                                bodies[i].mass, acc[i].x, acc[i].y, acc[i].z};
        std::fwrite(rec, sizeof rec, 1, out);
    }
    const double tail[14] = {root ? (double)root->N : -1.0, root ? root->mass : -1.0,
                             ch_key, missing ? 1.0 : 0.0, threw ? 1.0 : 0.0,
                             (double)before, mac_child, dec.x, dec.y, dec.z, (double)after_prune,
                             gone ? 1.0 : 0.0, (double)after_delete,
                             chn ? chn->baryCenter.x : 0.0};
    std::fwrite(tail, sizeof tail, 1, out);
    std::fclose(out);
    delete[] bodies;
    delete[] acc;
    std::printf("dropin_frame ok n=%zu frames=%d nodes=%zu perbody=%d\n", n, frames, before, (int)perbody);
    return 0;
}
