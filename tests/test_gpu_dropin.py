"""The drop-in C++ API (include/bhx_dropin.hpp) driven the way the reference's
ofApp drives its own functions, checked against the oracle's frames.

The reference's RadixSortBodies does not reorder the caller's acceleration
array (ofApp.cpp:176,248), so the oracle runs with permute_acc off.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import load_bhx
from tests.kats import MAC_KATS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "dropin_frame")


def test_dropin_driver_built():
    assert os.path.exists(DRIVER), "build() compiles tests/cpp/dropin_frame"


@pytest.mark.gpu
def test_dropin_frames_match_oracle(tmp_path):
    n, frames, seed, theta, dt = 2048, 3, 31, 0.5, 1e-3
    out = tmp_path / "frames.bin"
    print("dropin_frame n=%d frames=%d" % (n, frames), flush=True)
    r = subprocess.run([DRIVER, str(n), str(frames), str(seed), str(theta), str(dt), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    body = raw[:10 * n].reshape(n, 10)
    tail = raw[10 * n:]
    bhx = load_bhx()
    pos, vel, mass = bhx.plummer(n, seed=seed)
    p, v, m, a = pos, vel, mass, np.zeros_like(pos)
    for _ in range(frames):
        p, v, m, a, keys, st = orc.frame(p, v, m, a, dt, theta, 0.01, permute_acc=False)
    assert np.array_equal(body[:, 6], m)
    assert np.allclose(body[:, 0:3], p, rtol=0, atol=1e-13)
    assert np.allclose(body[:, 3:6], v, rtol=0, atol=1e-12)
    assert np.allclose(body[:, 7:10], a, rtol=1e-10, atol=1e-12)
    root_n, root_mass, child_key, missing, threw, nodes = tail[:6]
    assert root_n == n and root_mass == pytest.approx(1.0, rel=1e-14)
    assert child_key > 1 and missing == 0.0 and threw == 1.0
    assert nodes == st.nodes
    _, dx, dy, dz, after_prune, gone, after_delete, _ = tail[6:]
    assert (dx, dy, dz) == (0.0, 0.0, 0.0)
    assert after_prune == nodes and gone == 0.0 and after_delete == nodes - 1


@pytest.mark.gpu
def test_dropin_per_body_frames_bitexact(tmp_path):
    """The driver with ComputeHOTOctreeForce's loop written out (TraverseHOTInteractionList,
    then ComputeHOTForceInteractionList, per body): the lists come in the reference's
    order and sum with its FP64 operation sequence, so three whole frames equal the
    oracle's bit for bit."""
    n, frames, seed, theta, dt = 1024, 3, 33, 0.5, 1e-3
    out = tmp_path / "frames.bin"
    print("dropin_frame perbody n=%d frames=%d" % (n, frames), flush=True)
    r = subprocess.run([DRIVER, str(n), str(frames), str(seed), str(theta), str(dt), str(out), "perbody"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    body = raw[:10 * n].reshape(n, 10)
    tail = raw[10 * n:]
    bhx = load_bhx()
    pos, vel, mass = bhx.plummer(n, seed=seed)
    p, v, m, a = pos, vel, mass, np.zeros_like(pos)
    for f in range(frames):
        c_last, s_last = orc.bounds(p)
        p, v, m, a, keys, st = orc.frame(p, v, m, a, dt, theta, 0.01, permute_acc=False)
    assert np.array_equal(body[:, 7:10], a)
    assert np.array_equal(body[:, 0:3], p) and np.array_equal(body[:, 3:6], v) and np.array_equal(body[:, 6], m)
    # the device MAC of the root's first child, seen from the origin
    t = orc.build_insert(p, m, c_last, s_last)
    nd = t.nodes()
    child = int(tail[2])
    rec = nd[nd["key"] == child][0]
    assert tail[6] == float(orc.mac(rec["size"], rec["com"], np.zeros(3), theta)) and tail[13] == rec["com"][0]


def _as_keys(entries, nodes, tperm=None):
    """List entries as comparable items: ('n', key) for nodes, ('b', body) for bodies."""
    out = []
    for e in entries:
        e = int(e)
        out.append(("n", int(nodes["key"][e])) if e >= 0 else ("b", -1 - e))
    return out


def _oracle_items(t, nodes_o, point, theta):
    ids, kind = t.interaction_list(point, theta)
    return [("b", int(i)) if k else ("n", int(nodes_o["key"][i])) for i, k in zip(ids, kind)]


def _runs_sorted(items):
    """Sort each run of consecutive body entries (a leaf's bodies: tree order on the
    GPU, array order in the oracle)."""
    out, run = [], []
    for it in items:
        if it[0] == "b":
            run.append(it)
        else:
            out += sorted(run) + [it]
            run = []
    return out + sorted(run)


@pytest.mark.gpu
@pytest.mark.parametrize("cap,mode,quad", [(1, 0, 0), (4, 1, 1)])
def test_interaction_lists_and_forces_match_reference_order(cap, mode, quad):
    """TraverseHOTInteractionList on the device, for every body position and for
    points off the bodies: the same entries in the same order as the oracle's
    explicit-stack search; ComputeHOTForceInteractionList over those lists equals the
    oracle's walk (same terms, same order) -- bit for bit for the as-run payload, to
    the rounding of the intended moments otherwise."""
    bhx = load_bhx()
    # 2048 equal masses 2^-11: every node mass sum is exact, so the as-run payload
    # (which the engine sums in tree order, the reference in insertion order) is too
    n, theta, eps = 2048, 0.5, 0.025
    pos, vel, mass = bhx.plummer(n, seed=5)
    c, s = orc.bounds(pos)
    with bhx.Engine(n, theta=theta, eps=eps, leaf_capacity=cap, mode=mode, quadrupole=quad) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        g = e.nodes()
        pts = np.concatenate([pos, np.random.default_rng(2).normal(size=(50, 3))])
        off, ent = e.interaction_lists(pts, theta)
        acc = e.force_lists(pts, off, ent)
        acc0 = np.random.default_rng(3).normal(size=(pts.shape[0], 3)) * 1e-3
        acc_plus = e.force_lists(pts, off, ent, acc=acc0)       # adds to what acc holds
    t = orc.build_insert(pos, mass, c, s) if cap == 1 else orc.build_recursive(pos, mass, c, s, cap)
    if mode == 1:
        t.moments_intended()
    o = t.nodes()
    for q in range(0, pts.shape[0], 7):
        got = _as_keys(ent[off[q]:off[q + 1]], g)
        want = _oracle_items(t, o, pts[q], theta)
        if cap > 1:
            got, want = _runs_sorted(got), _runs_sorted(want)
        assert got == want, q
    oacc, ocnt = t.walk(theta, eps, quad=bool(quad))
    assert np.array_equal(np.diff(off)[:n].astype(np.int64), ocnt)
    # starting from acc0 changes the rounding of every partial sum: ~1e-16 of the
    # running magnitude per term, over lists of a few hundred terms
    assert np.abs(acc_plus - (acc0 + acc)).max() < 1e-12 * np.abs(acc).max()
    if mode == 0:
        assert np.array_equal(acc[:n], oacc)          # same terms in the same order
    else:
        err = np.abs(acc[:n] - oacc).max() / np.abs(oacc).max()
        print("intended, quadrupole: max relative difference %.2e" % err, flush=True)
        assert err < 1e-12                             # the moments' own rounding


@pytest.mark.gpu
def test_mac_known_answers_on_device():
    """LinearHashedOctree::MAC at the two theta = 0.6 boundary cases where the
    multiplied form decides the other way: a root of the KAT edge whose barycentre
    (as run: the first body) sits at the KAT position."""
    bhx = load_bhx()
    for com, size_hex, want in MAC_KATS:
        size = float.fromhex(size_hex)
        pos = np.array([com, [-0.1, 0.05, -0.2], [0.3, -0.3, 0.1]], dtype=np.float64)
        with bhx.Engine(8) as e:
            e.set_bodies(pos, None, np.ones(3))
            e.build(np.zeros(3), size)
            g = e.nodes()
            assert g["size"][0] == size and np.array_equal(g["com"][0], np.array(com))
            got = e.mac([0], np.zeros((1, 3)), 0.6)
        assert bool(got[0]) is want


@pytest.mark.gpu
def test_node_table_edits():
    """insertHOTNode / deleteNode / pruneEmptyNodes on the device table: a put node is
    found through the hash table and joins the traversal; the walks refuse until the
    next build; pruning removes the node-less childless node; a deleted node's lookups
    miss and traversals skip it; a rebuild restores everything."""
    bhx = load_bhx()
    n = 300
    pos, vel, mass = bhx.plummer(n, seed=8)
    c, s = orc.bounds(pos)
    with bhx.Engine(n, eps=0.025) as e:
        e.set_bodies(pos, vel, mass)
        nn = e.build(c, s)
        g = e.nodes()
        leaf = int(np.nonzero(g["child_byte"] == 0)[0][5])
        lk, ll = int(g["key"][leaf]), int(g["level"][leaf])
        # make the leaf a parent of an empty node in octant 3
        rec = g[leaf].copy()
        rec["child_byte"] = 1 << 3
        assert e.node_put(rec) == leaf
        new = np.zeros((), dtype=bhx.NODE_DTYPE)
        new["key"], new["level"], new["n"] = (lk << 3) | 3, ll + 1, 0
        new["com"] = [9.0, 9.0, 9.0]
        nid = e.node_put(new)
        assert nid == nn and e.lookup([(lk << 3) | 3])[0] == nid
        g2 = e.nodes()
        assert g2["key"][nid] == (lk << 3) | 3 and g2["size"][nid] == g["size"][leaf] * 0.5
        with pytest.raises(bhx.BhxError):
            e.walk()
        far = np.array([[50.0, 50.0, 50.0]])
        off, ent = e.interaction_lists(far, 0.0)      # theta 0: every node opened, every leaf listed
        assert nid in set(ent.tolist()) and leaf not in set(ent.tolist())
        assert e.prune_empty(1) == 1                    # PruneEmptyNodesFromTree
        assert e.lookup([(lk << 3) | 3])[0] == -1
        off, ent = e.interaction_lists(far, 0.0)
        assert nid not in set(ent.tolist()) and leaf not in set(ent.tolist())
        assert e.node_erase(int(g["key"][leaf]))
        assert not e.node_erase(int(g["key"][leaf]))
        assert e.lookup([int(g["key"][leaf])])[0] == -1 and e.nodes()["key"][leaf] == 0
        e.build(c, s)
        assert np.array_equal(e.nodes(), g)
        e.walk()


@pytest.mark.gpu
def test_decode_morton_keys_matches_oracle():
    """decodeMortonKey on the device: the survey's key KATs decode back to their
    points, and random keys at a non-dyadic size equal the oracle's bits."""
    bhx = load_bhx()
    rng = np.random.default_rng(9)
    keys = np.concatenate([np.array([0xF000000000000000, 0x8000000000000000, 0xDC00000000000000], np.uint64),
                           rng.integers(0, 2 ** 63, 1000, dtype=np.int64).astype(np.uint64) | np.uint64(1 << 63)])
    with bhx.Engine(8) as e:
        got1 = e.decode_morton_keys(keys[:3], 1.0)
        got = e.decode_morton_keys(keys, 1.3)
    assert np.array_equal(got1, [[0, 0, 0], [-1, -1, -1], [0.5, -0.5, 0]])
    ref = np.array([orc.decode_morton_key(k, 1.3) for k in keys])
    assert np.array_equal(got, ref)
