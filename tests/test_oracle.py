"""Pins the CPU oracle against known answers, brute force and the mathematics.

These run without a GPU.  Each check is chosen so that a plausible mistake in
the oracle (a wrong shift or mask, a non-stable pass, a wrong octant bit, a
dropped quadrupole term, a sign) fails it.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import (canonical_node_keys, effective21, interleave_bruteforce,
                           plummer_np)
from tests.kats import DRIFT_KAT, KICK_KAT, MAC_KATS, MEAN_KAT_C, MEAN_KAT_POS, TINY


# ---------------------------------------------------------------- Morton keys

@pytest.mark.parametrize("p,expected", [
    ((-1.0, -1.0, -1.0), 0x8000000000000000),
    ((0.0, 0.0, 0.0), 0xF000000000000000),
    ((0.5, -0.5, 0.0), 0xDC00000000000000),
    ((1.0, 0.0, 0.0), 0xB000000000000000),    # x = 2^21 wraps to 0
    ((-1.5, 0.0, 0.0), 0xF924000000000000),   # negative scaled x: x86-64 int64 truncation
])
def test_morton_kats(p, expected):
    assert orc.morton_key(np.array(p), 1.0) == expected


def test_x86_conversion_semantics():
    assert orc.x86_cvt_u64(5.99) == 5
    assert orc.x86_cvt_u64(-0.5) == 0
    assert orc.x86_cvt_u64(-1.0) == 0xFFFFFFFFFFFFFFFF
    assert orc.x86_cvt_u64(-524288.0) == (1 << 64) - 524288
    assert orc.x86_cvt_u64(2.0 ** 63) == 0x8000000000000000
    assert orc.x86_cvt_u64(float("nan")) == 0x8000000000000000
    assert orc.x86_cvt_u64(2.0 ** 62) == 1 << 62


def test_spread_matches_bit_loop_for_any_64bit_input():
    rng = np.random.default_rng(7)
    ts = [int(v) for v in rng.integers(0, 2 ** 63, 400, dtype=np.int64)]
    ts += [int(v) | (1 << 63) for v in rng.integers(0, 2 ** 63, 200, dtype=np.int64)]
    ts += [0, 1, (1 << 21) - 1, 1 << 21, (1 << 64) - 1, 0x1F << 48, 0x1F << 16]
    for t in ts:
        e = effective21(t)
        want = interleave_bruteforce(e, 0, 0) & ~(1 << 63)
        assert orc.spread21(t) << 2 == want, hex(t)


def test_morton_in_range_matches_bruteforce():
    rng = np.random.default_rng(3)
    size = 3.7
    pos = rng.uniform(-size, size * 0.999, (2000, 3))
    keys = orc.morton_keys(pos, size)
    frac = 2097152.0 / (2.0 * size)
    for i in range(0, 2000, 7):
        xi, yi, zi = (int(math.floor((pos[i, a] + size) * frac)) for a in range(3))
        assert int(keys[i]) == interleave_bruteforce(xi, yi, zi)


# ---------------------------------------------------------------- radix sort

def test_radix_sort_is_stable_full_sort():
    rng = np.random.default_rng(11)
    base = rng.integers(0, 2 ** 63, 64, dtype=np.int64).astype(np.uint64) | np.uint64(1 << 63)
    keys = base[rng.integers(0, 64, 5000)]                       # many exact duplicates
    keys[:100] ^= np.uint64(1 << 21)                              # digits straddling pass windows
    keys[100:200] ^= np.uint64(1 << 42)
    sk, idx = orc.radix_sort(keys)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(idx.astype(np.int64), ref)
    assert np.array_equal(sk, keys[ref])


def test_radix_sort_overlapping_windows_cover_all_bits():
    # keys differing only in bits 20, 41, 62 and 63 (edges of the nine windows)
    keys = np.array([1 << 63 | 1 << 62, 1 << 63 | 1 << 41, 1 << 63 | 1 << 20, 1 << 63, 1 << 62], dtype=np.uint64)
    sk, idx = orc.radix_sort(keys)
    assert list(idx) == [4, 3, 2, 1, 0]


# ---------------------------------------------------------------- bounds

def test_bounds_closed_form():
    pos = np.array([[1.0, 2.0, 3.0], [3.0, 4.0, 5.0]])
    c, s = orc.bounds(pos)
    assert list(c) == [2.0, 3.0, 4.0] and s == 1.0
    pos = np.array([[-4.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    c, s = orc.bounds(pos)                 # c = (-1,0,0); |-4|-(-1) = 5
    assert c[0] == -1.0 and s == 5.0


def test_bounds_is_sequential_left_to_right():
    pos = np.zeros((4, 3))
    pos[:, 0] = [1e16, 1.0, -1e16, 1.0]      # ((1e16+1)-1e16)+1 = 1 in FP64, exact sum 2
    c, _ = orc.bounds(pos)
    assert c[0] == 0.25


def test_bounds_empty():
    c, s = orc.bounds(np.zeros((0, 3)))
    assert list(c) == [0.0, 0.0, 0.0] and s == 0.0


# ---------------------------------------------------------------- octree

def _node_map(nodes):
    return {int(k): i for i, k in enumerate(nodes["key"])}


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (37, 2), (3000, 3)])
def test_insert_matches_recursive_definition(n, seed):
    pos, _, mass = plummer_np(n, seed)
    mass = mass * (1.0 + 0.1 * np.random.default_rng(seed).random(n))   # unequal masses
    c, s = orc.bounds(pos)
    ti = orc.build_insert(pos, mass, c, s)
    tr = orc.build_recursive(pos, mass, c, s, 1)
    a, b = ti.nodes(), tr.nodes()
    assert len(a) == len(b)
    a = np.sort(a, order="key")
    b = np.sort(b, order="key")
    for f in ("key", "center", "size", "com", "mass", "n", "first", "level", "child_byte"):
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("leaf_cap", [1, 4, 32])
def test_node_set_matches_path_key_definition(leaf_cap):
    pos, _, mass = plummer_np(4000, 5)
    c, s = orc.bounds(pos)
    t = orc.build_recursive(pos, mass, c, s, leaf_cap)
    pk = orc.tree_path_keys(pos, c, s)
    want = canonical_node_keys(pk, leaf_cap)
    got = set(int(k) for k in t.nodes()["key"])
    assert got == want


def test_tree_invariants():
    """Structure and as-run payload of the insertion build, each field against a
    definition from the path keys: child bits, counts, edge = S / 2^L, first member,
    barycentre = position of the first member in array order, and mass = the
    left-to-right sum of the members' masses in array order (insertion adds each
    body's mass to every node on its path, HashedOctree.cpp:450-711).  Masses are
    uneven, so the summation order shows in the bits."""
    n = 5000
    pos, _, _ = plummer_np(n, 9)
    mass = np.random.default_rng(9).uniform(0.1, 1.0, n) / n
    c, s = orc.bounds(pos)
    t = orc.build_insert(pos, mass, c, s)
    nd = t.nodes()
    m = _node_map(nd)
    pk = orc.tree_path_keys(pos, c, s)
    order_matters = 0
    for i, rec in enumerate(nd):
        k = int(rec["key"])
        kids = [m.get((k << 3) | o) for o in range(8)]
        bits = sum(1 << o for o in range(8) if kids[o] is not None)
        assert bits == rec["child_byte"]
        if bits:
            assert rec["n"] == sum(nd[j]["n"] for j in kids if j is not None)
        else:
            assert rec["n"] == 1 or t.unsplit
        L = int(rec["level"])
        assert rec["size"] == s * 0.5 ** L
        members = np.nonzero((pk >> np.uint64(63 - 3 * L)) == np.uint64(k))[0]
        assert members.size == rec["n"]
        assert rec["first"] == members.min()
        assert np.array_equal(rec["com"], pos[members.min()])           # as-run barycentre
        seq = 0.0
        for j in members:                                               # array order
            seq = seq + float(mass[j])
        assert rec["mass"] == seq
        order_matters += int(math.fsum(mass[members][::-1]) != seq or sum(mass[members][::-1]) != seq)
    assert order_matters > 0          # some node's mass depends on the order of its sum


def test_node_centres_follow_descent():
    pos, _, mass = plummer_np(500, 4)
    c, s = orc.bounds(pos)
    nd = orc.build_insert(pos, mass, c, s).nodes()
    for rec in nd:
        k = int(rec["key"])
        L = int(rec["level"])
        cc = np.array(c, dtype=np.float64)
        cs = s
        for lev in range(1, L + 1):
            o = (k >> (3 * (L - lev))) & 7
            cs = cs * 0.5
            d = np.array([1.0 if o & 4 else -1.0, 1.0 if o & 2 else -1.0, 1.0 if o & 1 else -1.0])
            cc = cc + (d * 0.5) * cs
        assert np.array_equal(cc, rec["center"])


def test_empty_and_single():
    t = orc.build_insert(np.zeros((0, 3)), np.zeros(0), np.zeros(3), 0.0)
    assert len(t) == 0
    pos = np.array([[0.3, 0.2, 0.1]])
    t = orc.build_insert(pos, np.array([2.0]), np.zeros(3), 1.0)
    nd = t.nodes()
    assert len(nd) == 1 and nd[0]["n"] == 1 and nd[0]["child_byte"] == 0
    acc, cnt = t.walk(0.5, 0.01)
    assert cnt[0] == 0 and np.all(acc == 0)


def test_coincident_bodies_stop_at_depth_cap():
    pos = np.array([[0.25, 0.25, 0.25]] * 3 + [[-0.3, 0.1, 0.2]])
    t = orc.build_insert(pos, np.ones(4), np.zeros(3), 1.0)
    nd = t.nodes()
    assert nd["level"].max() == 21 and t.unsplit == 1
    leaf = nd[np.argmax(nd["level"])]
    assert leaf["n"] == 3 and leaf["child_byte"] == 0


def test_intended_moments_match_member_sums():
    pos, _, mass = plummer_np(3000, 21)
    mass = mass * (0.5 + np.random.default_rng(2).random(3000))
    c, s = orc.bounds(pos)
    for cap in (1, 8):
        t = orc.build_recursive(pos, mass, c, s, cap)
        t.moments_intended()
        nd = t.nodes()
        pk = orc.tree_path_keys(pos, c, s)
        for rec in nd[::17]:
            L = int(rec["level"])
            mem = np.nonzero((pk >> np.uint64(63 - 3 * L)) == rec["key"])[0]
            M = mass[mem].sum()
            com = (mass[mem, None] * pos[mem]).sum(0) / M
            if mem.size == 1:
                com = pos[mem[0]]                # a one-body cell sits exactly on its body
            d = pos[mem] - com
            d2 = (d * d).sum(1)
            Q = np.zeros(6)
            for q, (a, b) in enumerate([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]):
                Q[q] = (mass[mem] * (3 * d[:, a] * d[:, b] - (d2 if a == b else 0))).sum()
            assert rec["mass"] == pytest.approx(M, rel=1e-13)
            assert np.allclose(rec["com"], com, rtol=0, atol=1e-12 * max(1.0, np.abs(com).max()))
            scale = (mass[mem] * d2).sum() + 1e-300
            assert np.allclose(rec["quad"], Q, rtol=0, atol=1e-10 * scale)
            assert abs(rec["quad"][0] + rec["quad"][3] + rec["quad"][5]) <= 1e-10 * scale


# ---------------------------------------------------------------- walk / forces

def test_two_body_closed_form():
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    mass = np.array([0.25, 0.5])
    eps = 0.1
    acc, pot = orc.direct(pos, mass, eps)
    f = 1.0 / (1.0 + eps * eps) ** 1.5
    assert acc[0] == pytest.approx([0.5 * f, 0, 0], rel=1e-15)
    assert acc[1] == pytest.approx([-0.25 * f, 0, 0], rel=1e-15)
    assert pot[0] == pytest.approx(-0.5 / math.sqrt(1 + eps * eps), rel=1e-15)
    c, s = orc.bounds(pos)
    t = orc.build_insert(pos, mass, c, s)
    a2, cnt = t.walk(0.5, eps)
    assert np.allclose(a2, acc, rtol=1e-15, atol=0) and list(cnt) == [1, 1]


def test_direct_momentum_conservation():
    pos, _, mass = plummer_np(800, 8)
    acc, _ = orc.direct(pos, mass, 0.01)
    ptot = (mass[:, None] * acc).sum(0)
    assert np.abs(ptot).max() < 1e-12 * np.abs(mass[:, None] * acc).sum()


@pytest.mark.parametrize("leaf_cap", [1, 6])
def test_theta_zero_reduces_to_direct_sum(leaf_cap):
    pos, _, mass = plummer_np(1500, 12)
    c, s = orc.bounds(pos)
    t = orc.build_recursive(pos, mass, c, s, leaf_cap) if leaf_cap > 1 else orc.build_insert(pos, mass, c, s)
    acc, cnt = t.walk(0.0, 0.01)
    ref, _ = orc.direct(pos, mass, 0.01)
    assert np.all(cnt == 1499)
    err = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < 1e-12


def test_walk_error_shrinks_with_theta_and_quadrupole():
    pos, _, mass = plummer_np(3000, 14)
    c, s = orc.bounds(pos)
    ref, _ = orc.direct(pos, mass, 0.01)
    t = orc.build_recursive(pos, mass, c, s, 1)
    t.moments_intended()
    errs = {}
    for theta in (1.0, 0.5, 0.3):
        for quad in (False, True):
            acc, _ = t.walk(theta, 0.01, quad=quad)
            e = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
            errs[(theta, quad)] = np.median(e)
    assert errs[(0.3, False)] < errs[(0.5, False)] < errs[(1.0, False)]
    assert errs[(1.0, True)] < errs[(1.0, False)]
    for theta in (0.5, 0.3):
        assert errs[(theta, True)] < 0.5 * errs[(theta, False)]


def test_walk_as_run_uses_first_body_barycentre():
    # Two far clusters: as run, a cluster's node sits at its first-inserted body.
    pos = np.array([[10.0, 0, 0], [10.5, 0, 0], [-10.0, 0, 0]])
    mass = np.array([1.0, 1.0, 1.0])
    c = np.array([0.0, 0, 0])
    t = orc.build_insert(pos, mass, c, 20.0)
    acc, cnt = t.walk(1.0, 0.0, targets=np.array([2]))
    # node of bodies 0,1 has mass 2 at x=10 (body 0), edge 10 -> MAC 100/400 < 1 passes
    assert cnt[0] == 1
    assert acc[0, 0] == pytest.approx(2.0 / 20.0 ** 2, rel=1e-15)


# ---------------------------------------------------------------- KDK

def test_kdk_closed_form_and_reversibility():
    pos = np.array([[0.0, 1.0, -2.0]])
    vel = np.array([[1.0, 0.5, 0.0]])
    acc = np.array([[2.0, -4.0, 8.0]])
    p1, v1 = orc.kick_drift(0.5, pos, vel, acc)
    assert list(v1[0]) == [1.5, -0.5, 2.0] and list(p1[0]) == [0.75, 0.75, -1.0]
    v2 = orc.kick(0.5, v1, acc)
    assert list(v2[0]) == [2.0, -1.5, 4.0]
    # with power-of-two values every operation is exact, so stepping back with the
    # negated velocity returns exactly: drift back (-v2 kicked by a: -v1), then kick
    p3, v3 = orc.kick_drift(0.5, p1, -v2, acc)
    assert list(v3[0]) == list(-v1[0]) and np.array_equal(p3, pos)
    v4 = orc.kick(0.5, v3, acc)
    assert np.array_equal(-v4, vel)


def test_frame_order_matches_composition():
    pos, vel, mass = plummer_np(700, 30)
    acc0 = np.random.default_rng(1).normal(size=(700, 3)) * 1e-3
    p, v, m, a, keys, st = orc.frame(pos, vel, mass, acc0, 1e-3, 0.5, 0.025)
    c, s = orc.bounds(pos)
    k = orc.morton_keys(pos, s)
    sk, idx = orc.radix_sort(k)
    idx = idx.astype(np.int64)
    pp, vv = orc.kick_drift(1e-3, pos[idx], vel[idx], acc0[idx])
    t = orc.build_insert(pp, mass[idx], c, s)
    aa, _ = t.walk(0.5, 0.025)
    vv = orc.kick(1e-3, vv, aa)
    assert np.array_equal(keys, sk)
    assert np.array_equal(p, pp) and np.array_equal(v, vv) and np.array_equal(a, aa)
    assert st.nodes == len(t)


# ---------------------------------------------------------------- trajectory gradient


def _traj_loss(p, v, m, gx, gv, eps, dt, T, init_acc=True, acc=None, ga=None):
    P, V, A = orc.traj_direct(p, v, m, eps, dt, T, init_acc=init_acc, acc=acc)
    out = (gx * P).sum() + (gv * V).sum()
    if ga is not None:
        out += (ga * A).sum()
    return out


@pytest.mark.parametrize("init_acc", [True, False])
def test_traj_adjoint_matches_central_differences(init_acc):
    """The direct-sum trajectory has no decisions, so central differences pin its adjoint."""
    rng = np.random.default_rng(3)
    n, eps, dt, T, h = 7, 0.05, 1e-2, 12, 1e-6
    p = rng.normal(size=(n, 3))
    v = 0.3 * rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 1.5, n) / n
    a0 = None if init_acc else 0.1 * rng.normal(size=(n, 3))
    gx, gv, ga = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    G = orc.traj_direct_adjoint(p, v, m, eps, dt, T, gx, gv, ga, init_acc=init_acc, acc=a0)
    L = lambda p_, v_, m_, a_: _traj_loss(p_, v_, m_, gx, gv, eps, dt, T, init_acc, a_, ga)
    for (i, a) in [(0, 0), (3, 2), (6, 1)]:
        for arg in range(2 if init_acc else 3):
            base = [p, v, a0][arg]
            hp, hm = base.copy(), base.copy()
            hp[i, a] += h
            hm[i, a] -= h
            args_p = [p, v, a0]
            args_m = [p, v, a0]
            args_p[arg], args_m[arg] = hp, hm
            fd = (L(args_p[0], args_p[1], m, args_p[2]) - L(args_m[0], args_m[1], m, args_m[2])) / (2 * h)
            got = G[arg][i, a]
            assert abs(fd - got) <= 1e-6 * abs(fd) + 1e-9, (arg, i, a, fd, got)
    for i in (0, 5):
        hh = 1e-7
        mp, mm = m.copy(), m.copy()
        mp[i] += hh
        mm[i] -= hh
        fd = (L(p, v, mp, a0) - L(p, v, mm, a0)) / (2 * hh)
        assert abs(fd - G[3][i]) <= 1e-5 * abs(fd) + 1e-9
    if init_acc:
        assert np.all(G[2] == 0.0)


def test_traj_adjoint_two_body_one_frame_closed_form():
    """One frame of two bodies: dL/dm for L = x_T[0, 0] from the closed form of the kick-drift-kick."""
    eps, dt = 0.1, 0.05
    p = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    v = np.zeros((2, 3))
    m = np.array([0.6, 0.4])
    gx = np.zeros((2, 3))
    gx[0, 0] = 1.0
    G = orc.traj_direct_adjoint(p, v, m, eps, dt, 1, gx, np.zeros((2, 3)), init_acc=True)
    # x_T[0,0] = x0 + (a0 dt/2) dt with a0_x = m1 r^-3 (r^2 = 1 + eps^2): d/dm1 = dt^2/2 r^-3
    r3 = (1.0 + eps * eps) ** 1.5
    assert G[3][1] == pytest.approx(0.5 * dt * dt / r3, rel=1e-12)
    assert G[3][0] == pytest.approx(0.0, abs=1e-15)


# ---------------------------------------------------------------- rounding-order known answers
#
# Each value below is derived by hand from IEEE-754 round-to-nearest-even of the
# reference's own expression, and each input is chosen so that the plausible
# alternative form rounds differently: the test then runs the alternative (a
# copy of the oracle compiled with that one mistake planted) and requires it to
# FAIL the same known answer.

def _mutant(tmp_path, define):
    return orc.build_mutant(define, tmp_path / (define.lower() + ".so"))


def _bounds_with(L, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    c = np.zeros(3)
    s = np.zeros(1)
    L.orc_bounds(pos.ctypes.data_as(ctypes.c_void_p), pos.shape[0], c.ctypes.data_as(ctypes.c_void_p),
                 s.ctypes.data_as(ctypes.c_void_p))
    return c, float(s[0])


def test_bounds_centre_kat_multiplies_by_reciprocal(tmp_path):
    """OctantBounds' centre is sum * fl(1/N) (Vec3D::operator/= multiplies by the
    reciprocal, MathUtils.cpp:81-90), N = 3."""
    c, s = orc.bounds(MEAN_KAT_POS)
    assert list(c) == MEAN_KAT_C
    assert s == 4.0 - MEAN_KAT_C[2]
    cm, _ = _bounds_with(_mutant(tmp_path, "ORC_MUTANT_MEAN_DIV"), MEAN_KAT_POS)
    assert list(cm) != MEAN_KAT_C and cm[0] == float.fromhex("0x1.ddddddddddddfp-3")


def _kdk_with(L, dt, pos, vel, acc, drift=True):
    pos, vel, acc = [np.array(a, np.float64, copy=True) for a in (pos, vel, acc)]
    P = ctypes.c_void_p
    if drift:
        L.orc_kick_drift(dt, pos.ctypes.data_as(P), vel.ctypes.data_as(P), acc.ctypes.data_as(P), pos.shape[0])
    else:
        L.orc_kick(dt, vel.ctypes.data_as(P), acc.ctypes.data_as(P), vel.shape[0])
    return pos, vel


def test_kick_drift_kat_rounding_order(tmp_path):
    """v + (a dt) 0.5 and x + v dt with each product rounded before the addition
    (Body.hpp:310-311,322): normal-range cases against fused multiply-adds, and a
    subnormal case against a (dt 0.5)."""
    dt = 1e-3
    v = orc.kick(dt, KICK_KAT["v"], KICK_KAT["a"])
    assert v[0, 0] == KICK_KAT["want"]
    x, _ = orc.kick_drift(dt, DRIFT_KAT["x"], np.zeros((1, 3)), np.zeros((1, 3)))
    assert x[0, 0] == -0.002
    # drift alone: zero acceleration leaves v, then x + fl(v dt)
    x, _ = orc.kick_drift(dt, DRIFT_KAT["x"], DRIFT_KAT["v"], np.zeros((1, 3)))
    assert x[0, 0] == DRIFT_KAT["want"]
    xs, vs = orc.kick_drift(2.75, np.zeros((1, 3)), np.zeros((1, 3)), [[TINY, 0.0, 0.0]])
    assert vs[0, 0] == 2 * TINY and xs[0, 0] == 6 * TINY

    fm = _mutant(tmp_path, "ORC_MUTANT_FMA")
    _, vm = _kdk_with(fm, dt, np.zeros((1, 3)), KICK_KAT["v"], KICK_KAT["a"], drift=False)
    assert vm[0, 0] != KICK_KAT["want"]
    xm, _ = _kdk_with(fm, dt, DRIFT_KAT["x"], DRIFT_KAT["v"], np.zeros((1, 3)))
    assert xm[0, 0] != DRIFT_KAT["want"]
    hm = _mutant(tmp_path, "ORC_MUTANT_HALF_DT")
    _, vh = _kdk_with(hm, 2.75, np.zeros((1, 3)), np.zeros((1, 3)), [[TINY, 0.0, 0.0]])
    assert vh[0, 0] == TINY


def test_mac_kat_divides(tmp_path):
    """MAC is (s*s)/(d.d) < theta*theta (HashedOctree.cpp:875-884), at two boundary
    cases where the multiplied form decides the other way."""
    mm = _mutant(tmp_path, "ORC_MUTANT_MAC_MUL")
    r = np.zeros(3)
    for com, size_hex, want in MAC_KATS:
        size = float.fromhex(size_hex)
        com = np.array(com)
        assert orc.mac(size, com, r, 0.6) is want
        got = bool(mm.orc_mac_raw(size, com.ctypes.data_as(ctypes.c_void_p), r.ctypes.data_as(ctypes.c_void_p),
                                  0.6))
        assert got is (not want)


# ---------------------------------------------------------------- decodeMortonKey, list order

def test_decode_morton_key_inverts_the_encoding():
    """decodeMortonKey (MortonKeys.hpp:204-228): the survey's encode KATs decode back
    to their points, and at a power-of-two size (every operation exact) any key
    decodes to its three 21-bit grid coordinates c as c * 2^-20 * S - S."""
    assert list(orc.decode_morton_key(0xF000000000000000, 1.0)) == [0.0, 0.0, 0.0]
    assert list(orc.decode_morton_key(0x8000000000000000, 1.0)) == [-1.0, -1.0, -1.0]
    assert list(orc.decode_morton_key(0xDC00000000000000, 1.0)) == [0.5, -0.5, 0.0]
    rng = np.random.default_rng(3)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(0, 1 << 21, 3))
        key = interleave_bruteforce(x, y, z)
        S = 2.0 ** int(rng.integers(-5, 6))
        got = orc.decode_morton_key(key, S)
        assert list(got) == [c * 2.0 ** -20 * S - S for c in (x, y, z)]
        # and encoding the decoded grid point gives the key back
        assert orc.morton_key(got, S) == key


def _reference_order_list(nodes, by_key, members, pos, p, theta):
    """The interaction list of point p written recursively: a cell emits its
    accepted children in octant order, then searches its opened children from
    octant 7 down to 0 (the explicit stack of HashedOctree.hpp:351-380 pops the
    last pushed child first); a leaf emits itself (or its member bodies)
    unless at p.  Items: ('n', key) or ('b', body)."""
    out = []
    theta2 = theta * theta

    def mac(rec):
        d = rec["com"] - p
        d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
        with np.errstate(divide="ignore"):      # d = 0 (the point itself): s^2/0 = inf, opened
            return (rec["size"] * rec["size"]) / d2 < theta2

    def visit(rec):
        if rec["child_byte"]:
            opened = []
            for o in range(8):
                if rec["child_byte"] >> o & 1:
                    ch = nodes[by_key[(int(rec["key"]) << 3) | o]]
                    if mac(ch):
                        out.append(("n", int(ch["key"])))
                    else:
                        opened.append(ch)
            for ch in reversed(opened):
                visit(ch)
        elif rec["member_count"] <= 1:
            if np.any(rec["com"] != p):
                out.append(("n", int(rec["key"])))
        else:
            for i in members[int(rec["key"])]:
                if np.any(pos[i] != p):
                    out.append(("b", int(i)))

    visit(nodes[by_key[1]])
    return out


@pytest.mark.parametrize("cap", [1, 4])
def test_interaction_list_order_matches_recursive_definition(cap):
    pos, _, mass = plummer_np(600, 12)
    c, s = orc.bounds(pos)
    t = orc.build_insert(pos, mass, c, s) if cap == 1 else orc.build_recursive(pos, mass, c, s, cap)
    nd = t.nodes()
    by_key = {int(k): i for i, k in enumerate(nd["key"])}
    pk = orc.tree_path_keys(pos, c, s)
    members = {}
    for rec in nd[nd["child_byte"] == 0]:
        L = int(rec["level"])
        members[int(rec["key"])] = np.nonzero((pk >> np.uint64(63 - 3 * L)) == rec["key"])[0]
    _, counts = t.walk(0.5, 0.01)
    for q in list(range(0, 600, 37)) + [-1]:
        p = pos[q] if q >= 0 else np.array([0.3, -0.1, 2.0])
        ids, kind = t.interaction_list(p, 0.5)
        got = [("b", int(i)) if k else ("n", int(nd["key"][i])) for i, k in zip(ids, kind)]
        assert got == _reference_order_list(nd, by_key, members, pos, p, 0.5)
        if q >= 0:
            assert len(got) == counts[q]
