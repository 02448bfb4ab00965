"""Bit-exact parity at the headline size: 1,048,576 Plummer bodies (seed 1), leaf
capacity 32, the precision workload's configuration, after two graph frames
(so the bodies are in the Morton order the step sees).  Through the C ABI on the
stepped state, against the oracle on the same positions: bounds, Morton keys,
the sort permutation, the tree path keys and tree permutation, and the node set
(key, centre, size, count, level, child bits, first body) for the as-run and
the intended payloads; as-run mass and barycentre bit for bit; intended masses,
centres of mass and quadrupoles field by field within stated bounds.  The
intended moments are also compared at 16,384 bodies with leaf capacity 1."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import load_bhx

pytestmark = pytest.mark.gpu

N = 1 << 20
CFG = dict(theta=0.5, eps=0.01, dt=5e-4, leaf_capacity=32, mode=1, precision=1, quadrupole=1, compensated=1,
           permute_acc=1, group_walk=1)
STRUCT = ["key", "center", "size", "n", "level", "child_byte", "first"]


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


def _by_key(a):
    return np.sort(a, order="key")


def _moment_errors(g, o):
    """Per-node errors of intended moments, each scaled by what sets its size:
    mass relative; centre of mass in units of the node's edge; quadrupole in units
    of M * edge^2 (the scale of its entries)."""
    g, o = _by_key(g), _by_key(o)
    assert np.array_equal(g["key"], o["key"])
    em = np.abs(g["mass"] - o["mass"]) / o["mass"]
    ec = np.abs(g["com"] - o["com"]).max(axis=1) / o["size"]
    eq = np.abs(g["quad"] - o["quad"]).max(axis=1) / (o["mass"] * o["size"] ** 2)
    return em, ec, eq


def test_headline_keys_sort_tree_bitexact(bhx):
    pos, vel, mass = bhx.plummer(N, seed=1)
    print("headline n=%d: two graph frames" % N, flush=True)
    with bhx.Engine(N, **CFG) as e:
        e.set_bodies(pos, vel, mass)
        e.step_graph(2)
        p, v, m, _ = e.get_bodies()
        c, s = e.bounds()
        path = e.bounds_path()
        keys = e.morton_keys(s)
        perm = e.sort_bodies()
        e.build(c, s)
        g_int = e.nodes()
        tperm, tkeys = e.tree_order()
        e.set_config(mode=0)
        e.build(c, s)
        g_run = e.nodes()
    print("bounds path %s, nodes %d" % (path, len(g_run)), flush=True)
    oc, os_ = orc.bounds(p)
    assert c.tobytes() == oc.tobytes() and s == os_
    assert path[3] == 0            # Morton-ordered positions: the chain over exact runs ran
    okeys = orc.morton_keys(p, os_)
    assert np.array_equal(keys, okeys)
    _, idx = orc.radix_sort(okeys)
    assert np.array_equal(perm, idx)
    ii = idx.astype(np.int64)
    ps, ms = p[ii], m[ii]
    pk = orc.tree_path_keys(ps, oc, os_)
    assert np.array_equal(tkeys, np.sort(pk))
    assert np.array_equal(tperm.astype(np.int64), np.argsort(pk, kind="stable"))
    t = orc.build_recursive(ps, ms, oc, os_, 32)
    o = _by_key(t.nodes())
    assert len(g_run) == len(o) == len(g_int)
    gr, gi = _by_key(g_run), _by_key(g_int)
    for f in STRUCT:
        assert np.array_equal(gr[f], o[f]), f
        assert np.array_equal(gi[f], o[f]), f
    # as-run payload: barycentre = first body; equal masses 2^-20 make every sum exact
    assert np.array_equal(gr["com"], o["com"]) and np.array_equal(gr["mass"], o["mass"])
    t.moments_intended()
    em, ec, eq = _moment_errors(g_int, t.nodes())
    print("intended moments (B=32, 2^20): mass %.2e com %.2e quad %.2e (max)" % (em.max(), ec.max(), eq.max()),
          flush=True)
    # the GPU sums a leaf's bodies in tree order and the oracle in array order:
    # equal masses 2^-20 keep mass sums exact; centres and quadrupoles differ by
    # rounding of sums of up to 2^20 terms
    assert em.max() == 0.0
    assert ec.max() < 1e-12 and eq.max() < 1e-10


@pytest.mark.parametrize("cap", [1, 32])
def test_intended_moments_field_by_field(bhx, cap):
    """Intended mass, centre of mass and quadrupole of every node against the
    oracle's textbook recursion, on uneven masses (every sum rounds)."""
    n = 16384
    pos, vel, _ = bhx.plummer(n, seed=3)
    mass = np.random.default_rng(3).uniform(0.5, 1.5, n) / n
    c, s = orc.bounds(pos)
    with bhx.Engine(n, leaf_capacity=cap, mode=1, quadrupole=1) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        g = e.nodes()
    t = orc.build_insert(pos, mass, c, s) if cap == 1 else orc.build_recursive(pos, mass, c, s, cap)
    t.moments_intended()
    o = t.nodes()
    gs, os_ = _by_key(g), _by_key(o)
    for f in STRUCT:
        assert np.array_equal(gs[f], os_[f]), f
    em, ec, eq = _moment_errors(g, o)
    print("intended moments (B=%d): mass %.2e com %.2e quad %.2e (max)" % (cap, em.max(), ec.max(), eq.max()),
          flush=True)
    # one-body cells: the body itself, exactly
    one = os_["n"] == 1
    assert np.array_equal(gs["com"][one], os_["com"][one]) and np.array_equal(gs["mass"][one], os_["mass"][one])
    assert np.all(gs["quad"][one] == 0)
    assert em.max() < 1e-13 and ec.max() < 1e-12 and eq.max() < 1e-10
