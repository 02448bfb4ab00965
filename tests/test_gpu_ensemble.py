"""Ensembles of independent systems in one context (bhx_set_systems), against
the same systems stepped one at a time in their own contexts and against the
oracle's frames.

Each system of an ensemble must evolve exactly as it does alone: the per-system
bounds, keys, sorts, trees and walk groups make every kernel do the same
arithmetic on the same operands, so the comparison with looped single-system
contexts is bitwise.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import load_bhx

pytestmark = pytest.mark.gpu

SIZES = [1000, 3001, 64, 129, 2500, 1, 2, 700]


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


def _systems(bhx, sizes, seed0=11):
    parts = [bhx.plummer(n, seed=seed0 + k, total_mass=1.0 + 0.1 * k, a=0.8 + 0.05 * k) for k, n in enumerate(sizes)]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return parts, pos, vel, mass, off


CONFIGS = {
    "as_run_b1_fp64": dict(theta=0.5, eps=0.01, dt=1e-3),
    "intended_b32_fp32_group": dict(theta=0.5, eps=0.01, dt=5e-4, leaf_capacity=32, mode=1, precision=1,
                                    quadrupole=1, compensated=1, group_walk=1),
    "intended_b8_fp64_group": dict(theta=0.6, eps=0.02, dt=1e-3, leaf_capacity=8, mode=1, precision=0,
                                   quadrupole=1, group_walk=1),
}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_ensemble_steps_equal_looped_systems_bitwise(bhx, name):
    cfg = CONFIGS[name]
    parts, pos, vel, mass, off = _systems(bhx, SIZES)
    steps = 3
    print("ensemble %s systems=%d n=%d" % (name, len(SIZES), pos.shape[0]), flush=True)
    with bhx.Engine(pos.shape[0], **cfg) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        assert e.nsys == len(SIZES)
        e.step(steps)
        P, V, M, A = e.get_bodies()
    for k, (p, v, m) in enumerate(parts):
        a, b = int(off[k]), int(off[k + 1])
        with bhx.Engine(p.shape[0], **cfg) as s:
            s.set_bodies(p, v, m)
            s.step(steps)
            p1, v1, m1, a1 = s.get_bodies()
        assert np.array_equal(P[a:b], p1), "positions of system %d" % k
        assert np.array_equal(V[a:b], v1), "velocities of system %d" % k
        assert np.array_equal(M[a:b], m1), "masses of system %d" % k
        assert np.array_equal(A[a:b], a1), "accelerations of system %d" % k


@pytest.mark.parametrize("prec", [0, 1])
def test_repeated_walks_cover_every_group(bhx, prec):
    """Every walk group of a ragged ensemble is walked on every call, with the same
    result: the accelerations are overwritten with NaN between walks."""
    parts, pos, vel, mass, off = _systems(bhx, SIZES + [20000], seed0=90)
    cfg = dict(theta=0.5, eps=0.01, leaf_capacity=16, mode=1, precision=prec, quadrupole=1, group_walk=1)
    print("repeated walks prec=%d n=%d" % (prec, pos.shape[0]), flush=True)
    with bhx.Engine(pos.shape[0], **cfg) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        C, S = e.system_bounds()
        e.build_systems(C, S)
        ref = e.walk()
        assert np.all(np.isfinite(ref))
        nan = np.full_like(ref, np.nan)
        for _ in range(3):
            e.set_accelerations(nan)
            assert np.array_equal(e.walk(), ref)


def test_ensemble_matches_oracle_frames(bhx):
    """As-run frames (leaf capacity 1, FP64 per-body walk) of every system against the oracle."""
    sizes = [300, 1, 77, 512]
    parts, pos, vel, mass, off = _systems(bhx, sizes, seed0=40)
    dt, theta, eps, frames = 1e-3, 0.5, 0.01, 2
    with bhx.Engine(pos.shape[0], theta=theta, eps=eps, dt=dt) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        e.step(frames)
        P, V, M, A = e.get_bodies()
    for k, (p, v, m) in enumerate(parts):
        a_, b_ = int(off[k]), int(off[k + 1])
        op, ov, om, oa = p, v, m, np.zeros_like(p)
        for _ in range(frames):
            op, ov, om, oa, _, _ = orc.frame(op, ov, om, oa, dt, theta, eps, permute_acc=True)
        assert np.array_equal(M[a_:b_], om)
        # tolerance: the walk sums each body's terms in another order (1e-12 relative on accelerations)
        assert np.allclose(P[a_:b_], op, rtol=0, atol=1e-13)
        assert np.allclose(V[a_:b_], ov, rtol=0, atol=1e-12)
        assert np.allclose(A[a_:b_], oa, rtol=1e-10, atol=1e-12)


def test_ensemble_bounds_tree_lookup(bhx):
    parts, pos, vel, mass, off = _systems(bhx, SIZES, seed0=70)
    cfg = dict(leaf_capacity=4, mode=1)
    with bhx.Engine(pos.shape[0], **cfg) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        C, S = e.system_bounds()
        total = e.build_systems(C, S)
        nodes = e.nodes()
        roots = e.system_roots().astype(np.int64)
        assert len(nodes) == total
        ends = np.append(roots[1:], total)
        for k, (p, v, m) in enumerate(parts):
            a, b = int(off[k]), int(off[k + 1])
            with bhx.Engine(p.shape[0], **cfg) as s:
                s.set_bodies(p, v, m)
                c1, s1 = s.bounds()
                assert np.array_equal(C[k], c1) and S[k] == s1, "bounds of system %d" % k
                n1 = s.build(c1, s1)
                g1 = s.nodes()
            g = nodes[roots[k]:ends[k]]
            assert len(g) == n1, "node count of system %d" % k
            for f in ("key", "n", "child_byte", "level", "center", "com", "mass", "size", "quad"):
                assert np.array_equal(g[f], g1[f]), "node field %s of system %d" % (f, k)
            assert np.array_equal(g["skip"].astype(np.int64) - roots[k], g1["skip"].astype(np.int64))
            assert np.array_equal(g["body_start"].astype(np.int64) - a, g1["body_start"].astype(np.int64))
            assert np.array_equal(g["first"].astype(np.int64) - a, g1["first"].astype(np.int64))
            ids = e.lookup_in(k, g["key"])
            assert np.array_equal(ids, np.arange(roots[k], ends[k]))
        # a key present only in another system's tree is missing here
        only1 = set(nodes[roots[1]:ends[1]]["key"].tolist()) - set(nodes[roots[0]:ends[0]]["key"].tolist())
        if only1:
            assert e.lookup_in(0, np.array(sorted(only1)[:5], np.uint64)).tolist() == [-1] * min(5, len(only1))


def test_ensemble_direct_sum_per_system(bhx):
    parts, pos, vel, mass, off = _systems(bhx, SIZES, seed0=90)
    rng = np.random.default_rng(5)
    targets = np.sort(rng.choice(pos.shape[0], 200, replace=False))
    with bhx.Engine(pos.shape[0]) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        acc, pot = e.direct(targets, 0.01)
    sysk = np.searchsorted(off, targets, side="right") - 1
    for k, (p, v, m) in enumerate(parts):
        sel = sysk == k
        if not sel.any():
            continue
        with bhx.Engine(p.shape[0]) as s:
            s.set_bodies(p, v, m)
            a1, p1 = s.direct(targets[sel] - int(off[k]), 0.01)
        assert np.array_equal(acc[sel], a1) and np.array_equal(pot[sel], p1)


def test_ensemble_argument_errors(bhx):
    parts, pos, vel, mass, off = _systems(bhx, [10, 20])
    with bhx.Engine(pos.shape[0]) as e:
        e.set_bodies(pos, vel, mass)
        for bad in ([0, 10], [1, 10, 30], [0, 10, 10, 30], [0, 40, 30]):
            with pytest.raises(bhx.BhxError):
                e.set_systems(np.array(bad, np.uint64))
        e.set_systems(off)
        with pytest.raises(bhx.BhxError):
            e.bounds()                      # one bounds box per system: bhx_system_bounds
        e.set_bodies(pos, vel, mass)        # resets to one system
        assert e.nsys == 1
        e.bounds()


def test_autotune_leaves_state_and_sets_capacity(bhx):
    parts, pos, vel, mass, off = _systems(bhx, [4000, 2500, 300], seed0=120)
    cfg = CONFIGS["intended_b32_fp32_group"]
    with bhx.Engine(pos.shape[0], **cfg) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        e.step(1)
        before = e.get_bodies()
        chosen, table = e.autotune(caps=(8, 32), steps=2, max_err=1e-2, nsample=512)
        assert chosen in (8, 32) and len(table) == 2
        for cap, ms, err in table:
            assert ms > 0 and 0 < err < 1e-2
        after = e.get_bodies()
        assert all(np.array_equal(x, y) for x, y in zip(before, after)), "state restored"
        # nothing is allowed: the configured capacity stays
        c2, _ = e.autotune(caps=(8, 32), steps=1, max_err=0.0, nsample=256)
        assert c2 == chosen
        # the group size too: every (capacity, group size) pair is measured
        c3, g3, table = e.autotune(caps=(chosen,), group_sizes=(64, 32, 16), steps=1, max_err=1e-2, nsample=256)
        assert c3 == chosen and g3 in (64, 32, 16) and [t[1] for t in table] == [64, 32, 16]
        assert all(t[2] > 0 and 0 < t[3] < 1e-2 for t in table)
        e.set_config(group_size=0)
        assert all(np.array_equal(x, y) for x, y in zip(before, e.get_bodies())), "state restored"
        e.step(2)
        P = e.get_bodies()
    with bhx.Engine(pos.shape[0], **dict(cfg, leaf_capacity=chosen)) as r:
        r.set_bodies(pos, vel, mass)
        r.set_systems(off)
        r.set_config(leaf_capacity=32)
        r.step(1)
        r.set_config(leaf_capacity=chosen)
        r.step(2)
        R = r.get_bodies()
    assert all(np.array_equal(x, y) for x, y in zip(P, R))
