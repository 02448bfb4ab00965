"""The tree order a frame builds, against its definition: the stable sort (in the
frame's new array order) of the path keys of the drifted positions in the bounds
of the positions the frame started from.  With accelerations moving with their
bodies the frame sorts those keys on a stream of its own, from the order before
the body sort, and rejoins through the inverse body permutation; bodies with
equal path keys (here: copies of bodies, so runs of equal keys exist) must come
out in the new array order.  Eager frames and graph frames, one system and an
ensemble.  Bit-exact.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import load_bhx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


def _with_copies(bhx, n, seed):
    pos, vel, mass = bhx.plummer(n, seed=seed)
    for k in range(24):                      # triples of identical bodies
        for off in (n // 3, 2 * n // 3):
            pos[off + k], vel[off + k], mass[off + k] = pos[k], vel[k], mass[k]
    return pos, vel, mass


def _expected(p_before, p_after):
    c, s = orc.bounds(p_before)
    pk = orc.tree_path_keys(p_after, c, s)
    return np.sort(pk), np.argsort(pk, kind="stable")


@pytest.mark.parametrize("graph", [False, True])
def test_frame_tree_order_with_equal_keys(bhx, graph):
    n = 6000
    pos, vel, mass = _with_copies(bhx, n, 50)
    print("frame tree order n=%d graph=%d" % (n, graph), flush=True)
    with bhx.Engine(n, theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=8, mode=1, precision=1, group_walk=1,
                    permute_acc=1) as e:
        e.set_bodies(pos, vel, mass)
        runs = 0
        for f in range(3):
            p0 = e.get_bodies()[0]
            if graph:
                e.step_graph(1)
            else:
                e.step(1)
            p1 = e.get_bodies()[0]
            tperm, tkeys = e.tree_order()
            want_k, want_p = _expected(p0, p1)
            runs += int(np.sum(want_k[1:] == want_k[:-1]))
            assert np.array_equal(tkeys, want_k), f
            assert np.array_equal(tperm.astype(np.int64), want_p), f
        assert runs > 0, "no equal-key runs exercised"


def test_frame_tree_order_ensemble(bhx):
    sizes = [1500, 700, 2300]
    parts = [_with_copies(bhx, k, 60 + i) for i, k in enumerate(sizes)]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    with bhx.Engine(pos.shape[0], theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=8, mode=1, precision=1,
                    group_walk=1, permute_acc=1) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        for f in range(2):
            p0 = e.get_bodies()[0]
            e.step_graph(1)
            p1 = e.get_bodies()[0]
            tperm, tkeys = e.tree_order()
            tperm = tperm.astype(np.int64)
            for k in range(len(sizes)):
                a, b = int(off[k]), int(off[k + 1])
                want_k, want_p = _expected(p0[a:b], p1[a:b])
                assert np.array_equal(tkeys[a:b], want_k), (f, k)
                assert np.array_equal(tperm[a:b] - a, want_p), (f, k)
