"""Frames replayed from captured HIP graphs (bhx_step_graph) against the same
frames run eagerly (bhx_step): bitwise equal on every replay, with host
downloads and synchronisations between the replays, across configuration
changes, for ensembles, and through the restart that follows a node-array
overflow."""
import numpy as np
import pytest

from tests.helpers import load_bhx

pytestmark = pytest.mark.gpu

FAST = dict(theta=0.5, eps=0.01, dt=5e-4, leaf_capacity=32, mode=1, precision=1, quadrupole=1, compensated=1,
            group_walk=1)


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


def _pair(bhx, n, cfg, pos, vel, mass, off=None):
    a = bhx.Engine(n, **cfg)
    b = bhx.Engine(n, **cfg)
    for e in (a, b):
        e.set_bodies(pos, vel, mass)
        if off is not None:
            e.set_systems(off)
    return a, b


def _same(a, b):
    A, B = a.get_bodies(), b.get_bodies()
    return all(np.array_equal(x, y) for x, y in zip(A, B))


@pytest.mark.parametrize("cfg", [FAST, dict(theta=0.5, eps=0.01, dt=1e-3)], ids=["fp32_group", "as_run_fp64"])
def test_graph_replays_equal_eager_steps(bhx, cfg):
    n = 20000
    pos, vel, mass = bhx.plummer(n, seed=3)
    print("graph vs eager n=%d cfg=%s" % (n, cfg), flush=True)
    eager, graph = _pair(bhx, n, cfg, pos, vel, mass)
    with eager, graph:
        for k in range(5):
            eager.step(1)
            graph.step_graph(1)
            assert _same(eager, graph), "replay %d" % k
            eager.sync()
        eager.step(4)
        graph.step_graph(4)
        assert _same(eager, graph)
        # a configuration change drops the captured graphs
        eager.set_config(theta=0.7)
        graph.set_config(theta=0.7)
        eager.step(3)
        graph.step_graph(3)
        assert _same(eager, graph)
        # and eager frames in between keep both in step
        graph.step(1)
        eager.step(1)
        graph.step_graph(2)
        eager.step(2)
        assert _same(eager, graph)
        assert np.array_equal(np.sort(eager.nodes(), order="key")["key"], np.sort(graph.nodes(), order="key")["key"])


def test_graph_ensemble_equals_eager(bhx):
    sizes = [5000, 64, 1, 3000, 777]
    parts = [bhx.plummer(n, seed=20 + k) for k, n in enumerate(sizes)]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    eager, graph = _pair(bhx, pos.shape[0], FAST, pos, vel, mass, off)
    with eager, graph:
        for k in range(3):
            eager.step(2)
            graph.step_graph(2)
            assert _same(eager, graph), "replay %d" % k


def test_graph_restarts_after_node_overflow(bhx):
    """Close pairs force long single-child chains: far more nodes than the
    context's initial arrays hold, so the first graph run overflows, restarts
    from its snapshot with larger arrays and must still match the eager run."""
    rng = np.random.default_rng(9)
    half = 3000
    base = rng.uniform(-1.0, 1.0, (half, 3))
    pos = np.concatenate([base, base + rng.uniform(-1e-6, 1e-6, (half, 3))])
    vel = rng.normal(0.0, 0.01, pos.shape)
    mass = np.full(pos.shape[0], 1.0 / pos.shape[0])
    cfg = dict(theta=0.5, eps=0.01, dt=1e-4)
    eager, graph = _pair(bhx, pos.shape[0], cfg, pos, vel, mass)
    with eager, graph:
        eager.step(3)
        graph.step_graph(3)
        assert len(graph.nodes()) > 2 * pos.shape[0] + 1024
        assert _same(eager, graph)
