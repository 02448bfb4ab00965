"""Engine-level snapshots, visualisation export, tree energies, and the
galaxy-pair workload (4,194,304 bodies: a 3:1 pair of disk + Hernquist-halo
galaxies, FP32 walk, theta 0.6, eps 0.02, bucket 32, dt 1e-3) checked through
size-independent properties."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import load_bhx
from tests.test_ic_io import _read_vtk

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


FAST = dict(theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=16, mode=1, precision=1, quadrupole=1, compensated=1,
            permute_acc=1, group_walk=1)


def test_save_load_round_trip_continues_bitwise(bhx, tmp_path):
    sizes = [700, 1300, 5]
    parts = [bhx.plummer(k, seed=50 + i) for i, k in enumerate(sizes)]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    path = str(tmp_path / "s.bhx")
    print("save/load n=%d systems=%d" % (pos.shape[0], len(sizes)), flush=True)
    with bhx.Engine(pos.shape[0], **FAST) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        e.step(3)
        e.save(path, time=0.003, step=3)
        before = e.get_bodies()
        comp = e.get_compensation()
        assert np.any(comp[0] != 0)   # the compensated leapfrog carries rounding terms
        e.step(4)
        ref = e.get_bodies()
    d = bhx.read_snapshot(path)
    assert d["step"] == 3 and d["time"] == 0.003 and np.array_equal(d["offsets"], off)
    assert all(np.array_equal(x, y) for x, y in zip((d["pos"], d["vel"], d["mass"], d["acc"]), before))
    assert np.array_equal(d["comp"][0], comp[0]) and np.array_equal(d["comp"][1], comp[1])
    with bhx.Engine(pos.shape[0], **FAST) as r:
        t, st = r.load(path)
        assert (t, st) == (0.003, 3)
        assert r.nsys == 3 and np.array_equal(r.get_systems(), off)
        assert all(np.array_equal(x, y) for x, y in zip(r.get_bodies(), before))
        r.step(4)
        # the bodies, accelerations and systems are the whole frame state
        assert all(np.array_equal(x, y) for x, y in zip(r.get_bodies(), ref))
    with bhx.Engine(10, **FAST) as small, pytest.raises(bhx.BhxError):
        small.load(path)   # more bodies than the context holds


def test_export_vtk_from_engine(bhx, tmp_path):
    n = 3000
    pos, vel, mass = bhx.plummer(n, seed=4)
    with bhx.Engine(n, **FAST) as e:
        e.set_bodies(pos, vel, mass)
        c, s = e.bounds()
        e.build(c, s)
        nodes = e.nodes()
        bp, np_ = str(tmp_path / "b.vtk"), str(tmp_path / "n.vtk")
        e.export_vtk(bp, np_, min_level=0, max_level=3)
    b = _read_vtk(bp)
    assert np.array_equal(b["points"], pos.astype(np.float32))
    assert np.array_equal(b["mass"], mass.astype(np.float32))
    t = _read_vtk(np_)
    sel = nodes[nodes["level"] <= 3]
    assert t["cell_types"].size == sel.size
    assert np.array_equal(t["level"], sel["level"]) and np.array_equal(t["count"], sel["n"])
    # the root box: centre +- size/2 (the box the reference draws)
    r = nodes[0]
    assert np.allclose(t["points"][0], (r["center"] - 0.5 * r["size"]).astype(np.float32))


@pytest.mark.parametrize("prec", [0, 1])
def test_energy_theta0_equals_direct_sum(bhx, prec):
    n = 2048
    pos, vel, mass = bhx.plummer(n, seed=8)
    print("energy theta=0 prec=%d" % prec, flush=True)
    with bhx.Engine(n, **dict(FAST, theta=0.0, quadrupole=0, precision=prec)) as e:
        e.set_bodies(pos, vel, mass)
        K, W, vir, phi = e.energy(potentials=True)
        acc_before = e.get_bodies()[3]
        assert np.all(acc_before == 0)   # accelerations untouched
    _, pot = orc.direct(pos, mass, 0.01)
    # theta = 0 opens every cell: the group walk is the direct sum (FP64: rounding
    # order only; FP32: single-precision terms)
    tol = 1e-12 if prec == 0 else 1e-5
    assert np.abs(phi - pot).max() <= tol * np.abs(pot).max()
    Wref = 0.5 * (mass * pot).sum()
    Kref = 0.5 * (mass * (vel ** 2).sum(1)).sum()
    assert abs(W - Wref) <= tol * abs(Wref)
    assert abs(K - Kref) <= 1e-13 * Kref
    assert abs(vir - 2 * Kref / abs(Wref)) <= 10 * tol


def test_energy_tree_accuracy(bhx):
    n = 20000
    pos, vel, mass = bhx.plummer(n, seed=9)
    with bhx.Engine(n, **FAST) as e:
        e.set_bodies(pos, vel, mass)
        K, W, vir, phi = e.energy(potentials=True)
    tg = np.arange(0, n, 10)
    _, pot = orc.direct(pos, mass, 0.01, targets=tg)
    # theta 0.5 with quadrupoles, FP32 walk: per-body potentials against the FP64 direct sum
    rel = np.abs(phi[tg] - pot) / np.abs(pot)
    assert np.median(rel) < 1e-4 and rel.max() < 2e-3
    # W is the FP64 sum of m phi / 2 over the returned potentials
    assert abs(W - 0.5 * (mass * phi).sum()) <= 1e-12 * abs(W)
    # an equilibrium Plummer sphere: virial ratio near 1
    assert abs(vir - 1) < 0.05


N_C3 = 4_194_304
# enclosing-cube cells: the secondary galaxy's centre lies beyond half the bounds
# size, outside every cell of the reference geometry
C3 = dict(theta=0.6, eps=0.02, dt=1e-3, leaf_capacity=32, mode=1, precision=1, quadrupole=1, compensated=1,
          permute_acc=1, group_walk=1, geometry=1)


def test_galaxy_pair_4m_steps(bhx):
    print("galaxy pair n=%d" % N_C3, flush=True)
    pos, vel, mass = bhx.galaxy_pair(N_C3, seed=2026, threads=16)
    p0 = (mass[:, None] * vel).sum(0)
    with bhx.Engine(N_C3, **C3) as e:
        e.set_bodies(pos, vel, mass)
        e.step(3)
        p, v, m, a = e.get_bodies()
        assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
        # bodies are permuted, never lost or duplicated
        assert np.array_equal(np.sort(m), np.sort(mass))
        assert np.isclose(m.sum(), mass.sum(), rtol=1e-14)
        # linear momentum: the tree forces are not pairwise antisymmetric, so it
        # drifts by the force error only
        p1 = (m[:, None] * v).sum(0)
        scale = (m * np.linalg.norm(v, axis=1)).sum()
        assert np.abs(p1 - p0).max() < 1e-5 * scale
        # the tree of the current bodies: root holds every body, children partition it
        c, s = e.bounds()
        e.build(c, s)
        nodes = e.nodes()
        assert nodes["n"][0] == N_C3
        lvl1 = nodes[nodes["level"] == 1]
        assert lvl1["n"].sum() == N_C3
        assert np.isclose(lvl1["mass"].sum(), m.sum(), rtol=1e-12)
        # tree order is sorted by path key
        tp, tk = e.tree_order()
        assert np.all(tk[1:] >= tk[:-1])
        assert np.array_equal(np.sort(tp), np.arange(N_C3, dtype=tp.dtype))
        # force accuracy on a seeded 1,024-body sample against the FP64 direct sum
        acc = e.walk()
        rng = np.random.default_rng(7)
        tg = np.sort(rng.choice(N_C3, 1024, replace=False))
        dacc, _ = e.direct(tg, C3["eps"])
        err = np.linalg.norm(acc[tg] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
        print("galaxy pair median force error %.3e p99 %.3e" % (np.median(err), np.quantile(err, 0.99)), flush=True)
        assert np.median(err) < 1e-3
        assert np.quantile(err, 0.99) < 1e-2
        # the reference geometry on the same bodies: cells that do not contain their
        # bodies pass the opening test too early
        e.set_config(geometry=0)
        c, s = e.bounds()
        e.build(c, s)
        acc0 = e.walk()
        err0 = np.linalg.norm(acc0[tg] - dacc, axis=1) / np.linalg.norm(dacc, axis=1)
        print("reference geometry median %.3e p99 %.3e" % (np.median(err0), np.quantile(err0, 0.99)), flush=True)
        assert np.quantile(err0, 0.99) > np.quantile(err, 0.99)
