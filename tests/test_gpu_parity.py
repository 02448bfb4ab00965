"""GPU parity: the HIP engine, called through its C ABI, against the CPU oracle
on the same seeded inputs.

Bit-exact: bounds, Morton keys, sort permutation, tree path keys, node set and
as-run node payload (equal power-of-two masses make every mass sum exact).
Tolerances (written at each check): accelerations, because the GPU sums the
same interaction terms in a different order.
"""
import zlib

import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import load_bhx
from tests.kats import DRIFT_KAT, KICK_KAT, MEAN_KAT_C, MEAN_KAT_POS, TINY

pytestmark = pytest.mark.gpu

N1 = 16384   # Plummer N of the CPU-runnable configuration


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


@pytest.fixture(scope="module")
def plummer16k(bhx):
    return bhx.plummer(N1, seed=1)


def _engine(bhx, n, **cfg):
    print("engine n=%d cfg=%s" % (n, cfg), flush=True)
    return bhx.Engine(n, **cfg)


def test_bounds_and_keys_bitexact(bhx, plummer16k):
    pos, vel, mass = plummer16k
    with _engine(bhx, N1) as e:
        e.set_bodies(pos, vel, mass)
        c, s = e.bounds()
        oc, os_ = orc.bounds(pos)
        assert np.array_equal(c, oc) and s == os_
        keys = e.morton_keys(s)
        assert np.array_equal(keys, orc.morton_keys(pos, s))


def _mean_case(bhx, name):
    """Positions whose left-to-right sums are hard for a parallel evaluation:
    random order (partial sums near zero: every tile added one by one), large
    offsets, steady drift across many binades, halfway cases, sums hovering at
    a binade edge, huge and subnormal magnitudes, non-finite values, ragged
    sizes, and the Morton-ordered Plummer sphere the step sees."""
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if name.startswith("size"):
        n = int(name[4:])
        return bhx.plummer(n, seed=n)[0] if n > 1 else np.array([[0.3, -0.2, 0.1]])
    n = 100003
    if name == "random_order":
        return rng.permutation(bhx.plummer(n, seed=2)[0])
    if name == "morton_1m":
        pos = bhx.plummer(1 << 20, seed=1)[0]
        _, s = orc.bounds(pos)
        return pos[np.argsort(orc.morton_keys(pos, s), kind="stable")]
    if name == "offset":
        return rng.standard_normal((n, 3)) + 1e8
    if name == "drift":
        return 1.0 + rng.uniform(-1, 1, (300000, 3))
    if name == "halves":
        return rng.integers(-1000, 1001, (n, 3)) * 0.25
    if name == "hover":
        x = np.empty((n, 3))
        for a in range(3):
            s = 0.0
            for k in range(n):
                v = 1024.0 if k == 0 else (1024.0 - s + 1e-9 * rng.uniform(-1, 1) if k % 2 else
                                           1e-7 * rng.uniform(-1, 1))
                x[k, a] = v
                s = s + v
        return x
    if name == "huge":
        return 1e300 * rng.uniform(-1, 1, (n, 3))
    if name == "subnormal":
        return 1e-310 * rng.uniform(-1, 1, (n, 3))
    if name == "up_down":
        sgn = np.where(np.arange(n) < n // 2, 1.0, -1.0)[:, None]
        return sgn * (1.0 + 1e-3 * rng.uniform(-1, 1, (n, 3)))
    if name == "nonfinite":
        x = rng.standard_normal((n, 3))
        x[5000, 0], x[70000, 1] = np.nan, np.inf
        return x
    raise ValueError(name)


@pytest.mark.parametrize("name", ["size1", "size2", "size1023", "size1024", "size1025", "size2049", "random_order",
                                  "morton_1m", "offset", "drift", "halves", "hover", "huge", "subnormal", "up_down",
                                  "nonfinite"])
def test_bounds_sequential_mean_bitexact(bhx, name):
    """The one-system centre is the left-to-right FP64 sum (OctantUtils.cpp:31-38)
    bit for bit, however the tiles split it into runs."""
    pos = _mean_case(bhx, name)
    n = pos.shape[0]
    with _engine(bhx, n) as e:
        e.set_bodies(pos, None, np.ones(n))
        c, s = e.bounds()
    oc, os_ = orc.bounds(pos)
    assert c.tobytes() == oc.tobytes() and np.float64(s).tobytes() == np.float64(os_).tobytes()


def test_keys_out_of_domain_x86_semantics(bhx):
    # bodies outside [-S, S] (negative and wrapped scaled coordinates)
    pos = np.array([[-1.5, 0, 0], [1.0, 0, 0], [0.5, -0.5, 0], [-1, -1, -1], [3.7, -2.2, 9.0]], float)
    with _engine(bhx, 8) as e:
        e.set_bodies(pos, None, np.ones(5))
        keys = e.morton_keys(1.0)
    assert [int(k) for k in keys] == [orc.morton_key(p, 1.0) for p in pos]
    assert int(keys[0]) == 0xF924000000000000 and int(keys[1]) == 0xB000000000000000


@pytest.mark.parametrize("n", [2, 1000, 4096, 4097, 100003])
def test_sort_keys_stable(bhx, n):
    rng = np.random.default_rng(n)
    base = rng.integers(0, 2 ** 63, 97, dtype=np.int64).astype(np.uint64) | np.uint64(1 << 63)
    keys = base[rng.integers(0, 97, n)]
    with _engine(bhx, n) as e:
        sk, idx = e.sort_keys(keys)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(idx.astype(np.int64), ref)
    assert np.array_equal(sk, keys[ref])


@pytest.mark.parametrize("case", ["narrow", "equal", "top_only"])
def test_sort_keys_constant_digits(bhx, case):
    """Digits that every key shares make their passes plain copies: the result is still the
    stable sort (numpy) — keys within a narrow range (the high bytes agree), all keys equal
    (every pass a copy), and keys that differ only in the top byte."""
    n = 70001
    rng = np.random.default_rng(7)
    if case == "narrow":
        keys = (np.uint64(1 << 63) | np.uint64(0x123456) << np.uint64(24)) + rng.integers(0, 1 << 20, n).astype(np.uint64)
    elif case == "equal":
        keys = np.full(n, np.uint64(0x8000000000000abc))
    else:
        keys = (rng.integers(0, 256, n).astype(np.uint64) << np.uint64(56)) | np.uint64(0x00ffffffffffff00)
    with _engine(bhx, n) as e:
        sk, idx = e.sort_keys(keys)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(idx.astype(np.int64), ref)
    assert np.array_equal(sk, keys[ref])


def test_sort_bodies_matches_reference_radix(bhx, plummer16k):
    pos, vel, mass = plummer16k
    with _engine(bhx, N1) as e:
        e.set_bodies(pos, vel, mass)
        _, s = e.bounds()
        e.morton_keys(s)
        perm = e.sort_bodies()
        gp, gv, gm, _ = e.get_bodies()
    keys = orc.morton_keys(pos, s)
    _, idx = orc.radix_sort(keys)
    assert np.array_equal(perm, idx)
    assert np.array_equal(gp, pos[idx.astype(np.int64)]) and np.array_equal(gv, vel[idx.astype(np.int64)])


def _compare_nodes(g, o, fields):
    g = np.sort(g, order="key")
    o = np.sort(o, order="key")
    assert len(g) == len(o)
    for f in fields:
        assert np.array_equal(g[f], o[f]), f


@pytest.mark.parametrize("n,seed", [(1, 3), (2, 4), (17, 5), (N1, 1)])
def test_build_as_run_node_set_bitexact(bhx, n, seed):
    pos, vel, mass = bhx.plummer(n, seed=seed)
    c, s = orc.bounds(pos)
    with _engine(bhx, n) as e:
        e.set_bodies(pos, vel, mass)
        nn = e.build(c, s)
        g = e.nodes()
        tperm, tkeys = e.tree_order()
        unsplit = e.unsplit_leaves()
    t = orc.build_insert(pos, mass, c, s)
    o = t.nodes()
    assert nn == len(o) and unsplit == t.unsplit
    _compare_nodes(g, o, ["key", "center", "size", "n", "level", "child_byte", "first", "com", "mass"])
    pk = orc.tree_path_keys(pos, c, s)
    assert np.array_equal(tkeys, np.sort(pk))
    assert np.array_equal(tperm.astype(np.int64), np.argsort(pk, kind="stable"))


@pytest.mark.parametrize("cap", [4, 32])
def test_build_bucket_node_set_bitexact(bhx, plummer16k, cap):
    pos, vel, mass = plummer16k
    c, s = orc.bounds(pos)
    with _engine(bhx, N1, leaf_capacity=cap) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        g = e.nodes()
    o = orc.build_recursive(pos, mass, c, s, cap).nodes()
    _compare_nodes(g, o, ["key", "center", "size", "n", "level", "child_byte", "first", "com", "mass"])


@pytest.mark.parametrize("cap", [1, 32])
def test_build_enclosing_geometry(bhx, plummer16k, cap):
    """geometry = 1 is the reference recurrence from a root edge of twice the bounds
    size: the node set equals the oracle's tree built with size 2S, and every body
    lies inside its leaf's cube (centre +- edge/2)."""
    pos, vel, mass = plummer16k
    c, s = orc.bounds(pos)
    with _engine(bhx, N1, leaf_capacity=cap, geometry=1) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        g = e.nodes()
        tperm, _ = e.tree_order()
    t = orc.build_insert(pos, mass, c, 2 * s) if cap == 1 else orc.build_recursive(pos, mass, c, 2 * s, cap)
    _compare_nodes(g, t.nodes(), ["key", "center", "size", "n", "level", "child_byte", "first", "com", "mass"])
    assert g["size"][0] == 2 * s
    leaves = g[g["child_byte"] == 0]
    tp = tperm.astype(np.int64)
    for lf in leaves[:: max(1, leaves.size // 2000)]:
        b = pos[tp[lf["body_start"]:lf["body_start"] + lf["n"]]]
        # (centres are rounded FP64 sums: allow their last bits)
        assert np.all(np.abs(b - lf["center"]) <= 0.5 * lf["size"] * (1 + 1e-12))
    # the reference geometry does not enclose its outer bodies (its root edge is S)
    with _engine(bhx, N1, leaf_capacity=cap) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        g0 = e.nodes()
        tp0 = e.tree_order()[0].astype(np.int64)
    outside = 0
    for lf in g0[g0["child_byte"] == 0]:
        b = pos[tp0[lf["body_start"]:lf["body_start"] + lf["n"]]]
        outside += int(np.any(np.abs(b - lf["center"]) > 0.5 * lf["size"]))
    assert outside > 0


def test_preorder_skip_structure(bhx, plummer16k):
    pos, vel, mass = plummer16k
    c, s = orc.bounds(pos)
    with _engine(bhx, N1) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        g = e.nodes()
    n = len(g)
    assert g["skip"][0] == n and g["level"][0] == 0
    # subtree [p, skip) = exactly the nodes whose key has p's key as prefix
    for p in range(0, n, max(1, n // 500)):
        k, L, sk = int(g["key"][p]), int(g["level"][p]), int(g["skip"][p])
        sub = g[p:sk]
        sh = (sub["level"] - L) * 3
        assert np.all((sub["key"] >> sh.astype(np.uint64)) == k)
        if sk < n:
            kk = int(g["key"][sk]) >> (3 * (int(g["level"][sk]) - L)) if g["level"][sk] >= L else None
            assert kk != k


def test_lookup_hash_table(bhx, plummer16k):
    pos, vel, mass = plummer16k
    c, s = orc.bounds(pos)
    with _engine(bhx, N1) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        g = e.nodes()
        ids = e.lookup(g["key"])
        assert np.array_equal(ids, np.arange(len(g)))
        missing = np.array([0, 2, (int(g["key"][1]) << 3) | 7, 1 << 62], np.uint64)
        present = set(int(k) for k in g["key"])
        exp = [-1 if int(k) not in present else None for k in missing]
        got = e.lookup(missing)
        for a, b in zip(got, exp):
            if b is not None:
                assert a == b


def test_walk_as_run_fp64_matches_reference_lists(bhx, plummer16k):
    pos, vel, mass = plummer16k
    c, s = orc.bounds(pos)
    with _engine(bhx, N1, theta=0.5, eps=0.01) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        acc, cnt = e.walk(counts=True)
    t = orc.build_insert(pos, mass, c, s)
    oacc, ocnt = t.walk(0.5, 0.01)
    assert np.array_equal(cnt.astype(np.int64), ocnt)          # same interaction-list lengths
    err = np.linalg.norm(acc - oacc, axis=1) / np.linalg.norm(oacc, axis=1)
    assert err.max() < 1e-12                                    # same terms, other summation order


def test_walk_theta_zero_is_direct_sum(bhx):
    pos, vel, mass = bhx.plummer(3000, seed=9)
    c, s = orc.bounds(pos)
    with _engine(bhx, 3000, theta=0.0, eps=0.01) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        acc, cnt = e.walk(counts=True)
        dacc, _ = e.direct(np.arange(3000), 0.01)
    ref, _ = orc.direct(pos, mass, 0.01)
    assert np.all(cnt == 2999)
    assert (np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)).max() < 1e-12
    assert (np.linalg.norm(dacc - ref, axis=1) / np.linalg.norm(ref, axis=1)).max() < 1e-12


@pytest.mark.parametrize("prec", [0, 1])
def test_walk_intended_quadrupole_accuracy(bhx, plummer16k, prec):
    pos, vel, mass = plummer16k
    c, s = orc.bounds(pos)
    with _engine(bhx, N1, theta=0.5, eps=0.01, mode=1, quadrupole=1, leaf_capacity=32, precision=prec) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        acc = e.walk()
    t = orc.build_recursive(pos, mass, c, s, 32)
    t.moments_intended()
    oacc, _ = t.walk(0.5, 0.01, quad=True)
    ref, _ = orc.direct(pos, mass, 0.01, targets=np.arange(0, N1, 8))
    sub = acc[::8]
    e_bh = np.linalg.norm(sub - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(e_bh) < 1e-3                                # BH + precision error vs FP64 direct sum
    e_or = np.linalg.norm(acc - oacc, axis=1) / np.linalg.norm(oacc, axis=1)
    tol = 1e-11 if prec == 0 else 1e-4                          # FP64: summation order; FP32: rounding
    assert np.median(e_or) < tol


@pytest.mark.parametrize("permute_acc", [1, 0])
def test_step_matches_reference_frames(bhx, permute_acc):
    n = 4096
    pos, vel, mass = bhx.plummer(n, seed=21)
    dt, theta, eps = 1e-3, 0.5, 0.01
    with _engine(bhx, n, theta=theta, eps=eps, dt=dt, permute_acc=permute_acc) as e:
        e.set_bodies(pos, vel, mass)
        e.step(3)
        gp, gv, gm, ga = e.get_bodies()
    p, v, m, a = pos, vel, mass, np.zeros_like(pos)
    for _ in range(3):
        p, v, m, a, keys, st = orc.frame(p, v, m, a, dt, theta, eps, permute_acc=bool(permute_acc))
    assert np.array_equal(gm, m)
    assert np.allclose(gp, p, rtol=0, atol=1e-13) and np.allclose(gv, v, rtol=0, atol=1e-12)
    assert np.allclose(ga, a, rtol=1e-10, atol=1e-12)


def test_kick_drift_kick_bitexact(bhx):
    pos, vel, mass = bhx.plummer(1000, seed=2)
    acc = np.random.default_rng(0).normal(size=(1000, 3))
    with _engine(bhx, 1000) as e:
        e.set_bodies(pos, vel, mass)
        e.set_accelerations(acc)
        e.kick_drift(1e-3)
        e.kick(1e-3)
        gp, gv, _, _ = e.get_bodies()
    p1, v1 = orc.kick_drift(1e-3, pos, vel, acc)
    v2 = orc.kick(1e-3, v1, acc)
    assert np.array_equal(gp, p1) and np.array_equal(gv, v2)


def test_rounding_order_kats(bhx):
    """The hand-derived known answers that fix the reference's rounding order
    (centre = sum * fl(1/N), kick (a dt) 0.5, drift x + fl(v dt), no fused
    multiply-adds, subnormal products kept), on the engine."""
    with _engine(bhx, 8) as e:
        e.set_bodies(MEAN_KAT_POS, None, np.ones(3))
        c, s = e.bounds()
        assert list(c) == MEAN_KAT_C and s == 4.0 - MEAN_KAT_C[2]
        e.set_bodies(np.zeros((1, 3)), KICK_KAT["v"], np.ones(1))
        e.set_accelerations(KICK_KAT["a"])
        e.kick(1e-3)
        assert e.get_bodies()[1][0, 0] == KICK_KAT["want"]
        e.set_bodies(DRIFT_KAT["x"], DRIFT_KAT["v"], np.ones(1))
        e.kick_drift(1e-3)
        assert e.get_bodies()[0][0, 0] == DRIFT_KAT["want"]
        e.set_bodies(np.zeros((1, 3)), None, np.ones(1))
        e.set_accelerations([[TINY, 0.0, 0.0]])
        e.kick_drift(2.75)
        p, v, _, _ = e.get_bodies()
        assert v[0, 0] == 2 * TINY and p[0, 0] == 6 * TINY


def test_empty_and_errors(bhx):
    with _engine(bhx, 16) as e:
        e.set_bodies(np.zeros((0, 3)), None, np.zeros(0))
        assert e.build(np.zeros(3), 1.0) == 0
        e.step(1)
        with pytest.raises(bhx.BhxError):
            e.set_bodies(np.zeros((17, 3)), None, np.ones(17))     # over capacity
        with pytest.raises(bhx.BhxError):
            e.build(np.array([np.nan, 0, 0]), 1.0)


@pytest.mark.parametrize("name", ["as_run_fp64", "group_fp32_quad", "adjoint_fp32"])
def test_nonfinite_state_is_reported(bhx, name):
    """A NaN velocity or an infinite mass makes the stepping calls (eager, graph,
    recorded) fail with BHX_E_NONFINITE (every loop in a frame is bounded by
    integer counts, so the frames still end); once reported the flag is clear,
    and finite bodies set afterwards step exactly as in a fresh engine."""
    adjoint = dict(theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=32, mode=1, precision=1, group_walk=1,
                   permute_acc=1)
    cfg = adjoint if name == "adjoint_fp32" else TINY_CFGS[name]
    n = 4096
    pos, vel, mass = bhx.plummer(n, seed=7)
    bad_v = vel.copy()
    bad_v[100, 1] = np.nan
    bad_m = mass.copy()
    bad_m[5] = np.inf
    with _engine(bhx, n, **cfg) as e:
        for v, m in ((bad_v, mass), (vel, bad_m)):
            calls = [("step", e.step), ("step_graph", e.step_graph)]
            if cfg is adjoint:
                calls.append(("record", lambda k: e.record(k, checkpoint_every=1)))
            for label, call in calls:
                print(name, label, flush=True)
                e.set_bodies(pos, v, m)
                with pytest.raises(bhx.BhxError) as err:
                    call(2)
                assert err.value.code == 5
        e.set_bodies(pos, vel, mass)
        e.step(2)
        e.step_graph(1)
        p1, v1, _, a1 = e.get_bodies()
    with _engine(bhx, n, **cfg) as f:
        f.set_bodies(pos, vel, mass)
        f.step(3)
        p2, v2, _, a2 = f.get_bodies()
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2) and np.array_equal(a1, a2)


TINY_CFGS = {
    "as_run_fp64": dict(theta=0.5, eps=0.01, dt=1e-3),
    "group_fp32_quad": dict(theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=32, mode=1, precision=1, quadrupole=1,
                            group_walk=1, compensated=1),
    "group_fp64": dict(theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=4, mode=1, precision=0, group_walk=1),
    "body_fp32": dict(theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=4, mode=1, precision=1, group_walk=0),
}


def test_clustered_bodies_depth_cap_and_big_leaves(bhx):
    """A Plummer sphere plus 200 bodies at one point and 100 within 1e-9 of another:
    the tree reaches the 21-level cap with multi-body leaves (the oracle's node set,
    bit for bit), a leaf of more than 64 bodies is queued in chunks, and both group
    walks at theta = 0 equal the direct sum."""
    base, _, _ = bhx.plummer(3000, seed=41)
    rng = np.random.default_rng(42)
    pos = np.concatenate([base, np.tile([[0.25, -0.5, 0.125]], (200, 1)),
                          np.array([-0.3, 0.2, 0.4]) + 1e-9 * rng.standard_normal((100, 3))])
    n = pos.shape[0]
    mass = np.full(n, 1.0 / 4096)
    c, s = orc.bounds(pos)
    with _engine(bhx, n, leaf_capacity=8) as e:
        e.set_bodies(pos, np.zeros_like(pos), mass)
        nn = e.build(c, s)
        g = e.nodes()
        unsplit = e.unsplit_leaves()
    t = orc.build_recursive(pos, mass, c, s, 8)
    o = t.nodes()
    assert g["level"].max() == 21 and unsplit == t.unsplit and unsplit >= 1
    assert nn == len(o)
    _compare_nodes(g, o, ["key", "center", "size", "n", "level", "child_byte", "first", "com", "mass"])
    ref, _ = orc.direct(pos, mass, 0.01)
    for prec in (0, 1):
        with _engine(bhx, n, theta=0.0, eps=0.01, group_walk=1, precision=prec, leaf_capacity=8, mode=1,
                     quadrupole=1) as e:
            e.set_bodies(pos, np.zeros_like(pos), mass)
            e.build(c, s)
            acc = e.walk()
            e.set_config(theta=0.5)
            acc5 = e.walk()
        err = np.linalg.norm(acc - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
        print("clustered prec %d: theta 0 max %.3e median %.3e" % (prec, err.max(), np.median(err)), flush=True)
        # FP32: offsets from a group's first body carry ~1e-7 of the group's extent, and
        # next to the 200-body point mass that is a large part of a body's net force
        # (measured: max 2.0e-4, typical 3e-6 to 6e-6)
        assert err.max() < (1e-12 if prec == 0 else 5e-4)
        assert np.median(err) < (1e-13 if prec == 0 else 1e-5)
        assert np.all(np.isfinite(acc5))


@pytest.mark.parametrize("name", sorted(TINY_CFGS))
def test_tiny_systems(bhx, name):
    """One body feels nothing; two bodies follow the softened two-body force (equal and
    opposite); two coincident bodies add nothing to each other, softened or not."""
    cfg = TINY_CFGS[name]
    eps = cfg["eps"]
    with _engine(bhx, 4, **cfg) as e:
        e.set_bodies(np.array([[0.3, -0.2, 0.1]]), np.array([[0.1, 0.0, 0.0]]), np.array([1.0]))
        e.step(3)
        p, v, m, a = e.get_bodies()
        assert np.all(a == 0) and np.allclose(p, [[0.3 + 3e-4, -0.2, 0.1]], rtol=0, atol=1e-15)
        x = np.array([[0.0, 0.0, 0.0], [0.7, -0.4, 0.2]])
        m2 = np.array([0.75, 0.25])
        e.set_bodies(x, np.zeros_like(x), m2)
        c, sz = e.bounds()
        e.build(c, sz)
        acc = e.walk()
        d = x[1] - x[0]
        f = d / (d @ d + eps * eps) ** 1.5
        tol = 1e-14 if cfg.get("precision", 0) == 0 else 1e-6
        assert np.allclose(acc[0], m2[1] * f, rtol=tol, atol=0) and np.allclose(acc[1], -m2[0] * f, rtol=tol, atol=0)
        e.step(5)
        assert all(np.all(np.isfinite(q)) for q in e.get_bodies())
        for soft in (eps, 0.0):
            e.set_config(eps=soft)
            e.set_bodies(np.array([[0.1, 0.1, 0.1]] * 2), np.zeros((2, 3)), np.array([0.5, 0.5]))
            c, sz = e.bounds()
            e.build(c, sz)
            acc = e.walk()
            assert np.all(acc == 0), (soft, acc)


def test_diagnostics_momentum(bhx, plummer16k):
    pos, vel, mass = plummer16k
    with _engine(bhx, N1) as e:
        e.set_bodies(pos, vel, mass)
        ek, mom, am = e.diagnostics()
    assert ek == pytest.approx(0.5 * (mass * (vel ** 2).sum(1)).sum(), rel=1e-12)
    assert np.allclose(mom, (mass[:, None] * vel).sum(0), atol=1e-14)
    assert np.allclose(am, (mass[:, None] * np.cross(pos, vel)).sum(0), atol=1e-13)


# ---------------------------------------------------------------- group walk (shared lists)

@pytest.mark.parametrize("prec", [0, 1])
def test_group_walk_theta_zero_is_direct_sum(bhx, prec):
    pos, vel, mass = bhx.plummer(3000, seed=19)
    c, s = orc.bounds(pos)
    with _engine(bhx, 3000, theta=0.0, eps=0.01, group_walk=1, precision=prec, leaf_capacity=8, mode=1) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        acc = e.walk()
    ref, _ = orc.direct(pos, mass, 0.01)
    err = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < (1e-12 if prec == 0 else 1e-4)      # FP64: summation order; FP32: rounding


@pytest.mark.parametrize("group", [1, 0])
@pytest.mark.parametrize("quad", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
def test_walk_unsoftened(bhx, prec, quad, group):
    """eps = 0 takes the walks' guarded form (a body meets itself at d = 0 and adds
    nothing): theta = 0 equals the unsoftened direct sum, and theta = 0.5 on the same
    tree stays as accurate as the softened walk (group and per-body walks)."""
    n = 3000
    pos, vel, mass = bhx.plummer(n, seed=29)
    c, s = orc.bounds(pos)
    with _engine(bhx, n, theta=0.0, eps=0.0, group_walk=group, precision=prec, leaf_capacity=8, mode=1,
                 quadrupole=quad) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        acc0 = e.walk()
        e.set_config(theta=0.5)
        acc5 = e.walk()
    ref, _ = orc.direct(pos, mass, 0.0)
    e0 = np.linalg.norm(acc0 - ref, axis=1) / np.linalg.norm(ref, axis=1)
    e5 = np.linalg.norm(acc5 - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print("eps 0: theta 0 max %.3e, theta 0.5 median %.3e" % (e0.max(), np.median(e5)), flush=True)
    assert np.all(np.isfinite(acc0)) and np.all(np.isfinite(acc5))
    assert e0.max() < (1e-12 if prec == 0 else 1e-4)      # FP64: summation order; FP32: rounding
    # the per-body walk is the reference's own criterion, coarser than the shared
    # group lists; smaller groups share less (measured at theta 0.5, monopole:
    # per body 2.0e-3, 16-body lists 1.04e-3, 64-body lists 6.3e-4)
    assert np.median(e5) < (1.5e-3 if group else 5e-3)


@pytest.mark.parametrize("scale_exp", [30, -30])
@pytest.mark.parametrize("prec", [0, 1])
def test_walk_scale_equivariance(bhx, prec, scale_exp):
    """Positions and softening times s = 2^k scale every FP64 operation of bounds,
    keys and tree exactly, so the tree is the same and accelerations scale by 1/s^2."""
    n = 4000
    pos, vel, mass = bhx.plummer(n, seed=31)
    s = 2.0 ** scale_exp
    acc = []
    for f in (1.0, s):
        with _engine(bhx, n, theta=0.5, eps=0.01 * f, group_walk=1, precision=prec, leaf_capacity=16, mode=1,
                     quadrupole=1) as e:
            e.set_bodies(pos * f, vel, mass)
            c, sz = e.bounds()
            e.build(c, sz)
            acc.append(e.walk() * f * f)
    d = np.abs(acc[1] - acc[0]).max() / np.abs(acc[0]).max()
    print("scale 2^%d prec %d: max relative difference %.3e" % (scale_exp, prec, d), flush=True)
    # FP64 scales exactly; the FP32 walk works in per-system power-of-two units,
    # so it scales exactly too (without them r^-5 and r^-7 leave the FP32 range)
    assert d == 0.0


def test_fp32_walk_in_si_units(bhx):
    """Solar-system units (lengths ~1e11 m, masses ~1e30 kg, G = 1): the FP32 group
    walk stays finite and agrees with the FP64 walk on the same tree."""
    n, L, M = 4000, 1.496e11, 1.989e30
    pos, vel, mass = bhx.plummer(n, seed=37)
    res = []
    for prec in (0, 1):
        # the same 64-body lists in both precisions: the difference is precision alone
        with _engine(bhx, n, theta=0.5, eps=0.01 * L, group_walk=1, precision=prec, leaf_capacity=16, mode=1,
                     quadrupole=1, group_size=64) as e:
            e.set_bodies(pos * L, vel, mass * M)
            c, sz = e.bounds()
            e.build(c, sz)
            res.append(e.walk())
    assert np.all(np.isfinite(res[1]))
    d = np.linalg.norm(res[1] - res[0], axis=1) / np.linalg.norm(res[0], axis=1)
    print("SI units: FP32 vs FP64 walk median %.3e max %.3e" % (np.median(d), d.max()), flush=True)
    assert np.median(d) < 1e-5 and d.max() < 1e-3


@pytest.mark.parametrize("prec", [0, 1])
def test_group_walk_is_at_least_as_accurate(bhx, plummer16k, prec):
    pos, vel, mass = plummer16k
    c, s = orc.bounds(pos)
    ref, _ = orc.direct(pos, mass, 0.01, targets=np.arange(0, N1, 8))
    kw = dict(theta=0.5, eps=0.01, mode=1, quadrupole=1, leaf_capacity=32, precision=prec)
    with _engine(bhx, N1, group_walk=1, **kw) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        acc_g, cnt_g = e.walk(counts=True)
        cnt_ref = e.interaction_counts()
        acc_again = e.walk()
    with _engine(bhx, N1, group_walk=0, **kw) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        acc_b, cnt_b = e.walk(counts=True)
    eg = np.linalg.norm(acc_g[::8] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    eb = np.linalg.norm(acc_b[::8] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(eg) < 1e-3
    assert np.median(eg) <= 1.05 * np.median(eb)
    assert np.array_equal(acc_g, acc_again)                    # deterministic, counts call left acc alone
    t = orc.build_recursive(pos, mass, c, s, 32)
    t.moments_intended()
    _, ocnt = t.walk(0.5, 0.01, quad=True)
    assert np.array_equal(cnt_ref.astype(np.int64), ocnt)     # reference list lengths
    assert np.all(cnt_g.astype(np.int64) >= ocnt)              # shared lists refine every body's list
    # the entries that take the quadrupole branch: nodes of more than one body in
    # the reference's own list of each sampled body
    with _engine(bhx, N1, group_walk=1, **kw) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        cnt_q, qc = e.interaction_counts(quad=True)
    assert np.array_equal(cnt_q, cnt_ref)
    nn = t.nodes()["n"]
    for i in range(0, N1, 512):
        ids, kind = t.interaction_list(pos[i], 0.5)
        assert int(qc[i]) == int(np.count_nonzero((kind == 0) & (nn[ids * (kind == 0)] > 1))), i
    assert 0 < qc.sum() < cnt_ref.sum()


def test_group_walk_step_energy(bhx):
    n = 8192
    pos, vel, mass = bhx.plummer(n, seed=23)
    with _engine(bhx, n, theta=0.5, eps=0.01, dt=1e-3, mode=1, quadrupole=1, leaf_capacity=16,
                 precision=1, group_walk=1, compensated=1) as e:
        e.set_bodies(pos, vel, mass)
        e.step(20)
        p, v, m, a = e.get_bodies()
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    assert np.array_equal(np.sort(m), np.sort(mass))
    ek0 = 0.5 * (mass * (vel ** 2).sum(1)).sum()
    ek1 = 0.5 * (m * (v ** 2).sum(1)).sum()
    _, pot0 = orc.direct(pos, mass, 0.01)
    _, pot1 = orc.direct(p, m, 0.01)
    e0 = ek0 + 0.5 * (mass * pot0).sum()
    e1 = ek1 + 0.5 * (m * pot1).sum()
    assert abs(e1 - e0) / abs(e0) < 1e-4


@pytest.mark.parametrize("group_size", [64, 32, 16])
def test_fp32_walk_group_sizes(bhx, group_size):
    """The FP32 group walk with 1, 2 or 4 interaction lists per 64-body wave.  At
    theta = 0 on 40,000 bodies every cell is opened, so the cell stack overflows and
    the cells it cannot hold are walked one at a time: the result is still the direct
    sum.  At theta = 0.5 every body's list refines its reference list (counts at least
    the reference's) and the force error against the FP64 direct sum stays small."""
    n = 40000
    pos, vel, mass = bhx.plummer(n, seed=61)
    c, s = orc.bounds(pos)
    tg = np.random.default_rng(62).choice(n, 512, replace=False)
    with _engine(bhx, n, theta=0.0, eps=0.01, group_walk=1, precision=1, leaf_capacity=16, mode=1,
                 quadrupole=1, group_size=group_size) as e:
        e.set_bodies(pos, vel, mass)
        e.build(c, s)
        acc0, cnt0 = e.walk(counts=True)
        ref, _ = e.direct(tg, 0.01)
        e.set_config(theta=0.5)
        acc5, cnt5 = e.walk(counts=True)
        cref = e.interaction_counts()
    assert np.all(cnt0 == n)     # every body, its own included (zero separation adds nothing)
    e0 = np.linalg.norm(acc0[tg] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    e5 = np.linalg.norm(acc5[tg] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print("group %d: theta 0 max %.2e; theta 0.5 median %.2e, pairs/body %.0f (reference %.0f)"
          % (group_size, e0.max(), np.median(e5), cnt5.mean(), cref.mean()), flush=True)
    assert e0.max() < 1e-4                                   # FP32 rounding over 40,000 terms
    assert np.median(e5) < 1e-3
    assert np.all(cnt5 >= cref)


@pytest.mark.parametrize("prec,comp", [(0, 0), (0, 1), (1, 1)])
def test_group_walk_frame_closing_kick_matches_oracle(bhx, prec, comp):
    """One frame from zero acceleration with a group walk, which applies the closing half
    kick itself as it writes each acceleration: positions are x0 + v0 dt exactly (the first
    half kick adds nothing), and every velocity moved by (a dt) / 2 with a of this frame's walk —
    checked against the oracle's frame (the reference's per-body walk, whose accelerations
    the shared lists approximate to ~1e-4), so a kick missing, doubled or applied with the
    old acceleration would be off by half of a dt."""
    n, dt = 4096, 1e-3
    pos, vel, mass = bhx.plummer(n, seed=31)
    with _engine(bhx, n, theta=0.5, eps=0.01, dt=dt, mode=1, quadrupole=1, leaf_capacity=32, precision=prec,
                 group_walk=1, compensated=comp, permute_acc=1) as e:
        e.set_bodies(pos, vel, mass)
        e.step(1)
        gp, gv, _, ga = e.get_bodies()
    p, v, m, a, _, _ = orc.frame(pos, vel, mass, np.zeros_like(pos), dt, 0.5, 0.01, leaf_cap=32, mode=1,
                                 quad=True, permute_acc=True)
    gi, oi = np.argsort(gp[:, 0], kind="stable"), np.argsort(p[:, 0], kind="stable")
    assert np.array_equal(gp[gi], p[oi])                                 # x0 + v0 dt, bit for bit
    amag = np.linalg.norm(a[oi], axis=1)
    dv = np.linalg.norm(gv[gi] - v[oi], axis=1) / (amag * dt)             # in units of |a| dt
    da = np.linalg.norm(ga[gi] - a[oi], axis=1) / amag
    print("prec %d comp %d: max |dv| / (|a| dt) %.2e, median rel |da| %.2e" % (prec, comp, dv.max(), np.median(da)),
          flush=True)
    assert np.median(da) < 1e-3
    assert dv.max() < 0.05
    # and the kick used this frame's own accelerations: v = v0 + (a dt) / 2 per body (bodies
    # matched by their drifted positions, x0 + v0 dt, which the engine reproduces bit for bit)
    hi = np.argsort((pos + vel * dt)[:, 0], kind="stable")
    assert np.array_equal((pos + vel * dt)[hi], gp[gi])
    kick = np.linalg.norm(gv[gi] - vel[hi] - 0.5 * dt * ga[gi], axis=1)
    assert kick.max() <= 1e-12 * (np.linalg.norm(gv, axis=1).max() + 1.0)
