"""OctantBounds on large systems (the parallel exact mean) against the oracle's
left-to-right FP64 sum, bit for bit, on inputs chosen to stress the integer
grid: Plummer spheres (the running sum wanders across many binades), values on
coarse dyadic grids (many exact half-ulp ties), one-signed data, wide dynamic
range, non-finite values, and ensembles mixing large and small systems."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import load_bhx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


def _check(bhx, pos):
    n = pos.shape[0]
    print("bounds n=%d" % n, flush=True)
    with bhx.Engine(n) as e:
        e.set_bodies(pos, np.zeros_like(pos), np.full(n, 1.0 / n))
        c, s = e.bounds()
    oc, os_ = orc.bounds(pos)
    assert np.array_equal(c, oc, equal_nan=True), (c, oc)
    assert s == os_ or (np.isnan(s) and np.isnan(os_))


def test_plummer_large(bhx):
    for n, seed in ((1 << 17, 1), (100003, 2), (1 << 20, 3)):
        pos, _, _ = bhx.plummer(n, seed=seed)
        _check(bhx, pos)


def test_dyadic_grid_ties(bhx):
    rng = np.random.default_rng(4)
    for k in (0, 6, 12, 30):
        pos = rng.integers(-2000, 2000, size=(70001, 3)).astype(np.float64) * 2.0 ** -k
        _check(bhx, pos)


def test_one_signed_and_wide_range(bhx):
    rng = np.random.default_rng(5)
    n = 65536 + 17
    pos = np.abs(rng.normal(size=(n, 3))) * 3.0 + 1.0
    _check(bhx, pos)
    _check(bhx, -pos)
    wide = rng.normal(size=(n, 3)) * 10.0 ** rng.integers(-12, 12, size=(n, 3))
    _check(bhx, wide)
    tiny = rng.normal(size=(n, 3)) * 1e-300
    _check(bhx, tiny)


def test_non_finite(bhx):
    rng = np.random.default_rng(6)
    n = 40000
    pos = rng.normal(size=(n, 3))
    pos[12345, 0] = np.nan
    pos[20000, 1] = np.inf
    pos[30000, 2] = -np.inf
    pos[30001, 2] = np.inf
    _check(bhx, pos)


def test_ensemble_mixed_sizes(bhx):
    sizes = [70000, 100, 33000, 5, 40000]
    parts = [bhx.plummer(n, seed=50 + k) for k, n in enumerate(sizes)]
    pos = np.concatenate([p[0] for p in parts])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    with bhx.Engine(pos.shape[0]) as e:
        e.set_bodies(pos, np.zeros_like(pos), np.full(pos.shape[0], 1e-5))
        e.set_systems(off)
        C, S = e.system_bounds()
    for k, (p, _, _) in enumerate(parts):
        oc, os_ = orc.bounds(p)
        assert np.array_equal(C[k], oc) and S[k] == os_, k


def _monotone_then_hover(n, seed):
    """About 80 % positive values (partial sums grow steadily: runs of whole binades)
    followed by about 20 % values that keep every partial sum within a few thousand
    ulps of the binade edge 2^17 (each such element needs its own addition, so those
    tiles are added element by element), for each axis."""
    rng = np.random.default_rng(seed)
    m = (4 * n) // 5
    x = np.empty((n, 3))
    for a in range(3):
        x[:m, a] = rng.uniform(0.5, 1.0, m)
        s = 0.0
        for k in range(m):
            s = s + x[k, a]
        edge = 131072.0
        for k in range(m, n):
            if (k - m) % 2 == 0:
                v = edge - s + 1e-7 * rng.uniform(-1, 1)
            else:
                v = 1e-5 * rng.uniform(-1, 1)
            x[k, a] = v
            s = s + v
    return x


def test_chain_over_runs_with_element_by_element_stretch(bhx):
    """The chain over exact runs (not the plain chain) handles a stretch of about 20
    consecutive tiles added element by element, bit for bit: the walker's merged
    ranges over several tiles and its multi-chunk staging through LDS.  The path
    taken is read back from the device."""
    n = 100003
    pos = _monotone_then_hover(n, 11)
    print("bounds monotone+hover n=%d" % n, flush=True)
    with bhx.Engine(n) as e:
        e.set_bodies(pos, np.zeros_like(pos), np.full(n, 1.0 / n))
        c, s = e.bounds()
        path = e.bounds_path()
        # a zero-mean sphere in generator order: the partial sums wander near zero,
        # (almost) every tile is added element by element and the plain chain runs
        rp = bhx.plummer(n, seed=2)[0]
        e.set_bodies(rp, np.zeros_like(rp), np.full(n, 1.0 / n))
        c2, s2 = e.bounds()
        path2 = e.bounds_path()
    tiles = (n + 1023) // 1024
    print("path %s (of %d tiles); shuffled %s" % (path, tiles, path2), flush=True)
    oc, os_ = orc.bounds(pos)
    assert c.tobytes() == oc.tobytes() and s == os_
    assert path[3] == 0 and all(10 <= path[a] <= tiles // 2 for a in range(3))
    oc2, os2 = orc.bounds(rp)
    assert c2.tobytes() == oc2.tobytes() and s2 == os2
    assert path2[3] == 1 and min(path2[:3]) > (3 * tiles) // 4
