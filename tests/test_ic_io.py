"""Host-side generators and file formats (no GPU): the galaxy-pair initial
conditions against closed forms they must follow, snapshot files, and the
legacy-VTK export writers read back by an independent parser."""
import os
import struct

import numpy as np
import pytest
from scipy import special

from tests.helpers import load_bhx


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


N_GAL = 240_000


@pytest.fixture(scope="module")
def pair(bhx):
    return bhx.galaxy_pair(N_GAL, seed=11, threads=4)


def _layout(n, ratio=3.0, fd=0.2):
    n1 = int(round(n * ratio / (1 + ratio)))
    out = []
    first = 0
    for cnt in (n1, n - n1):
        nd = int(round(cnt * fd))
        out.append((first, cnt - nd, nd))
        first += cnt
    return out


def _hernquist_sigma2(r, M, a):
    # isotropic one-dimensional dispersion, Hernquist (1990) eq. 10 (G = 1)
    x = r / a
    return M / (12 * a) * (12 * x * (1 + x) ** 3 * np.log((1 + x) / x)
                           - x / (1 + x) * (25 + 52 * x + 42 * x ** 2 + 12 * x ** 3))


def test_galaxy_pair_totals_and_determinism(bhx, pair):
    pos, vel, mass = pair
    assert np.allclose(mass, 1.0 / N_GAL, rtol=0, atol=0)
    # centre-of-mass frame
    assert np.abs(pos.mean(0)).max() < 1e-12
    assert np.abs(vel.mean(0)).max() < 1e-14
    # the draws depend only on (seed, body): any thread count gives the same bodies
    p1, v1, m1 = bhx.galaxy_pair(N_GAL, seed=11, threads=1)
    assert np.array_equal(p1, pos) and np.array_equal(v1, vel)
    p2, _, _ = bhx.galaxy_pair(N_GAL, seed=12, threads=4)
    assert not np.array_equal(p2, pos)


def test_galaxy_pair_parabolic_orbit(pair):
    pos, vel, mass = pair
    (f1, h1, d1), (f2, h2, d2) = _layout(N_GAL)
    c1, c2 = pos[f1:f1 + h1 + d1].mean(0), pos[f2:].mean(0)
    u1, u2 = vel[f1:f1 + h1 + d1].mean(0), vel[f2:].mean(0)
    r, v = c2 - c1, u2 - u1
    assert np.allclose(r, [30.0, 0.0, 0.0], atol=1e-10)
    M = 1.0
    # zero orbital energy: |v|^2 = 2 M / d; angular momentum sqrt(2 M r_p), r_p = 5
    assert abs(v @ v - 2 * M / 30.0) < 1e-12
    assert abs(np.linalg.norm(np.cross(r, v)) - np.sqrt(2 * M * 5.0)) < 1e-12
    assert v[0] < 0   # approaching
    # masses split 3 : 1
    assert abs((h1 + d1) / N_GAL - 0.75) < 1e-5


def test_galaxy_pair_halo_follows_hernquist_dispersion(pair):
    pos, vel, mass = pair
    (f, nh, nd), _ = _layout(N_GAL)
    g = slice(f, f + nh + nd)
    cen, vb = pos[g].mean(0), vel[g].mean(0)
    x = pos[f:f + nh] - cen
    u = vel[f:f + nh] - vb
    r = np.linalg.norm(x, axis=1)
    a = 3.0
    cut = 10.0
    Minf = nh * mass[0] / (cut ** 2 / (1 + cut) ** 2)
    for lo, hi in ((0.5, 1.5), (1.5, 3.0), (3.0, 6.0), (6.0, 12.0), (12.0, 24.0)):
        sel = (r >= lo) & (r < hi)
        assert sel.sum() > 5000
        got = (u[sel] ** 2).sum(1).mean() / 3.0
        want = _hernquist_sigma2(r[sel], Minf, a).mean()
        assert abs(got / want - 1) < 0.03, (lo, hi, got, want)
    # cumulative mass profile: fraction inside r is r^2/(r+a)^2 / X_cut
    xc = cut ** 2 / (1 + cut) ** 2
    for rr in (1.0, 3.0, 10.0):
        frac = (r < rr).mean()
        assert abs(frac - rr ** 2 / (rr + a) ** 2 / xc) < 0.005
    # the cut is at r = 30 about the generated centre; measured from the galaxy's
    # mean position it moves by that offset (a few hundredths here)
    assert r.max() <= 30.0 + 0.05


def test_galaxy_pair_disk_rotation_and_shape(pair):
    pos, vel, mass = pair
    (f, nh, nd), _ = _layout(N_GAL)
    g = slice(f, f + nh + nd)
    cen, vb = pos[g].mean(0), vel[g].mean(0)
    x = pos[f + nh:f + nh + nd] - cen
    u = vel[f + nh:f + nh + nd] - vb
    R = np.hypot(x[:, 0], x[:, 1])
    vphi = (x[:, 0] * u[:, 1] - x[:, 1] * u[:, 0]) / R
    Md = nd * mass[0]
    Minf = nh * mass[0] / (100.0 / 121.0)
    a, Rd = 3.0, 1.0
    y = R / (2 * Rd)
    vc = np.sqrt(Minf * R / (R + a) ** 2
                 + 2 * Md / Rd * y ** 2 * (special.i0(y) * special.k0(y) - special.i1(y) * special.k1(y)))
    for lo, hi in ((0.5, 1.0), (1.0, 2.0), (2.0, 4.0), (4.0, 8.0)):
        sel = (R >= lo) & (R < hi)
        assert abs(vphi[sel].mean() / vc[sel].mean() - 1) < 0.01, (lo, hi)
        # dispersion 0.1 v_c in each component
        vr = (x[sel, 0] * u[sel, 0] + x[sel, 1] * u[sel, 1]) / R[sel]
        assert abs(np.sqrt((vr ** 2).mean()) / (0.1 * np.sqrt((vc[sel] ** 2).mean())) - 1) < 0.05
    # exponential surface density: R / R_d ~ Gamma(2) below the cut (median 1.678...)
    assert abs(np.median(R) - 1.6783469900166608) < 0.02
    assert R.max() <= 10.0 + 0.05   # cut about the generated centre
    # sech^2 layer: median |z| = z_d atanh(1/2)
    assert abs(np.median(np.abs(x[:, 2])) - 0.1 * np.arctanh(0.5)) < 0.002


def test_galaxy_pair_secondary_scale_and_inclination(pair):
    pos, vel, mass = pair
    (f1, h1, d1), (f2, h2, d2) = _layout(N_GAL)
    s = ((h2 + d2) / (h1 + d1)) ** (1.0 / 3.0)
    c2, v2 = pos[f2:].mean(0), vel[f2:].mean(0)
    x = pos[f2 + h2:] - c2
    u = vel[f2 + h2:] - v2
    L = np.cross(x, u).sum(0)
    ang = np.degrees(np.arccos(L[2] / np.linalg.norm(L)))
    assert abs(ang - 60.0) < 1.0
    # disk radius (in the disk plane) scales by (M2/M1)^(1/3)
    n = L / np.linalg.norm(L)
    R = np.linalg.norm(x - np.outer(x @ n, n), axis=1)
    assert abs(np.median(R) / (1.6783469900166608 * s) - 1) < 0.03


def test_galaxy_pair_argument_errors(bhx):
    with pytest.raises(bhx.BhxError):
        bhx.galaxy_pair(3)
    with pytest.raises(bhx.BhxError):
        bhx.galaxy_pair(1000, pericentre=40.0)   # beyond the separation
    with pytest.raises(bhx.BhxError):
        bhx.galaxy_pair(1000, disk_fraction=1.5)


# ---- snapshots

def _bodies(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.random(n), rng.normal(size=(n, 3))


def test_snapshot_round_trip(bhx, tmp_path):
    pos, vel, mass, acc = _bodies(1001, 1)
    off = np.array([0, 10, 500, 1001], np.uint64)
    cfg = bhx.default_config(theta=0.7, eps=0.02, leaf_capacity=16, precision=bhx.FP32, group_walk=1, geometry=1)
    path = str(tmp_path / "a.bhx")
    comp = np.random.default_rng(9).normal(size=(2, 1001, 3)) * 1e-17
    bhx.write_snapshot(path, pos, vel, mass, acc=acc, offsets=off, time=1.25, step=77, config=cfg, comp=comp)
    assert os.path.getsize(path) == 256 + 8 * (16 * 1001 + 4)
    d = bhx.read_snapshot(path)
    for k, v in (("pos", pos), ("vel", vel), ("mass", mass), ("acc", acc), ("comp", comp)):
        assert np.array_equal(d[k], v), k
    assert np.array_equal(d["offsets"], off)
    assert d["time"] == 1.25 and d["step"] == 77
    c = d["config"]
    assert (c.theta, c.eps, c.leaf_capacity, c.precision, c.group_walk, c.geometry) == (0.7, 0.02, 16, bhx.FP32, 1, 1)
    # without accelerations and systems
    path2 = str(tmp_path / "b.bhx")
    bhx.write_snapshot(path2, pos, vel, mass)
    d2 = bhx.read_snapshot(path2)
    assert d2["acc"] is None and d2["comp"] is None and list(d2["offsets"]) == [0, 1001]
    assert np.array_equal(d2["pos"], pos)


def test_snapshot_header_layout(bhx, tmp_path):
    pos, vel, mass, _ = _bodies(5, 2)
    path = str(tmp_path / "h.bhx")
    bhx.write_snapshot(path, pos, vel, mass, time=2.5, step=9)
    raw = open(path, "rb").read()
    assert raw[:8] == b"BHXSNAP1"
    hb, ver, n, nsys = struct.unpack_from("<IIQQ", raw, 8)
    t, step, flags = struct.unpack_from("<dQI", raw, 32)
    assert (hb, ver, n, nsys, t, step, flags) == (256, 1, 5, 1, 2.5, 9, 0)
    assert np.array_equal(np.frombuffer(raw, "<f8", 15, 256).reshape(5, 3), pos)
    assert np.array_equal(np.frombuffer(raw, "<f8", 5, 256 + 8 * 30), mass)


def test_snapshot_rejects_bad_files(bhx, tmp_path):
    pos, vel, mass, _ = _bodies(50, 3)
    path = str(tmp_path / "c.bhx")
    bhx.write_snapshot(path, pos, vel, mass)
    raw = open(path, "rb").read()
    bad = str(tmp_path / "bad.bhx")
    for blob in (raw[:-8], raw + b"\0" * 8, b"XXXXXXXX" + raw[8:], raw[:12] + b"\x02" + raw[13:], b""):
        open(bad, "wb").write(blob)
        with pytest.raises(bhx.BhxError) as ei:
            bhx.read_snapshot(bad)
        assert ei.value.code == 11
    with pytest.raises(bhx.BhxError):
        bhx.read_snapshot(str(tmp_path / "missing.bhx"))
    with pytest.raises(bhx.BhxError) as ei:
        bhx.write_snapshot(path, pos, vel, mass, offsets=np.array([0, 60], np.uint64))
    assert ei.value.code == 2


# ---- legacy VTK

def _read_vtk(path):
    """Independent reader for the legacy-VTK binary files: returns the ASCII
    section keywords with their big-endian payloads."""
    data = open(path, "rb").read()
    lines, pos = [], 0

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode()
        pos = e + 1
        return s

    assert line() == "# vtk DataFile Version 3.0"
    line()
    assert line() == "BINARY"
    out = {"dataset": line().split()[1]}
    while pos < len(data):
        s = line()
        if not s:
            continue
        w = s.split()
        if w[0] == "POINTS":
            k = int(w[1])
            out["points"] = np.frombuffer(data, ">f4", 3 * k, pos).reshape(k, 3)
            pos += 12 * k
        elif w[0] in ("VERTICES", "CELLS"):
            k = int(w[2])
            out[w[0].lower()] = np.frombuffer(data, ">i4", k, pos)
            pos += 4 * k
        elif w[0] == "CELL_TYPES":
            k = int(w[1])
            out["cell_types"] = np.frombuffer(data, ">i4", k, pos)
            pos += 4 * k
        elif w[0] in ("POINT_DATA", "CELL_DATA"):
            out["_n"] = int(w[1])
        elif w[0] == "SCALARS":
            assert line() == "LOOKUP_TABLE default"
            dt = ">f4" if w[2] == "float" else ">i4"
            out[w[1]] = np.frombuffer(data, dt, out["_n"], pos)
            pos += 4 * out["_n"]
        else:
            raise AssertionError(s)
    return out


def test_vtk_bodies(bhx, tmp_path):
    pos, vel, mass, _ = _bodies(333, 4)
    path = str(tmp_path / "b.vtk")
    bhx.write_vtk_bodies(path, pos, vel, mass)
    d = _read_vtk(path)
    assert d["dataset"] == "POLYDATA"
    assert np.array_equal(d["points"], pos.astype(np.float32))
    v = d["vertices"].reshape(-1, 2)
    assert np.array_equal(v[:, 0], np.ones(333)) and np.array_equal(v[:, 1], np.arange(333))
    assert np.array_equal(d["mass"], mass.astype(np.float32))
    assert np.allclose(d["speed"], np.linalg.norm(vel, axis=1), rtol=1e-6)


def test_vtk_nodes(bhx, tmp_path):
    nodes = np.zeros(5, bhx.NODE_DTYPE)
    nodes["center"] = [[0, 0, 0], [1, 2, 3], [-1, 0, 0.5], [4, 4, 4], [0, 1, 0]]
    nodes["size"] = [2.0, 1.0, 0.5, 0.25, 1.0]
    nodes["level"] = [0, 1, 2, 3, 1]
    nodes["n"] = [100, 40, 7, 1, 60]
    nodes["mass"] = [1.0, 0.4, 0.07, 0.01, 0.6]
    path = str(tmp_path / "n.vtk")
    bhx.write_vtk_nodes(path, nodes, min_level=1, max_level=2)
    d = _read_vtk(path)
    assert d["dataset"] == "UNSTRUCTURED_GRID"
    sel = nodes[(nodes["level"] >= 1) & (nodes["level"] <= 2)]
    m = sel.size
    assert m == 3 and np.array_equal(d["cell_types"], np.full(m, 11))
    cells = d["cells"].reshape(m, 9)
    assert np.array_equal(cells[:, 0], np.full(m, 8))
    assert np.array_equal(cells[:, 1:].ravel(), np.arange(8 * m))
    pts = d["points"].reshape(m, 8, 3)
    for j in range(m):
        # a cube of edge `size` around the centre, corners in voxel order (x fastest)
        h = 0.5 * sel["size"][j]
        c = sel["center"][j]
        for k in range(8):
            want = c + np.array([h if k & 1 else -h, h if k & 2 else -h, h if k & 4 else -h])
            assert np.allclose(pts[j, k], want)
    assert np.array_equal(d["level"], sel["level"])
    assert np.array_equal(d["count"], sel["n"])
    assert np.allclose(d["mass"], sel["mass"])
