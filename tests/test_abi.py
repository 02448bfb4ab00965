"""Host-side checks of the engine library: it loads, exports every entry point
include/bhx.h declares, and its host-only initial-condition generator is sound.
No GPU compute is called here."""
import os
import re

import numpy as np
import pytest

from tests.helpers import load_bhx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "bhx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(bhx_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    bhx = load_bhx()
    L = bhx.lib()
    names = declared_entry_points()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_config_default_matches_reference_constants():
    bhx = load_bhx()
    cfg = bhx.default_config()
    assert cfg.eps == 0.025          # SOFTENING, HashedOctree.hpp:151
    assert cfg.dt == 1e-3            # ofApp.h:62
    assert cfg.leaf_capacity == 1    # one body per leaf, HashedOctree.cpp:480-504
    assert cfg.mode == bhx.AS_RUN and cfg.permute_acc == 1


def test_plummer_generator_is_seeded_and_centred():
    bhx = load_bhx()
    p1, v1, m1 = bhx.plummer(4096, seed=5)
    p2, v2, m2 = bhx.plummer(4096, seed=5)
    p3, _, _ = bhx.plummer(4096, seed=6)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)
    assert not np.array_equal(p1, p3)
    assert np.all(m1 == 1.0 / 4096) and m1.sum() == 1.0
    assert np.abs((m1[:, None] * p1).sum(0)).max() < 1e-12
    assert np.abs((m1[:, None] * v1).sum(0)).max() < 1e-12


def test_plummer_generator_profile_and_virial():
    from oracle import oracle as orc
    bhx = load_bhx()
    n = 4096
    pos, vel, mass = bhx.plummer(n, seed=11)
    r = np.linalg.norm(pos, axis=1)
    # half-mass radius of a Plummer sphere with a = 1: 1/sqrt(2^(2/3) - 1) = 1.3048
    assert abs(np.median(r) - 1.3048) < 0.06
    ek = 0.5 * (mass * (vel ** 2).sum(1)).sum()
    _, pot = orc.direct(pos, mass, 0.0)
    ep = 0.5 * (mass * pot).sum()
    assert abs(2 * ek / -ep - 1.0) < 0.08      # virial equilibrium 2K = -W


def test_torch_wrapper_imports():
    from tests.helpers import load_bhx_torch
    bt = load_bhx_torch()
    assert callable(bt.simulate) and callable(bt.make_engine)


def test_binding_rejects_mismatched_shapes(tmp_path):
    """The binding checks array shapes before the C side reads n (or 3n) values
    through each pointer: no GPU call is made for a rejected call."""
    bhx = load_bhx()
    e = bhx.Engine.__new__(bhx.Engine)     # no context: the checks come first
    e.h = None
    e.record_generation = 0
    pos = np.zeros((4, 3))
    with pytest.raises(ValueError):
        e.set_bodies(pos, None, np.ones(3))
    with pytest.raises(ValueError):
        e.set_bodies(pos, np.zeros((3, 3)), np.ones(4))
    with pytest.raises(ValueError):
        e.set_bodies(np.zeros(12), None, np.ones(4))
    with pytest.raises(ValueError):
        bhx.write_snapshot(str(tmp_path / "x.snap"), pos, np.zeros((3, 3)), np.ones(4))
    with pytest.raises(ValueError):
        bhx.write_snapshot(str(tmp_path / "x.snap"), pos, np.zeros((4, 3)), np.ones(4), acc=np.zeros((5, 3)))
    bhx.write_snapshot(str(tmp_path / "ok.snap"), pos, np.zeros((4, 3)), np.ones(4))
    assert bhx.read_snapshot(str(tmp_path / "ok.snap"))["mass"].size == 4
