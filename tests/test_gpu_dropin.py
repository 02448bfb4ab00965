"""The drop-in C++ API (include/bhx_dropin.hpp) driven the way the reference's
ofApp drives its own functions, checked against the oracle's frames.

The reference's RadixSortBodies does not reorder the caller's acceleration
array (ofApp.cpp:176,248), so the oracle runs with permute_acc off.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import load_bhx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "dropin_frame")


def test_dropin_driver_built():
    assert os.path.exists(DRIVER), "build() compiles tests/cpp/dropin_frame"


@pytest.mark.gpu
def test_dropin_frames_match_oracle(tmp_path):
    n, frames, seed, theta, dt = 2048, 3, 31, 0.5, 1e-3
    out = tmp_path / "frames.bin"
    print("dropin_frame n=%d frames=%d" % (n, frames), flush=True)
    r = subprocess.run([DRIVER, str(n), str(frames), str(seed), str(theta), str(dt), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, dtype=np.float64)
    body = raw[:10 * n].reshape(n, 10)
    tail = raw[10 * n:]
    bhx = load_bhx()
    pos, vel, mass = bhx.plummer(n, seed=seed)
    p, v, m, a = pos, vel, mass, np.zeros_like(pos)
    for _ in range(frames):
        p, v, m, a, keys, st = orc.frame(p, v, m, a, dt, theta, 0.01, permute_acc=False)
    assert np.array_equal(body[:, 6], m)
    assert np.allclose(body[:, 0:3], p, rtol=0, atol=1e-13)
    assert np.allclose(body[:, 3:6], v, rtol=0, atol=1e-12)
    assert np.allclose(body[:, 7:10], a, rtol=1e-10, atol=1e-12)
    root_n, root_mass, child_key, missing, threw, nodes = tail
    assert root_n == n and root_mass == pytest.approx(1.0, rel=1e-14)
    assert child_key > 1 and missing == 0.0 and threw == 1.0
    assert nodes == st.nodes
