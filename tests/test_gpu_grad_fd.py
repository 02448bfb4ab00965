"""The differentiation workload checked by central finite differences: a Plummer
sphere of 65,536 equal-mass bodies (total mass 1), 1,000 KDK frames at dt = 1e-3,
theta = 0.5, eps = 0.01, leaf bucket 32, checkpoints every 100 frames; loss = mean
|x_T - x_T*|^2 against the trajectory of seeded perturbed initial conditions
(sigma 1e-3 on positions and velocities, 1% on masses).  Gradients of the FP64 walk
(quadrupole cells) with respect to 32 random starting coordinates: 11 positions,
11 velocities, 10 masses, against central differences with h = 1e-6 (positions,
velocities) and h = 1e-6 m_i (masses: an absolute 1e-6 is 6.6% of m = 2^-16, and
such a step moves opening decisions in every frame).

Every perturbed recording reports the base recording's decision digest — the same
trees and interaction lists in all 1,001 walks — so the difference quotients are
of the frozen-decision function the adjoint differentiates.  Tolerances are
written at the check.
"""
import numpy as np
import pytest

from tests.helpers import load_bhx

pytestmark = pytest.mark.gpu

N, T, EVERY, H = 65536, 1000, 100, 1e-6
CFG = dict(theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=32, mode=1, precision=0, quadrupole=1, compensated=1,
           group_walk=1)


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


@pytest.fixture(scope="module")
def base(bhx):
    pos, vel, mass = bhx.plummer(N, seed=21)
    rng = np.random.default_rng(22)
    pt = pos + 1e-3 * rng.normal(size=pos.shape)
    vt = vel + 1e-3 * rng.normal(size=vel.shape)
    mt = mass * (1.0 + 0.01 * rng.normal(size=mass.shape))
    picks = [(0, int(i), int(a)) for i, a in zip(rng.choice(N, 11, replace=False), rng.integers(0, 3, 11))]
    picks += [(1, int(i), int(a)) for i, a in zip(rng.choice(N, 11, replace=False), rng.integers(0, 3, 11))]
    picks += [(2, int(i), 0) for i in rng.choice(N, 10, replace=False)]
    print("differentiation workload: n=%d T=%d every=%d" % (N, T, EVERY), flush=True)
    e = bhx.Engine(N, **CFG)
    try:
        e.set_bodies(pt, vt, mt)
        e.record(T, EVERY, init_acc=True)
        target = e.recorded_bodies()[0]
        e.set_bodies(pos, vel, mass)
        e.record(T, EVERY, init_acc=True)
        d0 = e.record_digest()
        xT = e.recorded_bodies()[0]
        G = e.adjoint(2.0 * (xT - target) / N)
        yield dict(e=e, pos=pos, vel=vel, mass=mass, target=target, d0=d0, G=G, picks=picks)
    finally:
        e.close()


def _loss(x, target):
    return float(np.mean(np.sum((x - target) ** 2, axis=1)))


@pytest.mark.parametrize("arg", [0, 1, 2], ids=["positions", "velocities", "masses"])
def test_adjoint_matches_frozen_fd_over_1000_frames(base, arg):
    e, pos, vel, mass, G = base["e"], base["pos"], base["vel"], base["mass"], base["G"]
    got_all = G[arg] if arg < 2 else G[3]
    scale = np.abs(got_all).max()
    for (kind, i, a) in base["picks"]:
        if kind != arg:
            continue
        step = H * (mass[i] if arg == 2 else 1.0)
        vals = []
        for sgn in (1.0, -1.0):
            p2, v2, m2 = pos.copy(), vel.copy(), mass.copy()
            if arg == 0:
                p2[i, a] += sgn * step
            elif arg == 1:
                v2[i, a] += sgn * step
            else:
                m2[i] += sgn * step
            e.set_bodies(p2, v2, m2)
            e.record(T, EVERY, init_acc=True)
            assert e.record_digest() == base["d0"], ("a decision moved", arg, i, a, sgn)
            vals.append(_loss(e.recorded_bodies()[0], base["target"]))
        fd = (vals[0] - vals[1]) / (2 * step)
        got = G[arg][i, a] if arg < 2 else G[3][i]
        print("arg=%d i=%d a=%d fd=%.9e adjoint=%.9e" % (arg, i, a, fd, got), flush=True)
        # O(h^2) truncation is ~1e-12 relative; the roundoff of L (~1e-16 |L| per
        # frame, amplified by the trajectory) over the step gives ~1e-7 relative for
        # positions and velocities (step 1e-6) and ~1e-5 for masses (step 1.5e-11)
        rel, floor = (1e-5, 1e-7) if arg < 2 else (2e-4, 2e-6)
        assert abs(fd - got) <= rel * abs(got) + floor * scale, (arg, i, a, fd, got, scale)
