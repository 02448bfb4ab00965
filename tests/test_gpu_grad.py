"""Reverse-mode gradients of recorded trajectories (bhx_record / bhx_adjoint)
against the oracle's direct-sum adjoint (theta = 0: the group walk opens
every cell, so the walk is the direct sum) and against central finite
differences of the engine's own frames with the decisions frozen (theta > 0).

Tolerances are written at each check.  Loss: L = sum(gx . x_T) + sum(gv . v_T)
with fixed random weights, so dL/dx_T = gx and dL/dv_T = gv.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.helpers import load_bhx

pytestmark = pytest.mark.gpu

DIFF = dict(theta=0.5, eps=0.05, dt=2e-3, leaf_capacity=8, mode=1, precision=0, quadrupole=0, group_walk=1)


@pytest.fixture(scope="module")
def bhx():
    return load_bhx()


def _bodies(bhx, n, seed):
    return bhx.plummer(n, seed=seed)


def _weights(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, 3)), rng.normal(size=(n, 3))


def _loss(e, gx, gv):
    P, V, M, A = e.recorded_bodies()
    return float((gx * P).sum() + (gv * V).sum())


@pytest.mark.parametrize("prec", [0, 1])
def test_adjoint_theta0_matches_oracle_direct_sum(bhx, prec):
    n, T = 300, 12
    pos, vel, mass = _bodies(bhx, n, 5)
    gx, gv = _weights(n, 6)
    cfg = dict(DIFF, theta=0.0, precision=prec)
    print("adjoint theta=0 prec=%d n=%d T=%d" % (prec, n, T), flush=True)
    with bhx.Engine(n, **cfg) as e:
        e.set_bodies(pos, vel, mass)
        e.record(T, checkpoint_every=5, init_acc=True)
        P, V, M, A = e.recorded_bodies()
        G = e.adjoint(gx, gv)
    oP, oV, oA = orc.traj_direct(pos, vel, mass, cfg["eps"], cfg["dt"], T, init_acc=True)
    OG = orc.traj_direct_adjoint(pos, vel, mass, cfg["eps"], cfg["dt"], T, gx, gv, init_acc=True)
    # trajectory: the same frames up to summation order (FP64) or FP32 walk rounding
    tol_x = 1e-12 if prec == 0 else 1e-6
    assert np.allclose(P, oP, rtol=0, atol=tol_x)
    # gradients: FP64 1e-9 relative to the largest component; FP32 walk 1e-4
    tol = 1e-9 if prec == 0 else 1e-4
    for k, name in ((0, "x0"), (1, "v0"), (3, "m")):
        ref = OG[k]
        err = np.abs(G[k] - ref).max() / np.abs(ref).max()
        assert err < tol, (name, err)
    assert np.all(G[2] == 0.0)


def test_adjoint_matches_frozen_finite_differences(bhx):
    """theta = 0.5 with cells: central differences of the engine's own frames.  A step
    of 1e-6 moves no opening decision here (checked: the same trajectory structure),
    so the differences see the frozen-decision derivative."""
    n, T, h = 400, 10, 1e-6
    pos, vel, mass = _bodies(bhx, n, 9)
    gx, gv = _weights(n, 10)
    with bhx.Engine(n, **DIFF) as e:
        e.set_bodies(pos, vel, mass)
        e.record(T, checkpoint_every=4, init_acc=True)
        G = e.adjoint(gx, gv)
        rng = np.random.default_rng(11)
        picks = [(int(i), int(a)) for i, a in zip(rng.choice(n, 8, replace=False), rng.integers(0, 3, 8))]
        for (i, a) in picks:
            for arg in (0, 1):
                vals = []
                for sgn in (1.0, -1.0):
                    p2, v2 = pos.copy(), vel.copy()
                    (p2 if arg == 0 else v2)[i, a] += sgn * h
                    e.set_bodies(p2, v2, mass)
                    e.record(T, checkpoint_every=100, init_acc=True)
                    vals.append(_loss(e, gx, gv))
                fd = (vals[0] - vals[1]) / (2 * h)
                got = G[arg][i, a]
                # central differences: O(h^2) truncation plus ~1e-16 |L| / h roundoff
                assert abs(fd - got) <= 2e-6 * max(1.0, abs(fd)), (arg, i, a, fd, got)


def test_adjoint_independent_of_checkpoint_interval(bhx):
    n, T = 500, 9
    pos, vel, mass = _bodies(bhx, n, 12)
    gx, gv = _weights(n, 13)
    res = []
    for every in (1, 4, 100):
        with bhx.Engine(n, **DIFF) as e:
            e.set_bodies(pos, vel, mass)
            e.record(T, checkpoint_every=every, init_acc=True)
            res.append(e.adjoint(gx, gv))
    for r in res[1:]:
        for k in (0, 1, 3):
            # FP64 atomics sum entry terms in any order: last-bit differences only
            assert np.allclose(r[k], res[0][k], rtol=1e-12, atol=1e-12 * np.abs(res[0][k]).max())


def test_record_matches_plain_frames_and_leaves_end_state(bhx):
    n, T = 700, 7
    pos, vel, mass = _bodies(bhx, n, 14)
    gx, gv = _weights(n, 15)
    with bhx.Engine(n, **DIFF) as e, bhx.Engine(n, **DIFF) as r:
        e.set_bodies(pos, vel, mass)
        e.record(T, checkpoint_every=3, init_acc=False)
        P, V, M, A = e.recorded_bodies()
        end = e.get_bodies()
        e.adjoint(gx, gv)
        assert all(np.array_equal(x, y) for x, y in zip(end, e.get_bodies())), "adjoint leaves the end state"
        r.set_bodies(pos, vel, mass)
        r.step(T)
        rp, rv, rm, ra = r.get_bodies()
        # same frames: the same multiset of bodies (the recording reports them in the starting order)
        assert np.array_equal(np.sort(P[:, 0]), np.sort(rp[:, 0]))
        assert np.array_equal(np.sort(V[:, 1]), np.sort(rv[:, 1]))
        assert all(np.array_equal(x, y) for x, y in zip(end, (rp, rv, rm, ra)))


def test_adjoint_ensemble_equals_single_systems(bhx):
    sizes = [200, 333, 64]
    parts = [bhx.plummer(k, seed=30 + i) for i, k in enumerate(sizes)]
    pos = np.concatenate([p[0] for p in parts])
    vel = np.concatenate([p[1] for p in parts])
    mass = np.concatenate([p[2] for p in parts])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    gx, gv = _weights(pos.shape[0], 31)
    T = 6
    with bhx.Engine(pos.shape[0], **DIFF) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        e.record(T, checkpoint_every=2, init_acc=True)
        G = e.adjoint(gx, gv)
    for k, (p, v, m) in enumerate(parts):
        a, b = int(off[k]), int(off[k + 1])
        with bhx.Engine(p.shape[0], **DIFF) as s:
            s.set_bodies(p, v, m)
            s.record(T, checkpoint_every=2, init_acc=True)
            g1 = s.adjoint(gx[a:b], gv[a:b])
        for q in (0, 1, 3):
            assert np.allclose(G[q][a:b], g1[q], rtol=1e-12, atol=1e-12 * np.abs(g1[q]).max())


def test_differentiation_argument_errors(bhx):
    pos, vel, mass = _bodies(bhx, 100, 1)
    with bhx.Engine(100, **DIFF) as e:
        e.set_bodies(pos, vel, mass)
        with pytest.raises(bhx.BhxError):
            e.adjoint(np.zeros((100, 3)))           # nothing recorded
        e.set_config(quadrupole=1)
        with pytest.raises(bhx.BhxError):
            e.record(3)                              # monopole walk only
        e.set_config(quadrupole=0)
        with pytest.raises(bhx.BhxError):
            e.record(0)


def test_torch_autograd_through_engine(bhx):
    import torch
    from tests.helpers import load_bhx_torch
    bt = load_bhx_torch()
    n, T = 512, 8
    pos, vel, mass = _bodies(bhx, n, 40)
    rng = np.random.default_rng(41)
    target = pos + 1e-3 * rng.normal(size=pos.shape)
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    x0 = torch.tensor(pos, device=dev, requires_grad=True)
    v0 = torch.tensor(vel, device=dev, requires_grad=True)
    m = torch.tensor(mass, device=dev, requires_grad=True)
    tgt = torch.tensor(target, device=dev)
    with bt.make_engine(n, eps=0.05, dt=2e-3, leaf_capacity=8, precision=0) as e:
        xT, vT = bt.simulate(x0, v0, m, T, e, checkpoint_every=3)
        loss = ((xT - tgt) ** 2).sum(dim=1).mean() + 0.5 * (vT ** 2).sum()
        loss.backward()
        # the same gradients straight from the engine
        gx = 2.0 * (xT.detach().cpu().numpy() - target) / n
        gv = vT.detach().cpu().numpy()
        e.set_bodies(pos, vel, mass)
        e.record(T, checkpoint_every=3, init_acc=True)
        G = e.adjoint(gx, gv)
        for got, ref in ((x0.grad, G[0]), (v0.grad, G[1]), (m.grad, G[3])):
            # FP64 atomics: last-bit differences only
            assert np.allclose(got.cpu().numpy(), ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
        # a small step against the gradient lowers the loss
        with torch.no_grad():
            x1 = x0 - 1e-3 * x0.grad / x0.grad.norm()
        xT1, vT1 = bt.simulate(x1, v0.detach(), m.detach(), T, e, checkpoint_every=3)
        loss1 = ((xT1 - tgt) ** 2).sum(dim=1).mean() + 0.5 * (vT1 ** 2).sum()
        assert float(loss1.detach()) < float(loss.detach())


def test_torch_backward_after_recording_replaced_raises(bhx):
    """A second simulate on the same engine replaces the recording; the first
    graph's backward must raise rather than return the other trajectory's
    gradients."""
    import torch
    from tests.helpers import load_bhx_torch
    bt = load_bhx_torch()
    n, T = 256, 3
    pos, vel, mass = _bodies(bhx, n, 44)
    x0 = torch.tensor(pos, requires_grad=True)
    v0 = torch.tensor(vel, requires_grad=True)
    m = torch.tensor(mass, requires_grad=True)
    with bt.make_engine(n, eps=0.05, dt=2e-3, leaf_capacity=8, precision=0) as e:
        xa, _ = bt.simulate(x0, v0, m, T, e, checkpoint_every=3)
        xb, _ = bt.simulate(x0 * 1.001, v0, m, T, e, checkpoint_every=3)
        with pytest.raises(RuntimeError):
            xa.sum().backward()
        xb.sum().backward()                   # the current recording still works
        assert torch.isfinite(x0.grad).all()
