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


def _perturbed(pos, vel, mass, arg, i, a, dh):
    p2, v2, m2 = pos.copy(), vel.copy(), mass.copy()
    if arg == 0:
        p2[i, a] += dh
    elif arg == 1:
        v2[i, a] += dh
    else:
        m2[i] += dh
    return p2, v2, m2


def _frozen_fd(e, pos, vel, mass, T, picks, loss, digest0, h=1e-6):
    """Central differences of loss(recorded bodies) for each pick (arg 0: x, 1: v, 2: m).
    Every perturbed recording must report the base recording's decision digest: the
    same trees and interaction lists in every frame, so the difference quotient is
    one of the frozen-decision function the adjoint differentiates."""
    out = []
    for (arg, i, a) in picks:
        vals = []
        for sgn in (1.0, -1.0):
            e.set_bodies(*_perturbed(pos, vel, mass, arg, i, a, sgn * h))
            e.record(T, checkpoint_every=T, init_acc=True)
            assert e.record_digest() == digest0, ("a decision moved", arg, i, a, sgn)
            vals.append(loss(e.recorded_bodies()))
        out.append((vals[0] - vals[1]) / (2 * h))
    return out


def _grad_of(G, arg, i, a):
    return G[arg][i, a] if arg < 2 else G[3][i]


@pytest.mark.parametrize("quad", [0, 1])
def test_adjoint_matches_frozen_finite_differences(bhx, quad):
    """theta = 0.5 with cells, monopole and quadrupole: central differences of the
    engine's own frames on positions, velocities and masses, every perturbed run
    proven on the same decisions by its digest."""
    n, T, h = 400, 10, 1e-6
    pos, vel, mass = _bodies(bhx, n, 9)
    gx, gv = _weights(n, 10)
    print("frozen FD n=%d T=%d quadrupole=%d" % (n, T, quad), flush=True)
    with bhx.Engine(n, **dict(DIFF, quadrupole=quad)) as e:
        e.set_bodies(pos, vel, mass)
        e.record(T, checkpoint_every=4, init_acc=True)
        d0 = e.record_digest()
        G = e.adjoint(gx, gv)
        rng = np.random.default_rng(11)
        picks = [(arg, int(i), int(a)) for arg in (0, 1, 2)
                 for i, a in zip(rng.choice(n, 8, replace=False), rng.integers(0, 3, 8))]
        fds = _frozen_fd(e, pos, vel, mass, T, picks, lambda r: _loss_of(r, gx, gv), d0, h)
    for (arg, i, a), fd in zip(picks, fds):
        got = _grad_of(G, arg, i, a)
        # central differences: O(h^2) truncation plus ~1e-16 |L| / h roundoff
        assert abs(fd - got) <= 2e-6 * max(1.0, abs(fd)), (arg, i, a, fd, got)


def _loss_of(rec, gx, gv):
    P, V, M, A = rec
    return float((gx * P).sum() + (gv * V).sum())


def test_record_digest_tracks_decisions(bhx):
    """Equal starting states give equal digests; a body moved across the cluster
    changes the trees and lists and so the digest; the FP32 walk collects none."""
    n, T = 500, 5
    pos, vel, mass = _bodies(bhx, n, 40)
    with bhx.Engine(n, **DIFF) as e:
        e.set_bodies(pos, vel, mass)
        e.record(T, checkpoint_every=2, init_acc=True)
        d0 = e.record_digest()
        e.set_bodies(pos, vel, mass)
        e.record(T, checkpoint_every=5, init_acc=True)
        assert e.record_digest() == d0
        p2 = pos.copy()
        p2[7] = -p2[7] + 0.3
        e.set_bodies(p2, vel, mass)
        e.record(T, checkpoint_every=5, init_acc=True)
        assert e.record_digest() != d0
        e.set_config(precision=1)
        with pytest.raises(bhx.BhxError):
            e.record_digest()                        # the recording was dropped with the config
        e.set_bodies(pos, vel, mass)
        e.record(T, checkpoint_every=5, init_acc=True)
        with pytest.raises(bhx.BhxError):
            e.record_digest()                        # FP64 walk only


def test_adjoint_quadrupole_fp32_walk_matches_fp64_walk(bhx):
    """The FP32 walk's quadrupole adjoint (scaled units) against the FP64 walk's on
    the same starting state: the same function up to FP32 rounding and the rare list
    entry whose opening test rounds the other way."""
    n, T = 3000, 6
    pos, vel, mass = _bodies(bhx, n, 41)
    gx, gv = _weights(n, 42)
    res = []
    for prec in (0, 1):
        with bhx.Engine(n, **dict(DIFF, quadrupole=1, precision=prec, eps=0.01, leaf_capacity=16)) as e:
            e.set_bodies(pos, vel, mass)
            e.record(T, checkpoint_every=3, init_acc=True)
            res.append(e.adjoint(gx, gv))
    for k, name in ((0, "x0"), (1, "v0"), (3, "m")):
        err = np.abs(res[1][k] - res[0][k]).max() / np.abs(res[0][k]).max()
        # FP32 terms: ~1e-6 relative each; a flipped list entry moves a few terms by the
        # monopole-quadrupole truncation (~1e-3 of a cell term)
        assert err < 2e-3, (name, err)


def test_adjoint_independent_of_checkpoint_interval(bhx):
    n, T = 500, 9
    pos, vel, mass = _bodies(bhx, n, 12)
    gx, gv = _weights(n, 13)
    res = []
    # every frame kept by the recording (default memory), or the segments re-run (none kept)
    for every, memory in ((1, None), (4, None), (100, None), (4, 0), (100, 0)):
        with bhx.Engine(n, **DIFF) as e:
            e.set_record_memory(memory)
            e.set_bodies(pos, vel, mass)
            e.record(T, checkpoint_every=every, init_acc=True)
            assert e.record_keeps_frames() == (memory is None)
            res.append(e.adjoint(gx, gv))
    for r in res[1:]:
        for k in (0, 1, 3):
            # every sum has a fixed order: the same bits whatever the interval
            assert np.array_equal(r[k], res[0][k]), k


@pytest.mark.parametrize("prec,quad", [(1, 0), (1, 1), (0, 1)])
def test_adjoint_repeats_bitwise(bhx, prec, quad):
    """The entry sums are added in a fixed order (sorted by entry, no atomics),
    so the gradients repeat bit for bit: twice on one recording, with the
    frames kept or re-run, and for any checkpoint interval (FP32 and FP64)."""
    n, T = 3000, 8
    pos, vel, mass = _bodies(bhx, n, 40 + quad)
    gx, gv = _weights(n, 41)
    cfg = dict(DIFF, precision=prec, quadrupole=quad, leaf_capacity=16)
    res = []
    for every, memory in ((3, None), (3, 0), (8, 0), (1, None)):
        with bhx.Engine(n, **cfg) as e:
            e.set_record_memory(memory)
            e.set_bodies(pos, vel, mass)
            e.record(T, checkpoint_every=every, init_acc=True)
            assert e.record_keeps_frames() == (memory is None)
            g = e.adjoint(gx, gv)
            again = e.adjoint(gx, gv)
            for k in (0, 1, 3):
                assert np.array_equal(g[k], again[k]), (every, memory, k)
            res.append(g)
    for r in res[1:]:
        for k in (0, 1, 3):
            assert np.array_equal(r[k], res[0][k]), k


@pytest.mark.parametrize("prec", [1, 0])
def test_adjoint_overflowing_lists(bhx, prec):
    """Lists longer than their stored region are evaluated by the traversal
    itself (entry sums with atomics): a 64- and a 640-entry region (every
    group overflows, after one or after several flushes) give the stored
    lists' gradients to rounding; bad capacities are refused."""
    n, T = 2500, 5
    pos, vel, mass = _bodies(bhx, n, 50 + prec)
    gx, gv = _weights(n, 51)
    cfg = dict(DIFF, precision=prec, quadrupole=1, leaf_capacity=16)
    res = []
    for cap in (0, 64, 640):
        with bhx.Engine(n, **cfg) as e:
            e.set_adjoint_list_capacity(cap)
            e.set_bodies(pos, vel, mass)
            e.record(T, checkpoint_every=T, init_acc=True)
            res.append(e.adjoint(gx, gv))
    for r in res[1:]:
        for k in (0, 1, 3):
            scale = np.abs(res[0][k]).max()
            err = np.abs(r[k] - res[0][k]).max() / scale
            # the same terms summed in another order (FP32 walk: FP32 group sums)
            assert err < (1e-5 if prec == 1 else 1e-12), (k, err)
    with bhx.Engine(64, **cfg) as e:
        for bad in (1, 63, 100):
            with pytest.raises(bhx.BhxError):
                e.set_adjoint_list_capacity(bad)


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
            # each system's entries are summed in the same order as alone: bit for bit
            assert np.array_equal(G[q][a:b], g1[q]), (k, q)


def test_differentiation_argument_errors(bhx):
    pos, vel, mass = _bodies(bhx, 100, 1)
    with bhx.Engine(100, **DIFF) as e:
        e.set_bodies(pos, vel, mass)
        with pytest.raises(bhx.BhxError):
            e.adjoint(np.zeros((100, 3)))           # nothing recorded
        e.set_config(group_size=32)
        with pytest.raises(bhx.BhxError):
            e.record(3)                              # one list per 64-body group
        e.set_config(group_size=64)
        with pytest.raises(bhx.BhxError):
            e.record(0)
        e.record(3)
        e.set_config(theta=0.4)
        with pytest.raises(bhx.BhxError):
            e.adjoint(np.zeros((100, 3)))           # a new configuration drops the recording


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
            # the loss gradient autograd hands in can differ from the numpy expression in its last bits
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
