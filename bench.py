"""Barnes-Hut full-step throughput on one MI355X.

Workload (the precision configuration of the headline metric): Plummer sphere,
N = 1,048,576 equal masses summing to 1, G = 1, scale radius 1, seeded
virial-equilibrium velocities; theta = 0.5, Plummer softening 0.01, leaf
bucket 32, dt = 5e-4.  The walk runs in FP32 with group-relative coordinates
and quadrupole moments; positions and velocities stay FP64 with compensated
leapfrog accumulation.  A step is one full frame: bounds, Morton keys, body
sort, kick+drift, tree keys, tree sort, node build, hash table, moments,
walk, kick.

Prints ONE JSON line.  value = body-updates/s = N * steps / wall time of the
timed steps (bodies resident in HBM).  The same steps are then timed with the
FP64 walk, the force error of both walks is measured against an FP64 direct
sum on a seeded 4,096-body subsample, the walk kernel is timed with HIP events
for the roofline, and the serial CPU oracle is timed on a bounded sample of
the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "body-updates/s at N=1M theta=0.5 (FP32 vs FP64 walk); fwd+adjoint/fwd time ratio"
PEAK_FP32_TFLOPS = 157.3     # MI355X published vector FP32
PEAK_FP64_TFLOPS = 78.6      # MI355X published vector FP64
PEAK_HBM_GBS = 8000.0
FLOPS_PER_INTERACTION = 20   # softened monopole term of one reference-list entry
FLOPS_PER_QUAD_TERM = 33     # the quadrupole branch an entry of more than one body adds


def load_bhx():
    import importlib.util
    path = os.path.join(ROOT, "serial-hashed-octree-barnes-hut-n-body-simulator-0_3_amd", "bhx.py")
    spec = importlib.util.spec_from_file_location("bhx", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bhx"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_steps(e, steps, warmup):
    e.step(warmup)
    e.sync()
    t0 = time.perf_counter()
    e.step(steps)
    e.sync()
    return time.perf_counter() - t0


def walk_kernel_ms(e, reps):
    """Average walk-kernel time from HIP events recorded around the launch on the engine's stream."""
    e.set_profiling(True)
    ms = []
    for _ in range(reps):
        e.step(1)
        ms.append(e.last_timing()["walk"])
    e.set_profiling(False)
    return float(np.mean(ms)), e.last_timing()


# Algorithmic HBM bytes per body of each timed phase of the FP32 step (compensated KDK, accelerations
# permuted with the bodies, one system); r = tree nodes per body.  Each entry names the event windows
# it covers: the tree branch (tree keys, their sort) is forked onto a stream of its own before the
# Morton keys and runs beside them, the body sort and the gather + kick + drift, so those four windows
# are one phase — apart, the tree-order window would be charged the branch's bytes for only the time
# it waits on the branch.
PHASE_BYTES = {
    # Morton keys: positions in, key and index out (36); body sort: 8 passes x (key + index in and
    # out) + one key read for the histograms (200); gather + kick + drift: index, pos, vel, mass, acc,
    # both compensations in, all but acc out (236); tree keys (36) + their sort (200) + tree-order
    # gather (68)
    "keys_sorts_kick_drift_tree": (("keys", "sort", "kick_drift", "treekeys"),
                                   lambda r: 36.0 + 200.0 + 236.0 + 304.0),
    "build": (("build",), lambda r: 40.0 + 208.0 * r),      # keys and tree-order bodies; nodes, hash, moments
    # bodies in, accelerations out, each node once; the closing kick's velocity and compensation in and
    # out (the group walk applies it as it writes each acceleration)
    "walk_kick": (("walk", "kick"), lambda r: 56.0 + 32.0 * r + 96.0),
}
BOUNDS_BYTES = 72.0                         # positions read three times (two for the mean, one for the size)
PEAK_HBM_GBPS = 8000.0


def walk_flops(r):
    """Algorithmic flops of one walk: every entry of the reference's own per-body lists on this tree
    at the monopole term's count, plus the quadrupole branch's for the entries that take it."""
    return r["inter_ref"] * FLOPS_PER_INTERACTION + r["quad_ref"] * FLOPS_PER_QUAD_TERM


def phase_roofline(phases, n, nodes, step_ms):
    """Per phase: algorithmic bytes, the rate its event windows give and the fraction of peak HBM; and
    the same for the whole step (bounds included: they run on a side stream inside the step)."""
    r = nodes / n
    out = {}
    for k, (windows, f) in PHASE_BYTES.items():
        ms = sum(phases.get(w, 0.0) for w in windows)
        if ms > 0:
            gbs = f(r) * n / (ms * 1e-3) / 1e9
            out[k] = {"bytes_per_body": round(f(r), 1), "ms": round(ms, 4), "GBps": round(gbs, 1),
                      "frac": round(gbs / PEAK_HBM_GBPS, 4)}
    per_body = BOUNDS_BYTES + sum(f(r) for _, f in PHASE_BYTES.values())
    gbs = per_body * n / (step_ms * 1e-3) / 1e9
    return out, {"bytes_per_body": round(per_body, 1), "nodes_per_body": round(r, 4), "GBps": round(gbs, 1),
                 "frac": round(gbs / PEAK_HBM_GBPS, 4)}


def force_error(e, bhx, nsub, seed, eps):
    n = e.n
    rng = np.random.default_rng(seed)
    tg = np.sort(rng.choice(n, size=min(nsub, n), replace=False))
    ref, _ = e.direct(tg, eps)
    acc = e.walk()[tg]
    err = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return {"median": float(np.median(err)), "p90": float(np.percentile(err, 90)),
            "p99": float(np.percentile(err, 99)), "subsample": int(tg.size)}


def total_energy(e, eps):
    """Kinetic + softened potential energy, the potential by FP64 direct sum over all bodies."""
    ek, _, _ = e.diagnostics()
    _, _, mass, _ = e.get_bodies()
    _, pot = e.direct(np.arange(e.n), eps)
    return ek + 0.5 * float(np.dot(mass, pot))


def energy_drift(bhx, n=16384, steps=100, seed=1):
    """Relative total-energy change over `steps` frames of an N=16384 Plummer sphere
    (theta 0.5, eps 0.01, dt 1e-3): the reference's own configuration (leaf capacity 1,
    as-run payload, FP64 per-body walk) and the fast one (bucket 32, FP32 group walk with
    quadrupoles, compensated FP64 leapfrog)."""
    pos, vel, mass = bhx.plummer(n, seed=seed)
    out = {"n": n, "steps": steps, "dt": 1e-3, "theta": 0.5, "eps": 0.01}
    cfgs = {"reference_fp64": dict(theta=0.5, eps=0.01, dt=1e-3),
            "fast_fp32": dict(theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=32, mode=bhx.INTENDED,
                              precision=bhx.FP32, quadrupole=1, compensated=1, group_walk=1)}
    for name, cfg in cfgs.items():
        with bhx.Engine(n, **cfg) as e:
            e.set_bodies(pos, vel, mass)
            e0 = total_energy(e, 0.01)
            e.step_graph(steps)
            e1 = total_energy(e, 0.01)
        out[name] = {"e0": e0, "e1": e1, "rel_drift": (e1 - e0) / abs(e0)}
    return out


def galaxy_bench(bhx, n=4194304, steps=20, warmup=3, seed=2026):
    """Two-galaxy workload: a 3:1 pair of exponential-disk + Hernquist-halo galaxies on a
    parabolic orbit (separation 30, pericentre 5), FP32 group walk with quadrupoles,
    theta 0.6, eps 0.02, bucket 32, dt 1e-3, enclosing-cube cells (the secondary lies
    beyond half the bounds size).  Energy from the tree potential before and after."""
    log("galaxy pair N=%d: generating" % n)
    pos, vel, mass = bhx.galaxy_pair(n, seed=seed, threads=16)
    cfg = dict(theta=0.6, eps=0.02, dt=1e-3, leaf_capacity=32, mode=bhx.INTENDED, precision=bhx.FP32,
               quadrupole=1, compensated=1, permute_acc=1, group_walk=1, geometry=1)
    res = {"n": n, "steps": steps, "warmup": warmup, "theta": 0.6, "eps": 0.02, "dt": 1e-3, "leaf_capacity": 32,
           "geometry": "enclosing cube", "data": "synthetic: bhx_ic_galaxy_pair seed %d" % seed}
    with bhx.Engine(n, **cfg) as e:
        e.set_bodies(pos, vel, mass)
        k0, w0, vir0 = e.energy()
        t = timed_steps(e, steps, warmup)
        k1, w1, vir1 = e.energy()
        err = force_error(e, bhx, 1024, 7, 0.02)
        _, phases = walk_kernel_ms(e, 2)
    ms = 1e3 * t / steps
    res.update({"ms_per_step": ms, "body_updates_per_s": n / (ms * 1e-3), "force_error": err,
                "phases_ms": {k: round(v, 4) for k, v in phases.items()},
                "energy": {"e0": k0 + w0, "e1": k1 + w1, "rel_change": (k1 + w1 - k0 - w0) / abs(k0 + w0),
                           "virial0": vir0, "frames_between": steps + warmup}})
    log("galaxy pair %.3f ms/step, median force error %.2e" % (ms, err["median"]))
    return res


def ensemble_bench(bhx, nsys, steps, warmup, seed=1):
    """Batched ensemble (all systems in one context, one graph launch per frame) against the
    same systems stepped one after another in one context: with graph frames (like for like:
    the same frames, one system at a time) and with eager frames.  All start from the same
    state and run the same frames; final states must agree bitwise."""
    t0 = time.perf_counter()
    pos, vel, mass, off = bhx.ensemble(nsys, seed=seed)
    n = pos.shape[0]
    sizes = np.diff(off).astype(np.int64)
    log("ensemble: %d systems, %d bodies (%.1f s to generate)" % (nsys, n, time.perf_counter() - t0))
    cfg = dict(theta=0.5, eps=0.01, dt=5e-4, leaf_capacity=32, mode=bhx.INTENDED, precision=bhx.FP32,
               quadrupole=1, compensated=1, group_walk=1)
    res = {"systems": nsys, "bodies": n, "sizes": {"min": int(sizes.min()), "max": int(sizes.max())},
           "theta": 0.5, "eps": 0.01, "dt": 5e-4, "warmup": warmup, "steps": steps}
    with bhx.Engine(n, **cfg) as e:
        e.set_bodies(pos, vel, mass)
        e.set_systems(off)
        chosen, gchosen, table = e.autotune(caps=(16, 32, 64), group_sizes=(64, 32), steps=2, max_err=1e-3)
        res["autotune"] = {"chosen_leaf_capacity": chosen, "chosen_group_size": gchosen,
                           "candidates": [{"leaf_capacity": c, "group_size": g, "ms_per_step": ms, "median_err": er}
                                          for c, g, ms, er in table]}
        cfg["leaf_capacity"] = chosen
        cfg["group_size"] = gchosen
        log("ensemble autotune: leaf capacity %d, group size %d %s" % (chosen, gchosen, table))
        e.step_graph(warmup)
        e.sync()
        t0 = time.perf_counter()
        e.step_graph(steps)
        e.sync()
        tb = time.perf_counter() - t0
        P, V, M, A = e.get_bodies()
    log("ensemble batched: %.2f ms/step" % (1e3 * tb / steps))
    looped = {}
    with bhx.Engine(int(sizes.max()), **cfg) as s:
        for mode in ("graph", "eager"):
            tl = 0.0
            same = True
            for k in range(nsys):
                a, b = int(off[k]), int(off[k + 1])
                s.set_bodies(pos[a:b], vel[a:b], mass[a:b])
                run = s.step_graph if mode == "graph" else s.step
                run(warmup)                  # graph frames: captured here, replayed below
                s.sync()
                t0 = time.perf_counter()
                run(steps)
                s.sync()
                tl += time.perf_counter() - t0
                p1, v1, m1, a1 = s.get_bodies()
                same = same and np.array_equal(P[a:b], p1) and np.array_equal(V[a:b], v1) and \
                    np.array_equal(A[a:b], a1)
            looped[mode] = (tl, same)
            log("ensemble looped (%s frames): %.2f ms/step, bitwise equal: %s" % (mode, 1e3 * tl / steps, same))
    tg, same_g = looped["graph"]
    te, same_e = looped["eager"]
    res.update({"batched_ms_per_step": 1e3 * tb / steps, "looped_ms_per_step": 1e3 * tg / steps,
                "looped_eager_ms_per_step": 1e3 * te / steps,
                "batched_body_updates_per_s": n * steps / tb, "looped_body_updates_per_s": n * steps / tg,
                "batched_over_looped": tg / tb, "batched_over_looped_eager": te / tb,
                "bitwise_equal_to_looped": bool(same_g and same_e),
                "looped": "each system alone in one context, bhx_step_graph per system (graphs captured in the "
                          "warm-up frames; one host wait per system)",
                "looped_eager": "each system alone in one context, bhx_step per system (eager frames)",
                "batched": "all systems in one context (bhx_set_systems), bhx_step_graph"})
    return res


def adjoint_bench(bhx, n=65536, steps=1000, every=100, seed=5, quadrupole=0):
    """Forward + reverse pass over a recorded trajectory against the forward alone.
    Loss: mean |x_T - x_T*|^2, the target trajectory from initial conditions perturbed by
    1e-3 (positions, velocities) and 1% (masses).  FP64 state, FP32 walk, monopole or
    quadrupole cells."""
    pos, vel, mass = bhx.plummer(n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    pos_t = pos + 1e-3 * rng.normal(size=pos.shape)
    vel_t = vel + 1e-3 * rng.normal(size=vel.shape)
    mass_t = mass * (1.0 + 0.01 * rng.normal(size=mass.shape))
    cfg = dict(theta=0.5, eps=0.01, dt=1e-3, leaf_capacity=32, mode=bhx.INTENDED, precision=bhx.FP32,
               quadrupole=quadrupole, compensated=1, group_walk=1)
    res = {"n": n, "steps": steps, "checkpoint_every": every, "theta": 0.5, "eps": 0.01, "dt": 1e-3,
           "walk": "FP32 group walk, " + ("quadrupole" if quadrupole else "monopole"),
           "loss": "mean |x_T - x_T*|^2"}
    with bhx.Engine(n, **cfg) as e:
        e.set_bodies(pos_t, vel_t, mass_t)
        e.record(steps, every)
        target = e.recorded_bodies()[0]
        e.set_bodies(pos, vel, mass)
        e.step_graph(5)                      # capture and warm the frame graphs
        e.set_bodies(pos, vel, mass)
        e.sync()
        t0 = time.perf_counter()
        e.step_graph(steps)
        e.sync()
        t_fwd = time.perf_counter() - t0
        e.set_bodies(pos, vel, mass)
        t0 = time.perf_counter()
        e.record(steps, every)
        xT = e.recorded_bodies()[0]
        t_rec = time.perf_counter() - t0
        loss = float(np.mean(np.sum((xT - target) ** 2, axis=1)))
        gx = 2.0 * (xT - target) / n
        t0 = time.perf_counter()
        gx0, gv0, _, gm = e.adjoint(gx)
        t_adj = time.perf_counter() - t0
    log("adjoint (quadrupole=%d): forward %.3f s, recorded forward %.3f s, reverse %.3f s"
        % (quadrupole, t_fwd, t_rec, t_adj))
    res.update({"forward_s": t_fwd, "recorded_forward_s": t_rec, "adjoint_s": t_adj, "loss_value": loss,
                "grad_norm_x0": float(np.linalg.norm(gx0)), "grad_norm_m": float(np.linalg.norm(gm)),
                "fwd_adjoint_over_fwd": (t_rec + t_adj) / t_fwd,
                "forward": "bhx_step_graph, no recording",
                "fwd_adjoint": "bhx_record (checkpoints) + bhx_adjoint (segments re-run, adjoint walks)"})
    return res


def _oracle_cflags():
    """The oracle's compile flags (oracle/Makefile)."""
    try:
        for line in open(os.path.join(ROOT, "oracle", "Makefile")):
            if line.startswith("CFLAGS"):
                return line.split("=", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pos, vel, mass, theta, eps, dt, cap):
    """Serial oracle (oracle/liboracle.so) for one full step of the workload on ONE core
    of this host (the process is pinned to a single CPU for the duration): bounds, keys,
    sort, kick/drift, bucket-32 build with intended moments, the quadrupole walk for every
    body (no sample, no extrapolation) and the closing kick."""
    from oracle import oracle as orc
    n = pos.shape[0]
    allowed = sorted(os.sched_getaffinity(0))
    core = allowed[0]
    os.sched_setaffinity(0, {core})
    try:
        t = {}
        t0 = time.perf_counter()
        c, s = orc.bounds(pos)
        t["bounds"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        keys = orc.morton_keys(pos, s)
        t["keys"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, idx = orc.radix_sort(keys)
        idx = idx.astype(np.int64)
        p, v, m = pos[idx], vel[idx], mass[idx]
        t["sort"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        p, v = orc.kick_drift(dt, p, v, np.zeros_like(p))
        t["kick_drift"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        tree = orc.build_recursive(p, m, c, s, cap)
        tree.moments_intended()
        t["build"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        acc, cnt = tree.walk(theta, eps, quad=True)
        t["walk"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        orc.kick(dt, v, acc)
        t["kick"] = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, set(allowed))
    step = sum(t.values())
    return {
        "value": n / step, "unit": "body-updates/s", "cores": 1, "kind": "port",
        "sample": "one full step of all %d bodies (no sampling): serial oracle, FP64, the walk with the "
                  "reference's per-body lists and quadrupoles" % n,
        "step_s": round(step, 3),
        "phase_s": {k: round(v_, 4) for k, v_ in t.items()},
        "cpu": _cpu_model(), "pinned_core": core, "nproc": os.cpu_count(), "affinity_cpus": len(allowed),
        "compiler_flags": "gcc " + _oracle_cflags(),
        "mean_interactions": float(np.mean(cnt)),
    }


def c2_error_series(bhx, pos, vel, mass, cfg, steps=50, nsub=4096, seed=4096):
    """Force error of the FP32 walk after each of `steps` frames of the precision workload,
    against the FP64 direct sum on a seeded sample of `nsub` array positions."""
    n = pos.shape[0]
    tg = np.sort(np.random.default_rng(seed).choice(n, size=nsub, replace=False))
    series = []
    with bhx.Engine(n, **cfg) as e:
        e.set_bodies(pos, vel, mass)
        for _ in range(steps):
            e.step_graph(1)
            ref, _ = e.direct(tg, cfg["eps"])
            acc = e.get_bodies()[3][tg]
            err = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
            series.append([float(np.median(err)), float(np.percentile(err, 90)), float(np.percentile(err, 99))])
    a = np.array(series)
    return {"steps": steps, "subsample": nsub, "columns": ["median", "p90", "p99"], "per_step": series,
            "max_median": float(a[:, 0].max()), "max_p99": float(a[:, 2].max())}


def copy_calibration(mib):
    """One device-to-device hipMemcpy of `mib` MiB (the runtime's copy kernel), so that a
    FETCH_SIZE / WRITE_SIZE counter pass can be scaled against a transfer of known size."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    rc = hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    if rc:
        raise RuntimeError("hipMemcpy failed: %d" % rc)
    log("copy calibration: %d MiB device-to-device" % mib)


def read_traffic(n):
    """HBM bytes per walk launch from the committed counter summary for this workload, if any."""
    path = os.path.join(ROOT, "profiles", "walk_traffic_fp32_n%d.txt" % n)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            try:
                vals[k.strip()] = float(v.split()[0])
            except (ValueError, IndexError):
                pass
    if "hbm_bytes_per_launch" not in vals:
        return None
    return vals["hbm_bytes_per_launch"]


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error-series", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-systems", type=int, default=512)
    ap.add_argument("--no-drift", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-galaxy", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (counter calibration runs)")
    ap.add_argument("--adjoint-only", type=int, default=0, metavar="FRAMES",
                    help="only the differentiation leg, monopole then quadrupole cells, over FRAMES frames "
                         "(kernel profiles of the reverse pass); prints one JSON line of its figures")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the whole Barnes-Hut step on a single MI355X "
              "(no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    bhx = load_bhx()
    if args.calibrate_copy_mib:
        copy_calibration(args.calibrate_copy_mib)
    if args.adjoint_only:
        res = {"adjoint_frames": args.adjoint_only}
        for q, key in ((0, "adjoint"), (1, "adjoint_quadrupole")):
            log("adjoint: %d frames, quadrupole %d" % (args.adjoint_only, q))
            res[key] = adjoint_bench(bhx, steps=args.adjoint_only, quadrupole=q)
        print(json.dumps(res), flush=True)
        return
    theta, eps, dt, cap = 0.5, 0.01, 5e-4, 32
    n = args.n
    log("generating Plummer N=%d seed=%d" % (n, args.seed))
    pos, vel, mass = bhx.plummer(n, seed=args.seed)
    cfg = dict(theta=theta, eps=eps, dt=dt, leaf_capacity=cap, mode=bhx.INTENDED, precision=bhx.FP32,
               quadrupole=1, compensated=1, permute_acc=1, group_walk=1)
    out = {}
    with bhx.Engine(n, **cfg) as e:
        e.set_bodies(pos, vel, mass)
        log("fp32: warmup %d, timing %d steps" % (args.warmup, args.steps))
        t32 = timed_steps(e, args.steps, args.warmup)
        ms32 = 1e3 * t32 / args.steps
        log("fp32 %.3f ms/step" % ms32)
        walk_ms, phases = walk_kernel_ms(e, 3)
        nodes = len(e.nodes())
        # algorithmic work: the reference's own per-body interaction lists on this tree;
        # the extra entries of the conservative group lists count as overhead, not work
        cnt_ref, qc_ref = e.interaction_counts(quad=True)
        inter_ref = float(cnt_ref.astype(np.float64).sum())
        quad_ref = float(qc_ref.astype(np.float64).sum())
        acc, cnt = e.walk(counts=True)
        inter = float(cnt.astype(np.float64).sum())
        err32 = force_error(e, bhx, 4096, 1234, eps)
        out["fp32"] = dict(ms=ms32, walk_ms=walk_ms, phases=phases, inter=inter, inter_ref=inter_ref,
                           quad_ref=quad_ref, err=err32, nodes=nodes)
        if not args.no_fp64:
            e.set_config(precision=bhx.FP64)
            log("fp64: warmup %d, timing %d steps" % (args.warmup, args.steps))
            t64 = timed_steps(e, args.steps, args.warmup)
            walk64_ms, phases64 = walk_kernel_ms(e, 3)
            cnt64_ref, qc64_ref = e.interaction_counts(quad=True)
            inter64_ref = float(cnt64_ref.astype(np.float64).sum())
            quad64_ref = float(qc64_ref.astype(np.float64).sum())
            _, cnt64 = e.walk(counts=True)
            err64 = force_error(e, bhx, 4096, 1234, eps)
            out["fp64"] = dict(ms=1e3 * t64 / args.steps, walk_ms=walk64_ms, phases=phases64,
                               inter=float(cnt64.astype(np.float64).sum()), inter_ref=inter64_ref,
                               quad_ref=quad64_ref, err=err64)
    value = n * args.steps / t32
    work32 = walk_flops(out["fp32"])
    achieved = work32 / (walk_ms * 1e-3) / 1e12
    achieved_mono = out["fp32"]["inter_ref"] * FLOPS_PER_INTERACTION / (walk_ms * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "body-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Plummer sphere (Philox4x32-10, seed %d) generated on the host" % args.seed,
        "config": {"workload": "Plummer N=%d, theta=%g, eps=%g, leaf bucket %d, FP32 group-relative walk with "
                               "quadrupoles, FP64 compensated KDK, dt=%g; full step = bounds+keys+sort+kick/drift"
                               "+tree keys+sort+build+hash+moments+walk+kick" % (n, theta, eps, cap, dt),
                   "n": n, "theta": theta, "eps": eps, "leaf_capacity": cap, "dt": dt},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": read_traffic(n),
                     "kernel": "k_walk_lp32<true, false, false, 1, false>", "kernel_ms": walk_ms,
                     "work": "%d reference-list interactions x %d flops + %d of them with the quadrupole term "
                             "x %d (group lists evaluate %d)" % (
                                 out["fp32"]["inter_ref"], FLOPS_PER_INTERACTION, out["fp32"]["quad_ref"],
                                 FLOPS_PER_QUAD_TERM, out["fp32"]["inter"]),
                     "achieved_monopole_only": achieved_mono, "frac_monopole_only": achieved_mono / PEAK_FP32_TFLOPS},
        "force_error_fp32": out["fp32"]["err"],
        "phases_ms_fp32": {k: round(v, 4) for k, v in out["fp32"]["phases"].items()},
    }
    line["phase_roofline_fp32"], line["step_hbm_fp32"] = phase_roofline(out["fp32"]["phases"], n,
                                                                        out["fp32"]["nodes"], ms32)
    if "fp64" in out:
        f64 = out["fp64"]
        ach64 = walk_flops(f64) / (f64["walk_ms"] * 1e-3) / 1e12
        line["fp64_walk"] = {"value": n * args.steps / (f64["ms"] * 1e-3 * args.steps), "ms_per_step": f64["ms"],
                             "walk_kernel_ms": f64["walk_ms"], "force_error": f64["err"],
                             "achieved_tflops": ach64, "frac_fp64_peak": ach64 / PEAK_FP64_TFLOPS,
                             "phases_ms": {k: round(v, 4) for k, v in f64["phases"].items()}}
        line["fp32_over_fp64_walk"] = f64["walk_ms"] / walk_ms
    if not args.no_error_series:
        log("force-error series over 50 steps")
        line["force_error_series_fp32"] = c2_error_series(bhx, pos, vel, mass, cfg)
    if not args.no_ensemble:
        line["ensemble"] = ensemble_bench(bhx, args.ensemble_systems, 10, 2)
    if not args.no_galaxy:
        line["galaxy_pair"] = galaxy_bench(bhx)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_bench(bhx)
        line["fwd_adjoint_over_fwd"] = line["adjoint"]["fwd_adjoint_over_fwd"]
        line["adjoint_quadrupole"] = adjoint_bench(bhx, quadrupole=1)
    if not args.no_drift:
        log("energy drift (N=16384, 100 steps)")
        line["energy_drift"] = energy_drift(bhx)
    if not args.no_cpu_baseline:
        log("cpu baseline (serial oracle, one full step, one core)")
        line["cpu_baseline"] = cpu_baseline(pos, vel, mass, theta, eps, dt, cap)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
