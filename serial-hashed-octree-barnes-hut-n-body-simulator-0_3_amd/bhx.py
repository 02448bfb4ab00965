"""Python binding of libbhx.so (the MI355X Barnes-Hut engine) over its C ABI.

A thin ctypes layer: every method calls one entry point of include/bhx.h and
raises BhxError on a non-zero status.  There is no CPU fallback: if the
library has not been built, loading raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbhx.so")

AS_RUN, INTENDED = 0, 1
FP64, FP32 = 0, 1

STATUS = {
    0: "BHX_OK", 1: "BHX_E_NULL", 2: "BHX_E_ARG", 3: "BHX_E_DEPTH", 4: "BHX_E_HASH_FULL",
    5: "BHX_E_NONFINITE", 6: "BHX_E_HIP", 7: "BHX_E_NOMEM", 8: "BHX_E_STATE", 9: "BHX_E_CAPACITY",
    10: "BHX_E_SORT", 11: "BHX_E_IO",
}

NODE_DTYPE = np.dtype([
    ("key", "<u8"), ("center", "<f8", (3,)), ("size", "<f8"), ("com", "<f8", (3,)), ("mass", "<f8"),
    ("quad", "<f8", (6,)), ("n", "<i8"), ("first", "<i8"), ("level", "<i4"), ("child_byte", "<u4"),
    ("skip", "<i8"), ("body_start", "<i8"),
])
assert NODE_DTYPE.itemsize == 160
TUNE_DTYPE = np.dtype([("leaf_capacity", "<i4"), ("group_size", "<i4"), ("ms_per_step", "<f8"),
                       ("median_err", "<f8")])


class BhxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("theta", ctypes.c_double), ("eps", ctypes.c_double), ("dt", ctypes.c_double),
                ("leaf_capacity", ctypes.c_int32), ("mode", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("quadrupole", ctypes.c_int32), ("permute_acc", ctypes.c_int32), ("compensated", ctypes.c_int32),
                ("device", ctypes.c_int32), ("group_walk", ctypes.c_int32), ("geometry", ctypes.c_int32),
                ("group_size", ctypes.c_int32)]


class Timing(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("bounds", "keys", "sort", "kick_drift", "treekeys", "build", "moments", "walk", "kick", "total")]


_lib = None


def lib():
    """Load libbhx.so; raises when it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libbhx.so not built: run build.py (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        P, S, D, I = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double, ctypes.c_int
        sigs = {
            "bhx_config_default": [P], "bhx_create": [P, S, P], "bhx_destroy": [P], "bhx_set_config": [P, P],
            "bhx_get_config": [P, P], "bhx_sync": [P], "bhx_set_bodies": [P, P, P, P, S],
            "bhx_set_accelerations": [P, P], "bhx_get_bodies": [P, P, P, P, P], "bhx_num_bodies": [P, P],
            "bhx_bounds": [P, P, P], "bhx_morton_keys": [P, D, P], "bhx_sort_keys": [P, P, P, S],
            "bhx_sort_bodies": [P, P], "bhx_reorder_bodies": [P, P], "bhx_kick_drift": [P, D], "bhx_kick": [P, D],
            "bhx_build": [P, P, D, P], "bhx_get_nodes": [P, P, S, P], "bhx_lookup": [P, P, S, P],
            "bhx_get_tree_order": [P, P, P], "bhx_unsplit_leaves": [P, P], "bhx_walk": [P, P, P],
            "bhx_interaction_counts": [P, P], "bhx_interaction_counts_quad": [P, P, P],
            "bhx_step": [P, I], "bhx_last_timing": [P, P], "bhx_set_profiling": [P, I],
            "bhx_direct": [P, P, S, D, P, P], "bhx_diagnostics": [P, P, P, P],
            "bhx_ic_plummer": [S, ctypes.c_uint64, D, D, P, P, P],
            "bhx_ic_ensemble_sizes": [S, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, D, P, P, P],
            "bhx_ic_ensemble": [S, ctypes.c_uint64, P, P, P, I, P, P, P],
            "bhx_set_systems": [P, P, S], "bhx_num_systems": [P, P], "bhx_system_bounds": [P, P, P],
            "bhx_build_systems": [P, P, P, P], "bhx_step_graph": [P, I],
            "bhx_autotune": [P, P, S, P, S, I, D, S, ctypes.c_uint64, P, P, P],
            "bhx_record": [P, I, I, I], "bhx_recorded_bodies": [P, P, P, P, P], "bhx_record_digest": [P, P],
            "bhx_adjoint": [P, P, P, P, P, P, P, P], "bhx_grad_timing": [P, P, P],
            "bhx_set_record_memory": [P, ctypes.c_uint64], "bhx_record_keeps_frames": [P, P],
            "bhx_set_adjoint_list_capacity": [P, ctypes.c_uint32], "bhx_system_roots": [P, P], "bhx_lookup_in": [P, S, P, S, P],
            "bhx_get_systems": [P, P], "bhx_energy": [P, P, P, P, P],
            "bhx_galaxy_defaults": [P], "bhx_ic_galaxy_pair": [S, ctypes.c_uint64, P, I, P, P, P],
            "bhx_snapshot_write": [ctypes.c_char_p, P, P, P, P, P, P, P],
            "bhx_snapshot_read_info": [ctypes.c_char_p, P],
            "bhx_snapshot_read": [ctypes.c_char_p, P, P, P, P, P, P],
            "bhx_get_compensation": [P, P, P], "bhx_set_compensation": [P, P, P],
            "bhx_save": [P, ctypes.c_char_p, D, ctypes.c_uint64], "bhx_load": [P, ctypes.c_char_p, P, P],
            "bhx_vtk_write_bodies": [ctypes.c_char_p, P, P, P, S],
            "bhx_vtk_write_nodes": [ctypes.c_char_p, P, S, I, I],
            "bhx_export_vtk": [P, ctypes.c_char_p, ctypes.c_char_p, I, I],
            "bhx_bounds_path": [P, P],
            "bhx_sort_keys_noalloc": [I, P, P, P, S],
            "bhx_interaction_lists": [P, P, S, D, P, P, S, P], "bhx_force_lists": [P, P, S, P, P, P],
            "bhx_mac": [P, P, P, S, D, P], "bhx_node_put": [P, P, P], "bhx_node_erase": [P, ctypes.c_uint64, P],
            "bhx_prune_empty": [P, ctypes.c_uint64, P], "bhx_decode_morton_keys": [P, P, S, D, P],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.bhx_last_error.argtypes = [P]
        L.bhx_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _f64_exact(a, shape, what):
    """float64 C-contiguous copy of `a`, which must have exactly `shape` (the C side
    reads that many values through the pointer)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != tuple(shape):
        raise ValueError("%s: expected shape %s, got %s" % (what, tuple(shape), a.shape))
    return a


def plummer(n, seed=1, total_mass=1.0, a=1.0):
    """Seeded Plummer sphere (host generator of the library): pos (n,3), vel (n,3), mass (n,)."""
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    mass = np.empty(n)
    rc = lib().bhx_ic_plummer(n, seed, total_mass, a, _p(pos), _p(vel), _p(mass))
    if rc:
        raise BhxError(rc, "bhx_ic_plummer")
    return pos, vel, mass


def ensemble(nsys, seed=1, nmin=1024, nmax=65536, jitter=0.2, threads=0):
    """Seeded ensemble of Plummer systems (host generator of the library).
    Returns pos (N,3), vel (N,3), mass (N,), offsets (nsys+1,) with N = sum of the sizes."""
    sizes = np.empty(nsys, np.uint64)
    ms = np.empty(nsys)
    rs = np.empty(nsys)
    rc = lib().bhx_ic_ensemble_sizes(nsys, seed, nmin, nmax, jitter, _p(sizes), _p(ms), _p(rs))
    if rc:
        raise BhxError(rc, "bhx_ic_ensemble_sizes")
    n = int(sizes.sum())
    pos, vel, mass = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    rc = lib().bhx_ic_ensemble(nsys, seed, _p(sizes), _p(ms), _p(rs), threads, _p(pos), _p(vel), _p(mass))
    if rc:
        raise BhxError(rc, "bhx_ic_ensemble")
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return pos, vel, mass, off


class GalaxyParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("total_mass", "mass_ratio", "disk_fraction", "disk_scale", "disk_height", "disk_cut",
                 "halo_scale", "halo_cut", "disk_sigma", "separation", "pericentre", "incl1", "incl2")]


def galaxy_params(**kw):
    """bhx_galaxy_params with the library's defaults, then the given fields."""
    p = GalaxyParams()
    lib().bhx_galaxy_defaults(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def galaxy_pair(n, seed=1, threads=0, **params):
    """Two disk + Hernquist-halo galaxies on a parabolic orbit (host generator of the
    library): pos (n,3), vel (n,3), mass (n,); the primary's bodies first (halo, disk)."""
    p = galaxy_params(**params)
    pos, vel, mass = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    rc = lib().bhx_ic_galaxy_pair(n, seed, ctypes.byref(p), threads, _p(pos), _p(vel), _p(mass))
    if rc:
        raise BhxError(rc, "bhx_ic_galaxy_pair")
    return pos, vel, mass


class SnapshotInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint64), ("nsys", ctypes.c_uint64), ("time", ctypes.c_double),
                ("step", ctypes.c_uint64), ("has_acc", ctypes.c_int32), ("has_comp", ctypes.c_int32),
                ("cfg", Config)]


def _path(path):
    return os.fsencode(path)


def write_snapshot(path, pos, vel, mass, acc=None, offsets=None, time=0.0, step=0, config=None, comp=None):
    """Write host arrays as a snapshot file (see bhx_snapshot_write); comp = (6n,)
    or (2, n, 3): position then velocity compensation terms."""
    mass = _f64(mass)
    n = mass.size
    if mass.ndim != 1:
        raise ValueError("mass: expected shape (n,), got %s" % (mass.shape,))
    pos = _f64_exact(pos, (n, 3), "pos")
    vel = _f64_exact(vel, (n, 3), "vel")
    acc = None if acc is None else _f64_exact(acc, (n, 3), "acc")
    comp = None if comp is None else _f64(comp).reshape(-1)
    if comp is not None and comp.size != 6 * n:
        raise ValueError("comp: expected 6n = %d values, got %d" % (6 * n, comp.size))
    off = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.uint64)
    info = SnapshotInfo()
    info.n = mass.size
    info.nsys = 1 if off is None else off.size - 1
    info.time, info.step = time, step
    info.cfg = config if config is not None else default_config()
    rc = lib().bhx_snapshot_write(_path(path), ctypes.byref(info), _p(pos), _p(vel), _p(mass), _p(acc), _p(comp),
                                  _p(off))
    if rc:
        raise BhxError(rc, "bhx_snapshot_write")


def read_snapshot(path):
    """Contents of a snapshot file: dict with pos, vel, mass, acc and comp (None if
    absent; comp shaped (2, n, 3)), offsets, time, step and config (a Config)."""
    info = SnapshotInfo()
    rc = lib().bhx_snapshot_read_info(_path(path), ctypes.byref(info))
    if rc:
        raise BhxError(rc, "bhx_snapshot_read_info")
    n = int(info.n)
    pos, vel, mass = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    acc = np.empty((n, 3)) if info.has_acc else None
    comp = np.empty((2, n, 3)) if info.has_comp else None
    off = np.empty(int(info.nsys) + 1, np.uint64)
    rc = lib().bhx_snapshot_read(_path(path), _p(pos), _p(vel), _p(mass), _p(acc), _p(comp), _p(off))
    if rc:
        raise BhxError(rc, "bhx_snapshot_read")
    return dict(pos=pos, vel=vel, mass=mass, acc=acc, comp=comp, offsets=off, time=info.time,
                step=int(info.step), config=info.cfg)


def write_vtk_bodies(path, pos, vel=None, mass=None):
    pos = _f64(pos)
    vel = None if vel is None else _f64(vel)
    mass = None if mass is None else _f64(mass)
    rc = lib().bhx_vtk_write_bodies(_path(path), _p(pos), _p(vel), _p(mass), pos.shape[0])
    if rc:
        raise BhxError(rc, "bhx_vtk_write_bodies")


def write_vtk_nodes(path, nodes, min_level=0, max_level=21):
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    rc = lib().bhx_vtk_write_nodes(_path(path), _p(nodes), nodes.size, min_level, max_level)
    if rc:
        raise BhxError(rc, "bhx_vtk_write_nodes")


def default_config(**kw):
    cfg = Config()
    lib().bhx_config_default(ctypes.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


class Engine:
    """One bhx context: bodies resident on the GPU, the per-step functions as methods."""

    def __init__(self, capacity, **cfg):
        self.cfg = default_config(**cfg)
        h = ctypes.c_void_p()
        rc = lib().bhx_create(ctypes.byref(self.cfg), capacity, ctypes.byref(h))
        if rc:
            raise BhxError(rc, "bhx_create")
        self.h = h
        self.capacity = capacity
        # bumped by every record() (and set_bodies()): a backward pass checks that the
        # recording it differentiates is still the one its forward pass made
        self.record_generation = 0

    def close(self):
        if getattr(self, "h", None):
            lib().bhx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc:
            msg = lib().bhx_last_error(self.h) if self.h else b""
            raise BhxError(rc, "%s (%s)" % (what, (msg or b"").decode()))

    # --- configuration
    def set_config(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.cfg, k):
                raise KeyError(k)
            setattr(self.cfg, k, v)
        self._chk(lib().bhx_set_config(self.h, ctypes.byref(self.cfg)), "bhx_set_config")

    # --- bodies
    @property
    def n(self):
        v = ctypes.c_size_t()
        self._chk(lib().bhx_num_bodies(self.h, ctypes.byref(v)), "bhx_num_bodies")
        return v.value

    def set_bodies(self, pos, vel, mass):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError("pos: expected shape (n, 3), got %s" % (pos.shape,))
        n = pos.shape[0]
        mass = _f64_exact(mass, (n,), "mass")
        vel = None if vel is None else _f64_exact(vel, (n, 3), "vel")
        self._chk(lib().bhx_set_bodies(self.h, _p(pos), _p(vel), _p(mass), n), "bhx_set_bodies")
        self.record_generation += 1    # the recording (if any) is of other bodies now

    def set_accelerations(self, acc):
        acc = _f64_exact(acc, (self.n, 3), "acc")
        self._chk(lib().bhx_set_accelerations(self.h, _p(acc)), "bhx_set_accelerations")

    def get_bodies(self):
        n = self.n
        pos, vel, acc, mass = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
        self._chk(lib().bhx_get_bodies(self.h, _p(pos), _p(vel), _p(mass), _p(acc)), "bhx_get_bodies")
        return pos, vel, mass, acc

    # --- the step functions
    def bounds(self):
        c = np.zeros(3)
        s = np.zeros(1)
        self._chk(lib().bhx_bounds(self.h, _p(c), _p(s)), "bhx_bounds")
        return c, float(s[0])

    def morton_keys(self, size):
        keys = np.empty(self.n, np.uint64)
        self._chk(lib().bhx_morton_keys(self.h, size, _p(keys)), "bhx_morton_keys")
        return keys

    def sort_keys(self, keys, idx=None):
        keys = np.array(keys, dtype=np.uint64, copy=True)
        idx = np.arange(keys.size, dtype=np.uint64) if idx is None else np.array(idx, np.uint64, copy=True)
        self._chk(lib().bhx_sort_keys(self.h, _p(keys), _p(idx), keys.size), "bhx_sort_keys")
        return keys, idx

    def sort_bodies(self):
        perm = np.empty(self.n, np.uint64)
        self._chk(lib().bhx_sort_bodies(self.h, _p(perm)), "bhx_sort_bodies")
        return perm

    def reorder_bodies(self, perm):
        perm = np.ascontiguousarray(perm, dtype=np.uint64)
        self._chk(lib().bhx_reorder_bodies(self.h, _p(perm)), "bhx_reorder_bodies")

    def kick_drift(self, dt):
        self._chk(lib().bhx_kick_drift(self.h, dt), "bhx_kick_drift")

    def kick(self, dt):
        self._chk(lib().bhx_kick(self.h, dt), "bhx_kick")

    def build(self, center, size):
        c = _f64(center)
        nn = ctypes.c_size_t()
        self._chk(lib().bhx_build(self.h, _p(c), size, ctypes.byref(nn)), "bhx_build")
        return nn.value

    def nodes(self):
        nn = ctypes.c_size_t()
        self._chk(lib().bhx_get_nodes(self.h, None, 0, ctypes.byref(nn)), "bhx_get_nodes")
        out = np.zeros(nn.value, NODE_DTYPE)
        self._chk(lib().bhx_get_nodes(self.h, _p(out), nn.value, ctypes.byref(nn)), "bhx_get_nodes")
        return out

    def lookup(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        ids = np.empty(keys.size, np.int64)
        self._chk(lib().bhx_lookup(self.h, _p(keys), keys.size, _p(ids)), "bhx_lookup")
        return ids

    # --- ensembles of independent systems
    def set_systems(self, offsets):
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        self._chk(lib().bhx_set_systems(self.h, _p(off), off.size - 1), "bhx_set_systems")

    @property
    def nsys(self):
        v = ctypes.c_size_t()
        self._chk(lib().bhx_num_systems(self.h, ctypes.byref(v)), "bhx_num_systems")
        return v.value

    def system_bounds(self):
        k = self.nsys
        c = np.empty((k, 3))
        sz = np.empty(k)
        self._chk(lib().bhx_system_bounds(self.h, _p(c), _p(sz)), "bhx_system_bounds")
        return c, sz

    def build_systems(self, centers, sizes):
        c = _f64(centers)
        sz = _f64(sizes)
        nn = ctypes.c_size_t()
        self._chk(lib().bhx_build_systems(self.h, _p(c), _p(sz), ctypes.byref(nn)), "bhx_build_systems")
        return nn.value

    def system_roots(self):
        r = np.empty(self.nsys, np.uint64)
        self._chk(lib().bhx_system_roots(self.h, _p(r)), "bhx_system_roots")
        return r

    def lookup_in(self, system, keys):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        ids = np.empty(keys.size, np.int64)
        self._chk(lib().bhx_lookup_in(self.h, system, _p(keys), keys.size, _p(ids)), "bhx_lookup_in")
        return ids

    def tree_order(self):
        perm = np.empty(self.n, np.uint64)
        keys = np.empty(self.n, np.uint64)
        self._chk(lib().bhx_get_tree_order(self.h, _p(perm), _p(keys)), "bhx_get_tree_order")
        return perm, keys

    def unsplit_leaves(self):
        v = ctypes.c_size_t()
        self._chk(lib().bhx_unsplit_leaves(self.h, ctypes.byref(v)), "bhx_unsplit_leaves")
        return v.value

    def walk(self, counts=False):
        n = self.n
        acc = np.empty((n, 3))
        cnt = np.empty(n, np.uint64) if counts else None
        self._chk(lib().bhx_walk(self.h, _p(acc), _p(cnt)), "bhx_walk")
        return (acc, cnt) if counts else acc

    def interaction_counts(self, quad=False):
        """Reference per-body interaction-list lengths on the last tree (accelerations untouched);
        with quad, also how many entries of each list take the quadrupole branch."""
        cnt = np.empty(self.n, np.uint64)
        if not quad:
            self._chk(lib().bhx_interaction_counts(self.h, _p(cnt)), "bhx_interaction_counts")
            return cnt
        qc = np.empty(self.n, np.uint64)
        self._chk(lib().bhx_interaction_counts_quad(self.h, _p(cnt), _p(qc)), "bhx_interaction_counts_quad")
        return cnt, qc

    def step(self, nsteps=1):
        self._chk(lib().bhx_step(self.h, nsteps), "bhx_step")

    def step_graph(self, nsteps=1):
        self._chk(lib().bhx_step_graph(self.h, nsteps), "bhx_step_graph")

    def autotune(self, caps=(8, 16, 24, 32, 48, 64), steps=3, max_err=1e-3, nsample=4096, seed=7, group_sizes=None):
        """Walk tiling with the fastest graph frames at median force error <= max_err: leaf
        capacity and (with group_sizes) bodies per interaction list, set on the context
        (bodies left as they were).  Returns (chosen_capacity, [(cap, ms_per_step,
        median_err)]); with group_sizes, (chosen_capacity, chosen_group_size,
        [(cap, group_size, ms_per_step, median_err)])."""
        caps = np.ascontiguousarray(caps, dtype=np.int32)
        gs = None if group_sizes is None else np.ascontiguousarray(group_sizes, dtype=np.int32)
        ng = 1 if gs is None else gs.size
        out = np.zeros(caps.size * ng, TUNE_DTYPE)
        ch, chg = ctypes.c_int32(), ctypes.c_int32()
        self._chk(lib().bhx_autotune(self.h, _p(caps), caps.size, _p(gs), 0 if gs is None else gs.size, steps,
                                     max_err, nsample, seed, _p(out), ctypes.byref(ch), ctypes.byref(chg)),
                  "bhx_autotune")
        self.cfg.leaf_capacity = ch.value
        self.cfg.group_size = chg.value
        if gs is None:
            return ch.value, [(int(r["leaf_capacity"]), float(r["ms_per_step"]), float(r["median_err"]))
                              for r in out]
        return ch.value, chg.value, [(int(r["leaf_capacity"]), int(r["group_size"]), float(r["ms_per_step"]),
                                      float(r["median_err"])) for r in out]

    # --- differentiation
    def record(self, nsteps, checkpoint_every=100, init_acc=True):
        """Run and record nsteps frames for adjoint(); see bhx_record."""
        self.record_generation += 1
        self._chk(lib().bhx_record(self.h, nsteps, checkpoint_every, int(init_acc)), "bhx_record")

    def set_record_memory(self, nbytes):
        """Device memory record() may use to keep every frame (0: none, re-run the checkpoint segments;
        None: half of the free device memory)."""
        v = (1 << 64) - 1 if nbytes is None else int(nbytes)
        self._chk(lib().bhx_set_record_memory(self.h, ctypes.c_uint64(v)), "bhx_set_record_memory")

    def set_adjoint_list_capacity(self, entries):
        """List entries per walk group the reverse pass stores (0: the default)."""
        self._chk(lib().bhx_set_adjoint_list_capacity(self.h, ctypes.c_uint32(int(entries))),
                  "bhx_set_adjoint_list_capacity")

    def record_keeps_frames(self):
        """True when the last recording kept every frame (its reverse pass re-runs nothing)."""
        k = ctypes.c_int()
        self._chk(lib().bhx_record_keeps_frames(self.h, ctypes.byref(k)), "bhx_record_keeps_frames")
        return bool(k.value)

    def recorded_bodies(self):
        """(pos, vel, mass, acc) after the recorded frames, in the order the recording started with."""
        n = self.n
        pos, vel, acc, mass = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
        self._chk(lib().bhx_recorded_bodies(self.h, _p(pos), _p(vel), _p(mass), _p(acc)), "bhx_recorded_bodies")
        return pos, vel, mass, acc

    def record_digest(self):
        """Decision digest of the last recording (FP64 walk): equal digests, equal trees and lists."""
        d = ctypes.c_uint64()
        self._chk(lib().bhx_record_digest(self.h, ctypes.byref(d)), "bhx_record_digest")
        return d.value

    def adjoint(self, gx=None, gv=None, ga=None):
        """dL/d(x0, v0, a0, m) from dL/d(x_T, v_T, a_T), all in the starting body order."""
        n = self.n
        g = [None if v is None else _f64(v, (n, 3)) for v in (gx, gv, ga)]
        gx0, gv0, ga0, gm = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
        self._chk(lib().bhx_adjoint(self.h, *[_p(v) for v in g], _p(gx0), _p(gv0), _p(gm), _p(ga0)), "bhx_adjoint")
        return gx0, gv0, ga0, gm

    def grad_timing(self):
        r, a = ctypes.c_double(), ctypes.c_double()
        self._chk(lib().bhx_grad_timing(self.h, ctypes.byref(r), ctypes.byref(a)), "bhx_grad_timing")
        return r.value, a.value

    def sync(self):
        self._chk(lib().bhx_sync(self.h), "bhx_sync")

    def set_profiling(self, on):
        self._chk(lib().bhx_set_profiling(self.h, int(on)), "bhx_set_profiling")

    def last_timing(self):
        t = Timing()
        self._chk(lib().bhx_last_timing(self.h, ctypes.byref(t)), "bhx_last_timing")
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    def get_compensation(self):
        pc, vc = np.empty((self.n, 3)), np.empty((self.n, 3))
        self._chk(lib().bhx_get_compensation(self.h, _p(pc), _p(vc)), "bhx_get_compensation")
        return pc, vc

    def set_compensation(self, pos_c=None, vel_c=None):
        pc = None if pos_c is None else _f64(pos_c, (self.n, 3))
        vc = None if vel_c is None else _f64(vel_c, (self.n, 3))
        self._chk(lib().bhx_set_compensation(self.h, _p(pc), _p(vc)), "bhx_set_compensation")

    def get_systems(self):
        off = np.empty(self.nsys + 1, np.uint64)
        self._chk(lib().bhx_get_systems(self.h, _p(off)), "bhx_get_systems")
        return off

    # --- snapshots and visualisation export
    def save(self, path, time=0.0, step=0):
        self._chk(lib().bhx_save(self.h, _path(path), time, step), "bhx_save")

    def load(self, path):
        """Restores bodies, accelerations and systems; returns (time, step)."""
        t = ctypes.c_double()
        st = ctypes.c_uint64()
        self._chk(lib().bhx_load(self.h, _path(path), ctypes.byref(t), ctypes.byref(st)), "bhx_load")
        return t.value, st.value

    def export_vtk(self, bodies_path=None, nodes_path=None, min_level=0, max_level=21):
        self._chk(lib().bhx_export_vtk(self.h, None if bodies_path is None else _path(bodies_path),
                                       None if nodes_path is None else _path(nodes_path), min_level, max_level),
                  "bhx_export_vtk")

    # --- diagnostics
    def direct(self, targets, eps):
        targets = np.ascontiguousarray(targets, dtype=np.int64)
        acc = np.empty((targets.size, 3))
        pot = np.empty(targets.size)
        self._chk(lib().bhx_direct(self.h, _p(targets), targets.size, eps, _p(acc), _p(pot)), "bhx_direct")
        return acc, pot

    def energy(self, potentials=False):
        """(K, W, 2K/|W|) with W from the tree walk; with potentials=True also phi per body."""
        k, w, v = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        phi = np.empty(self.n) if potentials else None
        self._chk(lib().bhx_energy(self.h, ctypes.byref(k), ctypes.byref(w), ctypes.byref(v), _p(phi)),
                  "bhx_energy")
        return (k.value, w.value, v.value, phi) if potentials else (k.value, w.value, v.value)

    # --- the reference's per-body tree API (one system, last build)
    def interaction_lists(self, points, theta):
        """TraverseHOTInteractionList for each point, in the reference's order:
        (offsets (nq+1,), entries) -- entry >= 0 a node index, -(1+i) body i."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        nq = pts.shape[0]
        off = np.zeros(nq + 1, np.uint64)
        tot = ctypes.c_size_t()
        self._chk(lib().bhx_interaction_lists(self.h, _p(pts), nq, theta, _p(off), None, 0, ctypes.byref(tot)),
                  "bhx_interaction_lists")
        ent = np.zeros(max(1, tot.value), np.int64)
        self._chk(lib().bhx_interaction_lists(self.h, _p(pts), nq, theta, _p(off), _p(ent), ent.size,
                                              ctypes.byref(tot)), "bhx_interaction_lists")
        return off, ent[:tot.value]

    def force_lists(self, points, offsets, entries, acc=None):
        """ComputeHOTForceInteractionList for each point: acc (default 0) plus its list's terms."""
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        nq = pts.shape[0]
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        if off.shape != (nq + 1,):
            raise ValueError("offsets: expected %d values" % (nq + 1))
        ent = np.ascontiguousarray(entries, dtype=np.int64)
        if ent.size < int(off[-1]):
            raise ValueError("entries: fewer values than offsets[-1]")
        a = np.zeros((nq, 3)) if acc is None else _f64_exact(acc, (nq, 3), "acc").copy()
        self._chk(lib().bhx_force_lists(self.h, _p(pts), nq, _p(off), _p(ent), _p(a)), "bhx_force_lists")
        return a

    def mac(self, node_ids, points, theta):
        ids = np.ascontiguousarray(node_ids, dtype=np.int64)
        pts = _f64_exact(points, (ids.size, 3), "points")
        out = np.zeros(ids.size, np.uint8)
        self._chk(lib().bhx_mac(self.h, _p(ids), _p(pts), ids.size, theta, _p(out)), "bhx_mac")
        return out.astype(bool)

    def node_put(self, record):
        """insertHOTNode: record is a NODE_DTYPE scalar/array of one; returns its node index."""
        rec = np.ascontiguousarray(np.asarray(record, dtype=NODE_DTYPE).reshape(1))
        nid = ctypes.c_int64()
        self._chk(lib().bhx_node_put(self.h, _p(rec), ctypes.byref(nid)), "bhx_node_put")
        return nid.value

    def node_erase(self, key):
        done = ctypes.c_int()
        self._chk(lib().bhx_node_erase(self.h, int(key), ctypes.byref(done)), "bhx_node_erase")
        return bool(done.value)

    def prune_empty(self, key=1):
        n = ctypes.c_size_t()
        self._chk(lib().bhx_prune_empty(self.h, int(key), ctypes.byref(n)), "bhx_prune_empty")
        return n.value

    def decode_morton_keys(self, keys, size):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.zeros((keys.size, 3))
        self._chk(lib().bhx_decode_morton_keys(self.h, _p(keys), keys.size, size, _p(out)),
                  "bhx_decode_morton_keys")
        return out

    def bounds_path(self):
        """(element-by-element tiles of x, y, z; 1 if the plain chain ran) of the last
        one-system bounds."""
        out = np.zeros(4, np.uint32)
        self._chk(lib().bhx_bounds_path(self.h, _p(out)), "bhx_bounds_path")
        return [int(v) for v in out]

    def diagnostics(self):
        ek = np.zeros(1)
        mom = np.zeros(3)
        am = np.zeros(3)
        self._chk(lib().bhx_diagnostics(self.h, _p(ek), _p(mom), _p(am)), "bhx_diagnostics")
        return float(ek[0]), mom, am
