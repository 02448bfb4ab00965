"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product path (libbhx) never
imports it.  The C sources restate the reference's serial step; see the
citations in bhx_oracle.c.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

AS_RUN = 0
INTENDED = 1

NODE_DTYPE = np.dtype([
    ("key", "<u8"), ("center", "<f8", (3,)), ("size", "<f8"), ("com", "<f8", (3,)),
    ("mass", "<f8"), ("quad", "<f8", (6,)), ("n", "<i8"), ("first", "<i8"),
    ("level", "<i4"), ("child_byte", "<u4"), ("member_begin", "<i8"), ("member_count", "<i8"),
])
assert NODE_DTYPE.itemsize == 160


class FrameStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("t_bounds", "t_keys", "t_sort", "t_kick_drift", "t_build", "t_walk", "t_kick")] + \
               [("nodes", ctypes.c_int64), ("interactions", ctypes.c_int64)]


def build():
    """Compile liboracle.so from oracle/bhx_oracle.c (make)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        S = ctypes.c_size_t
        D = ctypes.c_double
        I = ctypes.c_int
        U64 = ctypes.c_uint64
        L.orc_bounds.argtypes = [P, S, P, P]
        L.orc_x86_cvt_u64.argtypes = [D]; L.orc_x86_cvt_u64.restype = U64
        L.orc_spread21.argtypes = [U64]; L.orc_spread21.restype = U64
        L.orc_morton_key.argtypes = [P, D]; L.orc_morton_key.restype = U64
        L.orc_morton_keys.argtypes = [P, S, D, P]
        L.orc_radix_sort.argtypes = [P, P, S]
        L.orc_gather.argtypes = [P, S, P, S]
        L.orc_kick_drift.argtypes = [D, P, P, P, S]
        L.orc_kick.argtypes = [D, P, P, S]
        L.orc_tree_path_key.argtypes = [P, P, D]; L.orc_tree_path_key.restype = U64
        L.orc_tree_path_keys.argtypes = [P, S, P, D, P]
        L.orc_mac_raw.argtypes = [D, P, P, D]; L.orc_mac_raw.restype = I
        L.orc_decode_morton_key.argtypes = [U64, D, P]
        L.orc_tree_list.argtypes = [P, P, P, D, P, P]; L.orc_tree_list.restype = ctypes.c_int64
        L.orc_tree_build_insert.argtypes = [P, P, S, P, D]; L.orc_tree_build_insert.restype = P
        L.orc_tree_build_recursive.argtypes = [P, P, S, P, D, I]; L.orc_tree_build_recursive.restype = P
        L.orc_tree_moments_intended.argtypes = [P, P, P]
        L.orc_tree_size.argtypes = [P]; L.orc_tree_size.restype = S
        L.orc_tree_unsplit.argtypes = [P]; L.orc_tree_unsplit.restype = S
        L.orc_tree_nodes.argtypes = [P, P]
        L.orc_tree_lookup.argtypes = [P, U64]; L.orc_tree_lookup.restype = ctypes.c_int64
        L.orc_tree_free.argtypes = [P]
        L.orc_walk.argtypes = [P, P, P, S, D, D, I, P, P]
        L.orc_walk_subset.argtypes = [P, P, P, P, S, D, D, I, P, P]
        L.orc_direct.argtypes = [P, P, S, P, S, D, P, P]
        L.orc_frame.argtypes = [P, P, P, P, P, S, D, D, D, I, I, I, I, ctypes.POINTER(FrameStats)]
        L.orc_traj_direct.argtypes = [P, P, P, P, S, D, D, I, I]
        L.orc_traj_direct_adjoint.argtypes = [P, P, P, P, S, D, D, I, I, P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def bounds(pos):
    pos = _c(pos, np.float64)
    c = np.zeros(3)
    s = np.zeros(1)
    lib().orc_bounds(_p(pos), pos.shape[0], _p(c), _p(s))
    return c, float(s[0])


def morton_keys(pos, size):
    pos = _c(pos, np.float64)
    keys = np.empty(pos.shape[0], np.uint64)
    lib().orc_morton_keys(_p(pos), pos.shape[0], size, _p(keys))
    return keys


def morton_key(p, size):
    p = _c(p, np.float64)
    return int(lib().orc_morton_key(_p(p), size))


def spread21(t):
    return int(lib().orc_spread21(int(t) & 0xFFFFFFFFFFFFFFFF))


def x86_cvt_u64(u):
    return int(lib().orc_x86_cvt_u64(float(u)))


def radix_sort(keys, idx=None):
    keys = np.array(keys, dtype=np.uint64, copy=True)
    idx = np.arange(keys.size, dtype=np.uint64) if idx is None else np.array(idx, np.uint64, copy=True)
    lib().orc_radix_sort(_p(keys), _p(idx), keys.size)
    return keys, idx


def tree_path_keys(pos, c, s):
    pos = _c(pos, np.float64)
    c = _c(c, np.float64)
    out = np.empty(pos.shape[0], np.uint64)
    lib().orc_tree_path_keys(_p(pos), pos.shape[0], _p(c), s, _p(out))
    return out


def decode_morton_key(key, size):
    """decodeMortonKey (MortonKeys.hpp:204-228)."""
    out = np.zeros(3)
    lib().orc_decode_morton_key(int(key) & 0xFFFFFFFFFFFFFFFF, size, _p(out))
    return out


def mac(size, com, r, theta):
    """LinearHashedOctree::MAC (HashedOctree.cpp:872-885) for one node and point."""
    return bool(lib().orc_mac_raw(size, _p(_c(com, np.float64)), _p(_c(r, np.float64)), theta))


def build_mutant(define, out_path):
    """Compile a copy of the oracle with one planted mistake (ORC_MUTANT_*) and load it:
    used by the tests to show that a known-answer test tells the two forms apart."""
    subprocess.run(["make", "-s", "-C", HERE, "mutant", "MUTANT=" + define, "OUT=" + str(out_path)], check=True)
    L = ctypes.CDLL(str(out_path))
    P, S, D = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double
    L.orc_bounds.argtypes = [P, S, P, P]
    L.orc_kick_drift.argtypes = [D, P, P, P, S]
    L.orc_kick.argtypes = [D, P, P, S]
    L.orc_mac_raw.argtypes = [D, P, P, D]
    L.orc_mac_raw.restype = ctypes.c_int
    return L


class Tree:
    """Owning wrapper of an orc_tree."""

    def __init__(self, handle, pos, mass):
        self.h = handle
        self._pos = pos
        self._mass = mass

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def __len__(self):
        return int(lib().orc_tree_size(self.h))

    @property
    def unsplit(self):
        return int(lib().orc_tree_unsplit(self.h))

    def nodes(self):
        out = np.zeros(len(self), NODE_DTYPE)
        lib().orc_tree_nodes(self.h, _p(out))
        return out

    def lookup(self, key):
        return int(lib().orc_tree_lookup(self.h, int(key)))

    def moments_intended(self):
        lib().orc_tree_moments_intended(self.h, _p(self._pos), _p(self._mass))

    def interaction_list(self, point, theta):
        """TraverseHOTInteractionList for one point, in the reference's order:
        (ids, is_body): a node index, or a body index where is_body is 1."""
        p = _c(point, np.float64)
        cap = len(self) + self._pos.shape[0] + 8
        ids = np.zeros(cap, np.int64)
        kind = np.zeros(cap, np.int32)
        nl = lib().orc_tree_list(self.h, _p(self._pos), _p(p), theta, _p(ids), _p(kind))
        return ids[:nl], kind[:nl]

    def walk(self, theta, eps, quad=False, targets=None):
        n = self._pos.shape[0]
        if targets is None:
            acc = np.zeros((n, 3))
            cnt = np.zeros(n, np.int64)
            lib().orc_walk(self.h, _p(self._pos), _p(self._mass), n, theta, eps, int(quad), _p(acc), _p(cnt))
        else:
            targets = _c(targets, np.int64)
            acc = np.zeros((targets.size, 3))
            cnt = np.zeros(targets.size, np.int64)
            lib().orc_walk_subset(self.h, _p(self._pos), _p(self._mass), _p(targets), targets.size,
                                  theta, eps, int(quad), _p(acc), _p(cnt))
        return acc, cnt


def build_insert(pos, mass, c, s):
    pos = _c(pos, np.float64)
    mass = _c(mass, np.float64)
    c = _c(c, np.float64)
    return Tree(lib().orc_tree_build_insert(_p(pos), _p(mass), pos.shape[0], _p(c), s), pos, mass)


def build_recursive(pos, mass, c, s, leaf_cap=1):
    pos = _c(pos, np.float64)
    mass = _c(mass, np.float64)
    c = _c(c, np.float64)
    return Tree(lib().orc_tree_build_recursive(_p(pos), _p(mass), pos.shape[0], _p(c), s, leaf_cap), pos, mass)


def direct(pos, mass, eps, targets=None):
    pos = _c(pos, np.float64)
    mass = _c(mass, np.float64)
    n = pos.shape[0]
    if targets is None:
        tg = None
        nt = n
    else:
        tg = _c(targets, np.int64)
        nt = tg.size
    acc = np.zeros((nt, 3))
    pot = np.zeros(nt)
    lib().orc_direct(_p(pos), _p(mass), n, None if tg is None else _p(tg), nt, eps, _p(acc), _p(pot))
    return acc, pot


def kick_drift(dt, pos, vel, acc):
    pos = np.array(pos, np.float64, copy=True)
    vel = np.array(vel, np.float64, copy=True)
    acc = _c(acc, np.float64)
    lib().orc_kick_drift(dt, _p(pos), _p(vel), _p(acc), pos.shape[0])
    return pos, vel


def kick(dt, vel, acc):
    vel = np.array(vel, np.float64, copy=True)
    acc = _c(acc, np.float64)
    lib().orc_kick(dt, _p(vel), _p(acc), vel.shape[0])
    return vel


def frame(pos, vel, mass, acc, dt, theta, eps, leaf_cap=1, mode=AS_RUN, quad=False, permute_acc=True):
    """One reference frame in place on copies; returns (pos, vel, mass, acc, keys, stats)."""
    pos = np.array(pos, np.float64, copy=True)
    vel = np.array(vel, np.float64, copy=True)
    mass = np.array(mass, np.float64, copy=True)
    acc = np.array(acc, np.float64, copy=True)
    keys = np.zeros(pos.shape[0], np.uint64)
    st = FrameStats()
    lib().orc_frame(_p(pos), _p(vel), _p(mass), _p(acc), _p(keys), pos.shape[0], dt, theta, eps,
                    leaf_cap, mode, int(quad), int(permute_acc), ctypes.byref(st))
    return pos, vel, mass, acc, keys, st


def traj_direct(pos, vel, mass, eps, dt, nsteps, init_acc=True, acc=None):
    """Leapfrog with FP64 direct-sum forces; returns (pos, vel, acc) after nsteps frames."""
    pos = np.array(pos, np.float64, copy=True)
    vel = np.array(vel, np.float64, copy=True)
    mass = _c(mass, np.float64)
    acc = np.zeros_like(pos) if acc is None else np.array(acc, np.float64, copy=True)
    lib().orc_traj_direct(_p(pos), _p(vel), _p(mass), _p(acc), pos.shape[0], eps, dt, nsteps, int(init_acc))
    return pos, vel, acc


def traj_direct_adjoint(pos, vel, mass, eps, dt, nsteps, gx=None, gv=None, ga=None, init_acc=True, acc=None):
    """dL/d(x0, v0, a0, m) of traj_direct for dL/d(x_T, v_T, a_T) = (gx, gv, ga)."""
    pos = _c(pos, np.float64)
    vel = _c(vel, np.float64)
    mass = _c(mass, np.float64)
    n = pos.shape[0]
    a0 = None if acc is None else _c(acc, np.float64)
    g = [None if v is None else _c(v, np.float64) for v in (gx, gv, ga)]
    gx0, gv0, ga0, gm = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    lib().orc_traj_direct_adjoint(_p(pos), _p(vel), _p(mass), None if a0 is None else _p(a0), n, eps, dt, nsteps,
                                  int(init_acc), *[None if v is None else _p(v) for v in g],
                                  _p(gx0), _p(gv0), _p(ga0), _p(gm))
    return gx0, gv0, ga0, gm
