"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The checker for the HIP engine.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.  Each function below calls
the C restatement in nbody_oracle.c (which cites the reference kernel it
follows).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

RULES_REFERENCE, RULES_STANDARD = 0, 1
LAW_PLUMMER, LAW_REFERENCE_UNIT = 0, 1
MAX_LEVEL = 64
DQ_LEN = MAX_LEVEL + 1
PATH_STRIDE = 72
OK, E_TREE_OVERFLOW, E_DEPTH_OVERFLOW, E_NONFINITE = 0, 5, 6, 7

_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_U8P = C.POINTER(C.c_uint8)
_V = C.c_void_p

_SIG = {
    "orc_tree_new": (_V, [C.c_int, C.c_int]),
    "orc_tree_free": (None, [_V]),
    "orc_tree_ncells": (C.c_int, [_V]),
    "orc_tree_max_depth": (C.c_int, [_V]),
    "orc_bbox": (C.c_int, [C.c_int, _FP, _FP]),
    "orc_root_box": (None, [C.c_int, _FP, _FP, _FP]),
    "orc_path_of": (C.c_int, [C.c_int, _FP, _FP, C.c_int, _U8P]),
    "orc_build": (C.c_int, [_V, C.c_int, _FP, _FP]),
    "orc_summarize": (None, [_V, _FP, _FP]),
    "orc_sort": (None, [_V, _IP]),
    "orc_force_order": (None, [_V, _IP]),
    "orc_export": (C.c_int, [_V, _IP, _U8P, _IP, _IP, _IP, _IP, _FP]),
    "orc_dq_table": (None, [C.c_int, _FP, C.c_float, C.c_float, _FP]),
    "orc_force_bh": (None, [_V, _FP, _FP, _IP, _FP, C.c_float, C.c_float, C.c_int, C.c_int, _IP, C.c_int, _FP,
                            _DP, _I64P]),
    "orc_force_bf": (None, [C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, _FP, _FP, _IP, C.c_int, _FP, _DP]),
    "orc_direct_f64": (None, [C.c_int, C.c_double, C.c_double, C.c_int, _FP, _FP, _IP, C.c_int, _DP]),
    "orc_euler": (None, [C.c_int, C.c_float, _FP, _FP, _FP]),
    "orc_kick": (None, [C.c_int, C.c_float, _FP, _FP]),
    "orc_drift": (None, [C.c_int, C.c_float, _FP, _FP]),
    "orc_collision": (None, [C.c_int, _FP, _FP]),
    "orc_energy_momentum_f64": (None, [C.c_int, C.c_double, C.c_double, C.c_int, _FP, _FP, _FP, _DP]),
    "orc_bh_step": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                              C.c_int, C.c_int, _FP, _FP, _FP, _FP]),
    "orc_threads": (C.c_int, []),
    "orc_set_threads": (None, [C.c_int]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = C.CDLL(LIB_PATH)
        for k, (r, a) in _SIG.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _lib = lib
    return _lib


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def fp(a):
    return a.ctypes.data_as(_FP) if a is not None else None


def ip(a):
    return a.ctypes.data_as(_IP) if a is not None else None


def bbox(pos):
    pos = f32(pos).reshape(-1)
    box = np.empty(6, np.float32)
    st = load().orc_bbox(pos.size // 3, fp(pos), fp(box))
    return box, st


def root_box(rules, box):
    lo = np.empty(3, np.float32)
    hi = np.empty(3, np.float32)
    load().orc_root_box(rules, fp(f32(box)), fp(lo), fp(hi))
    return lo, hi


def path_of(rules, p3, box, maxlen=MAX_LEVEL + 1):
    d = np.empty(maxlen, np.uint8)
    load().orc_path_of(rules, fp(f32(p3)), fp(f32(box)), maxlen, d.ctypes.data_as(_U8P))
    return d


class Tree:
    """Sequentially built tree of the oracle (orc_build + orc_summarize + orc_sort)."""

    def __init__(self, pos, mass, rules, cap=None, box=None):
        self.lib = load()
        self.pos = f32(pos).reshape(-1)
        self.mass = f32(mass).reshape(-1)
        self.n = self.mass.size
        self.cap = cap if cap is not None else max(2 * self.n, 16)
        self.rules = rules
        self.box = bbox(self.pos)[0] if box is None else f32(box)
        self.h = self.lib.orc_tree_new(self.n, self.cap)
        self.status = self.lib.orc_build(self.h, rules, fp(self.pos), fp(self.box))
        self.sorted = None
        self.force_order = None
        if self.status == OK:
            self.lib.orc_summarize(self.h, fp(self.pos), fp(self.mass))
            self.sorted = np.empty(self.n, np.int32)
            self.lib.orc_sort(self.h, ip(self.sorted))
            self.force_order = np.empty(self.n, np.int32)
            self.lib.orc_force_order(self.h, ip(self.force_order))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_tree_free(self.h)
            self.h = None

    @property
    def ncells(self):
        return self.lib.orc_tree_ncells(self.h)

    @property
    def max_depth(self):
        return self.lib.orc_tree_max_depth(self.h)

    def export(self):
        n, nc = self.n, self.ncells
        s = np.empty(n, np.int32)
        paths = np.empty(n * PATH_STRIDE, np.uint8)
        ln = np.empty(n, np.int32)
        lv = np.empty(nc, np.int32)
        sl = np.empty(nc, np.int32)
        ct = np.empty(nc, np.int32)
        com = np.empty(4 * nc, np.float32)
        self.lib.orc_export(self.h, ip(s), paths.ctypes.data_as(_U8P), ip(ln), ip(lv), ip(sl), ip(ct), fp(com))
        return {"sorted": s, "paths": paths.reshape(n, PATH_STRIDE), "path_len": ln, "cell_level": lv,
                "cell_slot": sl, "cell_count": ct, "cell_com": com.reshape(-1, 4)}

    def dq(self, theta, eps):
        return dq_table(self.rules, self.box, theta, eps)

    def force(self, theta, eps, G, vote=64, kahan=False, groups=None, order=None, want_absacc=True):
        """Barnes-Hut accelerations (n,3); also sum |term| per body (None unless want_absacc)
        and interactions per body.  Vote groups are `vote` consecutive bodies of `order`
        (default: the force order, as the kernel forms them); `groups` selects group indices
        of that order."""
        dq = self.dq(theta, eps)
        order = self.force_order if order is None else np.ascontiguousarray(order, dtype=np.int32)
        acc = np.zeros(3 * self.n, np.float32)
        absacc = np.zeros(self.n, np.float64) if want_absacc else None
        ninter = np.zeros(self.n, np.int64)
        g = None if groups is None else np.ascontiguousarray(groups, dtype=np.int32)
        self.lib.orc_force_bh(self.h, fp(self.pos), fp(self.mass), ip(order), fp(dq), eps, G, vote,
                              int(kahan), ip(g), 0 if g is None else g.size, fp(acc),
                              None if absacc is None else absacc.ctypes.data_as(_DP),
                              ninter.ctypes.data_as(_I64P))
        return acc.reshape(-1, 3), absacc, ninter


def dq_table(rules, box, theta, eps):
    dq = np.empty(DQ_LEN, np.float32)
    load().orc_dq_table(rules, fp(f32(box)), theta, eps, fp(dq))
    return dq


def force_bf(pos, mass, law, eps, G, kahan=False, targets=None):
    pos = f32(pos).reshape(-1)
    mass = f32(mass).reshape(-1)
    n = mass.size
    acc = np.zeros(3 * n, np.float32)
    absacc = np.zeros(n, np.float64)
    t = None if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
    load().orc_force_bf(law, eps, G, int(kahan), n, fp(pos), fp(mass), ip(t), 0 if t is None else t.size,
                        fp(acc), absacc.ctypes.data_as(_DP))
    return acc.reshape(-1, 3), absacc


def direct_f64(pos, mass, law, eps, G, targets=None):
    pos = f32(pos).reshape(-1)
    mass = f32(mass).reshape(-1)
    n = mass.size
    t = None if targets is None else np.ascontiguousarray(targets, dtype=np.int32)
    nt = n if t is None else t.size
    acc = np.zeros(3 * nt, np.float64)
    load().orc_direct_f64(law, eps, G, n, fp(pos), fp(mass), ip(t), 0 if t is None else nt,
                          acc.ctypes.data_as(_DP))
    return acc.reshape(-1, 3)


def euler(acc, vel, pos, dt):
    acc, vel, pos = f32(acc).reshape(-1), f32(vel).reshape(-1).copy(), f32(pos).reshape(-1).copy()
    load().orc_euler(pos.size // 3, dt, fp(acc), fp(vel), fp(pos))
    return vel.reshape(-1, 3), pos.reshape(-1, 3)


def kick(acc, vel, h):
    acc, vel = f32(acc).reshape(-1), f32(vel).reshape(-1).copy()
    load().orc_kick(vel.size // 3, h, fp(acc), fp(vel))
    return vel.reshape(-1, 3)


def drift(vel, pos, dt):
    vel, pos = f32(vel).reshape(-1), f32(pos).reshape(-1).copy()
    load().orc_drift(pos.size // 3, dt, fp(vel), fp(pos))
    return pos.reshape(-1, 3)


def collision(pos, vel):
    """Velocities after the collision stage (reflection of pairs closer than 0.03)."""
    pos = f32(pos).reshape(-1)
    vel = f32(vel).reshape(-1).copy()
    load().orc_collision(pos.size // 3, fp(pos), fp(vel))
    return vel.reshape(-1, 3)


def energy_momentum(pos, vel, mass, eps, G, law=LAW_PLUMMER):
    """FP64 kinetic and potential energy, linear and angular momentum, total mass."""
    out = np.zeros(9, np.float64)
    pos, vel, mass = f32(pos).reshape(-1), f32(vel).reshape(-1), f32(mass).reshape(-1)
    load().orc_energy_momentum_f64(law, eps, G, mass.size, fp(pos), fp(vel), fp(mass), out.ctypes.data_as(_DP))
    return {"kinetic": out[0], "potential": out[1], "momentum": out[2:5], "angular": out[5:8], "mass": out[8]}


def bh_steps(pos, vel, mass, rules, nsteps, theta, eps, G, dt, vote=64, kahan=False, cap=None):
    pos = f32(pos).reshape(-1).copy()
    vel = f32(vel).reshape(-1).copy()
    mass = f32(mass).reshape(-1)
    n = mass.size
    acc = np.zeros(3 * n, np.float32)
    st = load().orc_bh_step(rules, nsteps, n, cap or max(2 * n, 16), theta, eps, G, dt, vote, int(kahan), fp(pos),
                            fp(vel), fp(mass), fp(acc))
    return st, pos.reshape(-1, 3), vel.reshape(-1, 3), acc.reshape(-1, 3)


def threads():
    return load().orc_threads()


def set_threads(n):
    load().orc_set_threads(int(n))
