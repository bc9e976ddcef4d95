"""Seeded initial conditions through libnbody.so's host generators (nbody_ic.h)."""
from __future__ import annotations

import numpy as np

from . import _lib as L


def _out(n):
    return np.empty(3 * n, np.float32), np.empty(3 * n, np.float32), np.empty(n, np.float32)


def generate(kind: int, n: int, seed: int, **kw):
    """kind: IC_NORMAL / IC_PLUMMER / IC_UNIFORM_BALL / IC_COLLISION; keyword
    overrides of nb_ic_params (total_mass, scale, mass_lo, mass_hi, cutoff, G,
    center, bulk_velocity)."""
    lib = L.load()
    p = L.NbIcParams()
    lib.nb_ic_params_default(p, kind)
    for k, v in kw.items():
        if k in ("center", "bulk_velocity"):
            arr = getattr(p, k)
            for i in range(3):
                arr[i] = float(v[i])
        else:
            setattr(p, k, v)
    pos, vel, mass = _out(n)
    L.check(lib.nb_ic_generate(p, n, seed, pos.ctypes.data_as(L._FP), vel.ctypes.data_as(L._FP),
                               mass.ctypes.data_as(L._FP)), "nb_ic_generate")
    return pos.reshape(-1, 3), vel.reshape(-1, 3), mass


def reference(n: int, seed: int, **kw):
    """The reference's InitBodyAttributes modes (src/MyApp.cpp:331-495), seeded."""
    lib = L.load()
    p = L.NbIcReferenceParams()
    lib.nb_ic_reference_default(p)
    for k, v in kw.items():
        setattr(p, k, v)
    pos, vel, mass = _out(n)
    L.check(lib.nb_ic_reference(p, n, seed, pos.ctypes.data_as(L._FP), vel.ctypes.data_as(L._FP),
                                mass.ctypes.data_as(L._FP)), "nb_ic_reference")
    return pos.reshape(-1, 3), vel.reshape(-1, 3), mass


def normal_cloud(n: int, seed: int):
    """x ~ Normal(0,1) per axis, m ~ U[0.5,1.5]/N, v = 0 (brute-force workload)."""
    return generate(L.IC_NORMAL, n, seed)


def plummer(n: int, seed: int, total_mass: float = 1.0, scale: float = 1.0, G: float = 1.0):
    """Plummer sphere, equal masses, virial-equilibrium velocities (AHW 1974), cutoff 0.999 of the mass."""
    return generate(L.IC_PLUMMER, n, seed, total_mass=total_mass, scale=scale, G=G)


def uniform_ball(n: int, seed: int, radius: float = 1.0, total_mass: float = 1.0):
    return generate(L.IC_UNIFORM_BALL, n, seed, scale=radius, total_mass=total_mass)


def collision(n: int, seed: int):
    """Two Plummer spheres of n/2 bodies each, a = 0.5, centres +-(2,0,0), bulk -+(0.3,0,0), M = 1 each."""
    return generate(L.IC_COLLISION, n, seed)
