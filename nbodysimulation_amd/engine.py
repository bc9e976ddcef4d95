"""numpy-facing wrapper over the C ABI: one NBodySim per nb_sim handle.

Mirrors the stage sequence of CMyApp::Update (src/MyApp.cpp:583-598):
bounds -> tree_reset -> tree_build -> summarize -> sort -> force -> update, and
the brute-force branch (src/MyApp.cpp:601-630).  All compute runs in
libnbody.so on the GPU; this module only moves numpy arrays across the ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _fp(a):
    return a.ctypes.data_as(L._FP) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(L._IP) if a is not None else None


def make_config(n: int, algorithm: int = L.BARNES_HUT, *, method: int = L.EULER, rules: int = L.RULES_REFERENCE,
                law: int = L.LAW_REFERENCE_UNIT, accumulation: int = L.ACCUM_PLAIN, G: float = 1.0,
                theta: float = 0.5, eps: float = 0.01, dt: float = 1e-3, cell_capacity: int = 0,
                use_graph: bool = True, time_stages: bool = False, collision: bool = False) -> L.NbConfig:
    lib = L.load()
    c = L.NbConfig()
    lib.nb_config_default(c)
    c.num_bodies, c.algorithm, c.method, c.rules = n, algorithm, method, rules
    c.law, c.accumulation, c.gravitational_constant = law, accumulation, G
    c.theta, c.epsilon, c.dt, c.cell_capacity = theta, eps, dt, cell_capacity
    c.use_graph, c.time_stages, c.collision = int(use_graph), int(time_stages), int(collision)
    return c


class NBodySim:
    def __init__(self, cfg: L.NbConfig, stream: int | None = None):
        self.lib = L.load()
        self.cfg = cfg
        self.n = cfg.num_bodies
        h = C.c_void_p()
        L.check(self.lib.nb_sim_create(C.byref(cfg), C.c_void_p(stream) if stream else None, C.byref(h)),
                "nb_sim_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.nb_sim_destroy(self.h)
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

    def _call(self, name, *args):
        L.check(getattr(self.lib, name)(self.h, *args), name)

    # state
    def upload(self, pos, vel, mass):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
        mass = np.ascontiguousarray(mass, dtype=np.float32).reshape(-1)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
        assert pos.size == 3 * self.n and mass.size == self.n
        self._call("nb_sim_upload", _fp(pos), _fp(vel), _fp(mass))

    def download(self):
        p = np.empty(3 * self.n, np.float32)
        v = np.empty(3 * self.n, np.float32)
        a = np.empty(3 * self.n, np.float32)
        if self.n:
            self._call("nb_sim_download", _fp(p), _fp(v), _fp(a))
        return p.reshape(-1, 3), v.reshape(-1, 3), a.reshape(-1, 3)

    def status(self) -> int:
        return self.lib.nb_sim_status(self.h)

    def clear_status(self):
        self._call("nb_sim_clear_status")

    def sync(self):
        self._call("nb_sim_sync")

    # stages
    def bounds(self):
        box = np.empty(6, np.float32)
        self._call("nb_stage_bounds", _fp(box))
        return box

    def tree_reset(self): self._call("nb_stage_tree_reset")
    def tree_build(self): self._call("nb_stage_tree_build")
    def summarize(self): self._call("nb_stage_summarize")
    def sort(self): self._call("nb_stage_sort")
    def force(self): self._call("nb_stage_force")
    def brute_force(self): self._call("nb_stage_brute_force")
    def collision(self): self._call("nb_stage_collision")

    def force_range(self, first: int, count: int):
        """Barnes-Hut force for the targets [first, first+count) of the force order only."""
        self._call("nb_stage_force_range", int(first), int(count))

    def brute_force_range(self, first: int, count: int):
        """Brute-force force for the bodies [first, first+count) only."""
        self._call("nb_stage_brute_force_range", int(first), int(count))
    def update(self): self._call("nb_stage_update")

    def force_pipeline(self):
        """bounds..force on the current positions; returns the box."""
        box = self.bounds()
        self.tree_reset()
        self.tree_build()
        self.summarize()
        self.sort()
        self.force()
        return box

    def step(self, nsteps: int = 1, grow: bool = False):
        """nsteps whole steps.  grow=True (Barnes-Hut, Euler): a step that overflows the cell
        pool is repeated after the pool has doubled (steps are then issued one at a time)."""
        if not grow:
            self._call("nb_step", int(nsteps))
            return
        if self.cfg.algorithm != L.BARNES_HUT or self.cfg.method != L.EULER:
            raise ValueError("elastic cell pools need a Barnes-Hut simulation with the Euler method")
        done = 0
        while done < nsteps:
            st = self.lib.nb_step(self.h, 1)
            if st == L.NB_E_TREE_OVERFLOW and self.cell_capacity < (1 << 24) - 1:
                self.clear_status()
                self.set_cell_capacity(min(2 * self.cell_capacity, (1 << 24) - 1))
                continue
            L.check(st, "nb_step")
            done += 1

    @property
    def cell_capacity(self) -> int:
        return self.tree_info()["capacity"]

    def set_cell_capacity(self, cells: int):
        """Re-create the arena with a cell pool of `cells`, keeping the bodies' state."""
        self._call("nb_sim_set_cell_capacity", int(cells))
        self.cfg.cell_capacity = int(cells)

    def step_async(self, nsteps: int = 1):
        self._call("nb_step_async", int(nsteps))

    # inspection
    def tree_info(self):
        info = L.NbTreeInfo()
        self._call("nb_tree_info_get", C.byref(info))
        return {"num_cells": info.num_cells, "max_depth": info.max_depth, "capacity": info.capacity,
                "status": info.status}

    def export_bodies(self, paths: bool = True):
        s = np.empty(self.n, np.int32)
        pth = np.empty(self.n * L.PATH_STRIDE, np.uint8) if paths else None
        ln = np.empty(self.n, np.int32) if paths else None
        self._call("nb_export_bodies", _ip(s), pth.ctypes.data_as(L._U8P) if paths else None, _ip(ln))
        return s, (pth.reshape(self.n, L.PATH_STRIDE) if paths else None), ln

    def export_cells(self):
        nc = self.tree_info()["num_cells"]
        lv = np.empty(nc, np.int32)
        sl = np.empty(nc, np.int32)
        ct = np.empty(nc, np.int32)
        com = np.empty(4 * nc, np.float32)
        self._call("nb_export_cells", _ip(lv), _ip(sl), _ip(ct), _fp(com))
        return lv, sl, ct, com.reshape(-1, 4)

    def export_force_order(self):
        """Body indices in force order (STANDARD rules: tree order with children along a Hilbert
        curve; REFERENCE rules: the tree order): the order whose 64-runs are the force pass's
        vote groups and in which force_range counts."""
        o = np.empty(self.n, np.int32)
        self._call("nb_export_force_order", _ip(o))
        return o

    def export_dq(self):
        dq = np.empty(L.DQ_LEN, np.float32)
        self._call("nb_export_dq", _fp(dq))
        return dq

    def interactions(self) -> int:
        t = C.c_int64()
        self._call("nb_force_interactions", C.byref(t))
        return t.value

    def interactions_per_body(self) -> np.ndarray:
        """Interactions of each body in the last force pass (body index order)."""
        out = np.zeros(self.n, np.int32)
        self._call("nb_export_interactions", out.ctypes.data_as(L._IP))
        return out

    def energy(self) -> dict:
        """FP64 kinetic/potential energy, momentum, angular momentum and mass of the current state (GPU)."""
        out = np.zeros(9, np.float64)
        self._call("nb_sim_energy", out.ctypes.data_as(L._DP))
        return {"kinetic": out[0], "potential": out[1], "total": out[0] + out[1], "momentum": out[2:5].copy(),
                "angular": out[5:8].copy(), "mass": out[8]}

    def export_vertices(self) -> np.ndarray:
        """(x, y, z, 1) per body, as the reference's copy_vertices kernel writes for rendering."""
        out = np.empty(4 * self.n, np.float32)
        self._call("nb_export_vertices", _fp(out), 0)
        return out.reshape(-1, 4)

    def step_count(self) -> int:
        k = C.c_int64()
        self._call("nb_sim_step_count", C.byref(k))
        return k.value

    def checkpoint(self) -> bytes:
        """The state (positions, velocities, masses, accelerations, step counter, config) as bytes."""
        size = C.c_size_t()
        self._call("nb_checkpoint_size", C.byref(size))
        buf = C.create_string_buffer(size.value)
        self._call("nb_checkpoint_save", buf, size)
        return buf.raw

    def restore(self, blob: bytes):
        buf = C.create_string_buffer(bytes(blob), len(blob))
        self._call("nb_checkpoint_load", buf, len(blob))

    def save(self, path: str):
        with open(path, "wb") as f:
            f.write(self.checkpoint())

    @classmethod
    def from_checkpoint(cls, path: str, stream: int | None = None) -> "NBodySim":
        """A new simulation with the configuration and state stored in a checkpoint file."""
        with open(path, "rb") as f:
            blob = f.read()
        h = L.NbCheckpointHeader.from_buffer_copy(blob[:C.sizeof(L.NbCheckpointHeader)])
        if h.magic != L.CHECKPOINT_MAGIC or h.version != L.CHECKPOINT_VERSION:
            raise ValueError(f"{path}: not a checkpoint of version {L.CHECKPOINT_VERSION}")
        sim = cls(h.config, stream=stream)
        sim.restore(blob)
        return sim

    def stage_times(self):
        ms = np.zeros(L.NUM_STAGES, np.float32)
        self._call("nb_stage_times", _fp(ms))
        return dict(zip(L.STAGE_NAMES, ms.tolist()))


def system_spec(n: int, algorithm: int = L.BARNES_HUT, *, rules: int = L.RULES_STANDARD, G: float = 1.0,
                theta: float = 0.5, eps: float = 0.01, dt: float = 1e-3, cell_capacity: int = 0,
                collision: bool = False) -> L.NbSystemSpec:
    sp = L.NbSystemSpec()
    sp.num_bodies, sp.algorithm, sp.rules, sp.cell_capacity = n, algorithm, rules, cell_capacity
    sp.collision = int(collision)
    sp.gravitational_constant, sp.theta, sp.epsilon, sp.dt = G, theta, eps, dt
    return sp


class NBodyEnsemble:
    """Many independent systems stepped by shared batched launches (one HIP graph per step)."""

    def __init__(self, specs, *, method: int = L.EULER, law: int = L.LAW_PLUMMER, accumulation: int = L.ACCUM_PLAIN,
                 use_graph: bool = True, stream: int | None = None):
        self.lib = L.load()
        self.specs = list(specs)
        arr = (L.NbSystemSpec * len(self.specs))(*self.specs)
        h = C.c_void_p()
        L.check(self.lib.nb_ens_create(arr, len(self.specs), method, law, accumulation, int(use_graph),
                                       C.c_void_p(stream) if stream else None, C.byref(h)), "nb_ens_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.nb_ens_destroy(self.h)
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

    def upload(self, i, pos, vel, mass):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
        mass = np.ascontiguousarray(mass, dtype=np.float32).reshape(-1)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
        L.check(self.lib.nb_ens_upload(self.h, i, _fp(pos), _fp(vel), _fp(mass)), "nb_ens_upload")

    def interactions_per_body(self, i) -> np.ndarray:
        """Interactions of each body of system i in the last force pass (body index order)."""
        out = np.zeros(self.specs[i].num_bodies, np.int32)
        L.check(self.lib.nb_ens_export_interactions(self.h, i, out.ctypes.data_as(L._IP)),
                "nb_ens_export_interactions")
        return out

    def download(self, i):
        n = self.specs[i].num_bodies
        p = np.empty(3 * n, np.float32)
        v = np.empty(3 * n, np.float32)
        a = np.empty(3 * n, np.float32)
        L.check(self.lib.nb_ens_download(self.h, i, _fp(p), _fp(v), _fp(a)), "nb_ens_download")
        return p.reshape(-1, 3), v.reshape(-1, 3), a.reshape(-1, 3)

    def step(self, nsteps: int = 1):
        L.check(self.lib.nb_ens_step(self.h, int(nsteps)), "nb_ens_step")

    def status(self):
        out = np.zeros(len(self.specs), np.int32)
        L.check(self.lib.nb_ens_status(self.h, _ip(out)), "nb_ens_status")
        return out

    def device_bytes(self) -> int:
        b = C.c_int64()
        L.check(self.lib.nb_ens_device_bytes(self.h, C.byref(b)), "nb_ens_device_bytes")
        return b.value


class NBodyServer:
    """Continuous-batching server: requests (system + step count) join and leave
    one shared batched step per tick inside a fixed HBM arena."""

    def __init__(self, hbm_budget_bytes: int, max_systems: int, *, method: int = L.EULER,
                 law: int = L.LAW_PLUMMER, accumulation: int = L.ACCUM_PLAIN, stream: int | None = None):
        self.lib = L.load()
        h = C.c_void_p()
        L.check(self.lib.nb_server_create(int(hbm_budget_bytes), int(max_systems), method, law, accumulation,
                                          C.c_void_p(stream) if stream else None, C.byref(h)), "nb_server_create")
        self.h = h
        self.sizes = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.nb_server_destroy(self.h)
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

    def submit(self, spec, steps: int, pos, vel, mass) -> int:
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
        mass = np.ascontiguousarray(mass, dtype=np.float32).reshape(-1)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
        rid = C.c_int64()
        L.check(self.lib.nb_server_submit(self.h, C.byref(spec), int(steps), _fp(pos), _fp(vel), _fp(mass),
                                          C.byref(rid)), "nb_server_submit")
        self.sizes[rid.value] = spec.num_bodies
        return rid.value

    def tick(self) -> int:
        done = C.c_int32()
        L.check(self.lib.nb_server_tick(self.h, C.byref(done)), "nb_server_tick")
        return done.value

    def poll(self):
        """(id, status) of the oldest retired request, or None."""
        rid, st = C.c_int64(), C.c_int32()
        if self.lib.nb_server_poll(self.h, C.byref(rid), C.byref(st)) != L.NB_OK:
            return None
        return rid.value, st.value

    def fetch(self, rid: int):
        """Final (pos, vel, acc) of a retired request; the server forgets it afterwards."""
        if rid not in self.sizes:  # not submitted here, or fetched already: the library reports it
            L.check(self.lib.nb_server_fetch(self.h, int(rid), None, None, None), "nb_server_fetch")
        n = self.sizes.get(rid, 0)
        p = np.empty(3 * n, np.float32)
        v = np.empty(3 * n, np.float32)
        a = np.empty(3 * n, np.float32)
        L.check(self.lib.nb_server_fetch(self.h, int(rid), _fp(p), _fp(v), _fp(a)), "nb_server_fetch")
        self.sizes.pop(rid, None)
        return p.reshape(-1, 3), v.reshape(-1, 3), a.reshape(-1, 3)

    def stats(self) -> dict:
        s = L.NbServerStats()
        L.check(self.lib.nb_server_stats_get(self.h, C.byref(s)), "nb_server_stats_get")
        return {k: getattr(s, k) for k, _ in L.NbServerStats._fields_ if k != "reserved"}


def device_info():
    lib = L.load()
    cu = C.c_int32()
    name = C.create_string_buffer(64)
    st = lib.nb_device_info(C.byref(cu), name, 64)
    if st != L.NB_OK:
        return None
    return {"compute_units": cu.value, "arch": name.value.decode()}
