"""Project-written autotuner: picks the brute-force / Barnes-Hut crossover N on
the device it runs on.

For each N of a geometric grid it builds one batched ensemble of equal systems
(Plummer ICs, theta and eps as given) totalling about `bodies` bodies, times
one batched step with brute force and one with Barnes-Hut, and reports the cost
per system-step.  The crossover is the smallest N from which Barnes-Hut is
cheaper.  Results are cached per device architecture in a JSON file.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import _lib as L
from . import ic
from .engine import NBodyEnsemble, device_info, system_spec

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "nbodysimulation_amd", "autotune.json")


def _time_step(specs, law, reps=3):
    with NBodyEnsemble(specs, law=law) as ens:
        for i, sp in enumerate(specs):
            p, v, m = ic.plummer(sp.num_bodies, 1000 + i)
            ens.upload(i, p, v, m)
        ens.step(1)  # warm-up (graph capture, first touch)
        t0 = time.perf_counter()
        ens.step(reps)
        return (time.perf_counter() - t0) / reps


def crossover(theta: float = 0.5, eps: float = 0.01, bodies: int = 1 << 20, grid=None, log=None) -> dict:
    grid = grid or [256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536]
    rows = []
    for n in grid:
        k = max(1, bodies // n)
        bf = _time_step([system_spec(n, L.BRUTE_FORCE_GLOBAL, eps=eps) for _ in range(k)], L.LAW_PLUMMER)
        bh = _time_step([system_spec(n, L.BARNES_HUT, rules=L.RULES_STANDARD, theta=theta, eps=eps)
                         for _ in range(k)], L.LAW_PLUMMER)
        rows.append({"n": n, "systems": k, "bf_us_per_system_step": 1e6 * bf / k,
                     "bh_us_per_system_step": 1e6 * bh / k})
        if log:
            log(f"autotune: N={n:6d} x{k:5d}  BF {1e6 * bf / k:9.2f} us  BH {1e6 * bh / k:9.2f} us per system-step")
    cross = None
    for r in rows:
        if r["bh_us_per_system_step"] < r["bf_us_per_system_step"]:
            cross = r["n"]
            break
    if cross is None:
        cross = 2 * grid[-1]
    return {"crossover_n": cross, "theta": theta, "eps": eps, "rows": rows}


def load_or_tune(theta: float = 0.5, eps: float = 0.01, log=None, retune: bool = False) -> dict:
    info = device_info() or {}
    key = f"{info.get('arch', 'unknown')}|theta={theta}|eps={eps}"
    cache = {}
    if os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if not retune and key in cache:
        return cache[key]
    res = crossover(theta, eps, log=log)
    cache[key] = res
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    except OSError:
        pass
    return res


def choose_algorithm(n: int, crossover_n: int) -> int:
    return L.BARNES_HUT if n >= crossover_n else L.BRUTE_FORCE_GLOBAL


def log_uniform_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    r = np.random.default_rng(seed)
    return np.exp(r.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)
