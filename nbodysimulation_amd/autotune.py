"""Project-written autotuner for the device it runs on.

1. Brute-force launch shape: for each candidate (targets per lane, minimum waves per
   system) it times batched brute-force steps on three workloads (one system of 4096
   bodies, one of 131,072, and 64 systems of 4096) and keeps the candidate with the
   lowest geometric mean of the times relative to the best per workload.
2. Brute-force / Barnes-Hut crossover N: for each N of a geometric grid it builds one
   batched ensemble of equal systems (Plummer ICs, theta and eps as given) totalling
   about `bodies` bodies, times one batched step with brute force and one with
   Barnes-Hut under the chosen shape, and reports the cost per system-step.  The
   crossover is the smallest N from which Barnes-Hut is cheaper.

Results are cached per device architecture in a JSON file; `load_or_tune` applies the
chosen shape process-wide (nb_set_tuning) before returning.
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from . import _lib as L
from . import ic
from .engine import NBodyEnsemble, device_info, system_spec
from .workloads import choose_algorithm, log_uniform_sizes  # noqa: F401  (re-exported)

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "nbodysimulation_amd", "autotune.json")
SHAPES = [(tpl, waves) for waves in (4096, 8192, 16384) for tpl in (0, 1, 2)]


def _time_step(specs, law, reps=3):
    with NBodyEnsemble(specs, law=law) as ens:
        for i, sp in enumerate(specs):
            p, v, m = ic.plummer(sp.num_bodies, 1000 + i)
            ens.upload(i, p, v, m)
        ens.step(1)  # warm-up (graph capture, first touch)
        t0 = time.perf_counter()
        ens.step(reps)
        return (time.perf_counter() - t0) / reps


def tune_brute(eps: float = 0.01, log=None) -> dict:
    workloads = {"1x4096": [4096], "1x131072": [131072], "64x4096": [4096] * 64}
    rows = []
    for tpl, waves in SHAPES:
        L.set_tuning(bf_targets_per_lane=tpl, bf_min_waves=waves)
        t = {name: _time_step([system_spec(n, L.BRUTE_FORCE_GLOBAL, eps=eps) for n in sizes], L.LAW_PLUMMER,
                              reps=5 if len(sizes) == 1 and sizes[0] < 10000 else 3)
             for name, sizes in workloads.items()}
        rows.append({"bf_targets_per_lane": tpl, "bf_min_waves": waves, "seconds": t})
        if log:
            log("autotune: brute tpl=%d waves=%5d  " % (tpl, waves)
                + "  ".join(f"{k} {1e6 * v:9.1f} us" for k, v in t.items()))
    best = {k: min(r["seconds"][k] for r in rows) for k in workloads}
    for r in rows:
        r["score"] = math.exp(sum(math.log(r["seconds"][k] / best[k]) for k in workloads) / len(workloads))
    pick = min(rows, key=lambda r: r["score"])
    return {"bf_targets_per_lane": pick["bf_targets_per_lane"], "bf_min_waves": pick["bf_min_waves"],
            "rows": rows}


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
    """The cached result for this device (tuning it first if absent); applies the brute-force shape."""
    info = device_info() or {}
    key = f"{info.get('arch', 'unknown')}|theta={theta}|eps={eps}"
    cache = {}
    if os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    res = None if retune else cache.get(key)
    if res is None or "brute_shape" not in res:
        shape = tune_brute(eps, log=log)
        L.set_tuning(bf_targets_per_lane=shape["bf_targets_per_lane"], bf_min_waves=shape["bf_min_waves"])
        res = crossover(theta, eps, log=log)
        res["brute_shape"] = shape
        cache[key] = res
        try:
            os.makedirs(os.path.dirname(CACHE), exist_ok=True)
            with open(CACHE, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    shape = res["brute_shape"]
    L.set_tuning(bf_targets_per_lane=shape["bf_targets_per_lane"], bf_min_waves=shape["bf_min_waves"])
    return res


