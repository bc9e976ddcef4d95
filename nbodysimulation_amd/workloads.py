"""The batched-ensemble and serving workloads, defined once for the benchmark and the tests.

Ensemble: 512 independent systems, N log-uniform in [1024, 65536], each a Plummer sphere,
a Normal(0,1) cloud or a uniform unit ball (equal odds), m = 1/N, eps = 0.01, dt = 1e-3,
seed 2; brute force below the crossover N, Barnes-Hut (STANDARD rules, theta = 0.5) from it on.

Serving: a stream of requests, each a new Plummer system with N log-uniform in
[4096, 2,097,152], steps uniform in [10, 200], theta from {0.3, 0.5, 0.7}, eps = 0.01,
dt = 1e-3, Euler, open-loop Poisson arrivals, seed 3; brute force below the crossover N,
Barnes-Hut (STANDARD) from it on.  The reference steps one system per frame
(CMyApp::Update, src/MyApp.cpp:550-656) and has neither workload.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib as L
from . import ic
from .engine import system_spec

ENSEMBLE_SYSTEMS = 512
ENSEMBLE_N = (1024, 65536)
ENSEMBLE_SEED = 2
SERVING_N = (4096, 2097152)
SERVING_STEPS = (10, 200)
SERVING_THETAS = (0.3, 0.5, 0.7)
SERVING_SEED = 3


def log_uniform_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    r = np.random.default_rng(seed)
    return np.exp(r.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)


def choose_algorithm(n: int, crossover_n: int) -> int:
    return L.BARNES_HUT if n >= crossover_n else L.BRUTE_FORCE_GLOBAL


# ---------------------------------------------------------------- ensemble

def ensemble_sizes_kinds(count: int = ENSEMBLE_SYSTEMS, seed: int = ENSEMBLE_SEED):
    """Per system: N and the kind of cloud (0 Plummer, 1 Normal(0,1), 2 uniform ball)."""
    sizes = log_uniform_sizes(count, ENSEMBLE_N[0], ENSEMBLE_N[1], seed)
    kinds = np.random.default_rng(seed).integers(0, 3, count)
    return sizes, kinds


def ensemble_ic(i: int, n: int, kind: int, seed: int = ENSEMBLE_SEED):
    """System i's initial state; masses 1/N for every kind."""
    s = seed * 100003 + i
    if kind == 0:
        return ic.plummer(int(n), s)
    if kind == 1:
        p, v, _ = ic.normal_cloud(int(n), s)
        return p, v, np.full(int(n), 1.0 / n, np.float32)
    return ic.uniform_ball(int(n), s)


def ensemble_specs(sizes, crossover_n: int, theta: float = 0.5, eps: float = 0.01, dt: float = 1e-3):
    return [system_spec(int(n), choose_algorithm(int(n), crossover_n), rules=L.RULES_STANDARD, theta=theta,
                        eps=eps, dt=dt) for n in sizes]


# ---------------------------------------------------------------- serving

@dataclass
class Requests:
    sizes: np.ndarray     # bodies per request
    steps: np.ndarray     # steps per request
    thetas: np.ndarray    # opening angle per request
    arrivals: np.ndarray  # arrival time (s) of each request, the first at 0
    specs: list           # nb_system_spec per request


def serving_requests(count: int, rate: float, crossover_n: int, seed: int = SERVING_SEED) -> Requests:
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(count, SERVING_N[0], SERVING_N[1], seed)
    steps = rng.integers(SERVING_STEPS[0], SERVING_STEPS[1] + 1, count)
    thetas = rng.choice(SERVING_THETAS, count)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    arrivals -= arrivals[0]
    specs = [system_spec(int(n), choose_algorithm(int(n), crossover_n), rules=L.RULES_STANDARD, theta=float(t),
                         eps=0.01, dt=1e-3) for n, t in zip(sizes, thetas)]
    return Requests(sizes, steps, thetas, arrivals, specs)


def serving_ic(i: int, n: int, seed: int = SERVING_SEED):
    """Request i's initial state (a Plummer sphere, M = 1, a = 1)."""
    return ic.plummer(int(n), seed * 1000003 + i)
