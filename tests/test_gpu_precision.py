"""Brute-force kernel at large N (source split, register blocking), Kahan accumulation, and
the FP64 conservation diagnostics, against the oracle."""
import sys

import numpy as np
import pytest

from nbodysimulation_amd import _lib as L
from nbodysimulation_amd import ic
from nbodysimulation_amd.engine import NBodyEnsemble, NBodySim, make_config, system_spec

pytestmark = pytest.mark.gpu

ACC_RTOL = 2e-5  # per body, relative to sum_j |a_ij|


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("accum", [L.ACCUM_PLAIN, L.ACCUM_KAHAN])
def test_brute_force_large_n_parity(orc, gpu, accum):
    """N = 140,001: one wave per target group, a ragged last tile and last group."""
    n = 140001
    say(f"case bf large n={n} accum={accum}")
    pos, vel, mass = ic.collision(n - 1, 9)
    pos = np.vstack([pos, [[0.1, 0.2, 0.3]]]).astype(np.float32)
    mass = np.append(mass, np.float32(1e-5)).astype(np.float32)
    eps = 0.005
    with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, law=L.LAW_PLUMMER, accumulation=accum, eps=eps)) as s:
        s.upload(pos, None, mass)
        s.brute_force()
        _, _, a_g = s.download()
    tg = np.concatenate([np.random.default_rng(1).choice(n, 500, replace=False), [0, n - 1]]).astype(np.int32)
    a_o, absacc = orc.force_bf(pos, mass, L.LAW_PLUMMER, eps, 1.0, kahan=bool(accum), targets=tg)
    err = np.linalg.norm(a_g[tg] - a_o[tg], axis=1)
    assert (err <= ACC_RTOL * absacc[tg]).all()


def test_kahan_is_closer_to_fp64(orc, gpu):
    """Compensated accumulation: median error against the FP64 direct sum drops."""
    n = 131072
    pos, vel, mass = ic.collision(n, 3)
    tg = np.random.default_rng(2).choice(n, 256, replace=False).astype(np.int32)
    a64 = orc.direct_f64(pos, mass, L.LAW_PLUMMER, 0.005, 1.0, targets=tg)
    med = {}
    for accum in (L.ACCUM_PLAIN, L.ACCUM_KAHAN):
        with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, law=L.LAW_PLUMMER, accumulation=accum,
                                  eps=0.005)) as s:
            s.upload(pos, vel, mass)
            s.brute_force()
            _, _, a = s.download()
        med[accum] = np.median(np.linalg.norm(a[tg] - a64, axis=1) / np.linalg.norm(a64, axis=1))
    say(f"median rel err plain {med[0]:.3e} kahan {med[1]:.3e}")
    assert med[L.ACCUM_KAHAN] < 0.5 * med[L.ACCUM_PLAIN]


def test_brute_force_batch_independent(gpu):
    """Systems whose batch runs two targets per lane give the bytes they give alone (one per lane)."""
    sizes = [300000, 200000, 50000]
    ics = [ic.normal_cloud(n, 70 + k) for k, n in enumerate(sizes)]
    specs = [system_spec(n, L.BRUTE_FORCE_GLOBAL, eps=0.01) for n in sizes]
    with NBodyEnsemble(specs, law=L.LAW_PLUMMER, accumulation=L.ACCUM_KAHAN) as ens:
        for k, (p, v, m) in enumerate(ics):
            ens.upload(k, p, v, m)
        ens.step(1)
        mixed = [ens.download(k) for k in range(len(sizes))]
    for k, (n, (p, v, m)) in enumerate(zip(sizes, ics)):
        with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, law=L.LAW_PLUMMER, accumulation=L.ACCUM_KAHAN,
                                  eps=0.01)) as s:
            s.upload(p, v, m)
            s.step(1)
            alone = s.download()
        for a, b in zip(alone, mixed[k]):
            assert a.tobytes() == b.tobytes(), f"system {k} differs"


@pytest.mark.parametrize("algo,law", [(L.BARNES_HUT, L.LAW_PLUMMER), (L.BRUTE_FORCE_GLOBAL, L.LAW_REFERENCE_UNIT),
                                      (L.BRUTE_FORCE_LOCAL, L.LAW_REFERENCE_UNIT)])
@pytest.mark.parametrize("n", [1, 2, 257, 3000])
def test_energy_matches_oracle(orc, gpu, algo, law, n):
    pos, vel, mass = ic.plummer(n, 20 + n)
    eps = 0.02
    with NBodySim(make_config(n, algo, law=law, eps=eps, G=1.5)) as s:
        s.upload(pos, vel, mass)
        e = s.energy()
    plummer = algo == L.BARNES_HUT or law == L.LAW_PLUMMER
    # the library holds eps as float32; the unit law's e is 0.1 or 0.01
    e_eps = float(np.float32(eps)) if plummer else (0.1 if algo == L.BRUTE_FORCE_LOCAL else 0.01)
    o = orc.energy_momentum(pos, vel, mass, e_eps, 1.5, law=0 if plummer else 1)
    for k in ("kinetic", "potential", "mass"):
        assert e[k] == pytest.approx(o[k], rel=1e-12, abs=1e-300), k
    scale = float(np.abs(mass[:, None] * vel).sum()) + 1e-300
    assert np.abs(e["momentum"] - o["momentum"]).max() <= 1e-12 * scale
    lscale = float((mass * np.linalg.norm(pos, axis=1) * np.linalg.norm(vel, axis=1)).sum()) + 1e-300
    assert np.abs(e["angular"] - o["angular"]).max() <= 1e-12 * lscale


@pytest.mark.parametrize("tpl,waves", [(1, 4096), (2, 16384)])
def test_tuned_brute_shape_keeps_parity(orc, gpu, tpl, waves):
    """Other launch shapes (as the autotuner may pick) stay within tolerance of the oracle and
    bit-identical between a system alone and in a batch."""
    L.set_tuning(bf_targets_per_lane=tpl, bf_min_waves=waves)
    try:
        sizes = [4096, 777, 20000]
        ics = [ic.normal_cloud(n, 90 + k) for k, n in enumerate(sizes)]
        specs = [system_spec(n, L.BRUTE_FORCE_GLOBAL, eps=0.01) for n in sizes]
        with NBodyEnsemble(specs, law=L.LAW_PLUMMER) as ens:
            for k, (p, v, m) in enumerate(ics):
                ens.upload(k, p, v, m)
            ens.step(1)
            mixed = [ens.download(k) for k in range(len(sizes))]
        for k, (n, (p, v, m)) in enumerate(zip(sizes, ics)):
            with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, law=L.LAW_PLUMMER, eps=0.01)) as s:
                s.upload(p, v, m)
                s.brute_force()
                _, _, a = s.download()
                s.upload(p, v, m)
                s.step(1)
                alone = s.download()
            a_o, absacc = orc.force_bf(p, m, L.LAW_PLUMMER, 0.01, 1.0)
            assert (np.linalg.norm(a - a_o, axis=1) <= ACC_RTOL * absacc).all()
            for x, y in zip(alone, mixed[k]):
                assert x.tobytes() == y.tobytes(), f"system {k} differs"
    finally:
        L.set_tuning()
    assert L.get_tuning() == {"bf_targets_per_lane": 0, "bf_min_waves": 8192}
