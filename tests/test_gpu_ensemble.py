"""Ensembles: every system of a batched ensemble gives bit-identical results to the
same system stepped alone (the batched kernels only add per-system offsets)."""
import numpy as np
import pytest

from nbodysimulation_amd import _lib as L
from nbodysimulation_amd import ic
from nbodysimulation_amd.engine import NBodyEnsemble, NBodySim, make_config, system_spec

pytestmark = pytest.mark.gpu


def mixed_systems():
    rng = np.random.default_rng(5)
    out = []
    kinds = [ic.plummer, ic.normal_cloud, ic.uniform_ball]
    for i, n in enumerate([1, 2, 100, 1024, 3000, 5000, 20000, 777]):
        algo = L.BARNES_HUT if i % 3 != 1 else L.BRUTE_FORCE_GLOBAL
        rules = int(rng.integers(0, 2))
        theta = float(rng.choice([0.3, 0.5, 0.7]))
        out.append((system_spec(n, algo, rules=rules, theta=theta, eps=0.01, dt=1e-3), kinds[i % 3](n, 40 + i)))
    return out


@pytest.mark.parametrize("method", [L.EULER, L.LEAPFROG])
def test_ensemble_equals_alone(gpu, method):
    systems = mixed_systems()
    steps = 5
    with NBodyEnsemble([s for s, _ in systems], method=method, law=L.LAW_PLUMMER) as ens:
        for i, (_, (p, v, m)) in enumerate(systems):
            ens.upload(i, p, v, m)
        ens.step(2)
        ens.step(steps - 2)
        assert (ens.status() == 0).all()
        mixed = [ens.download(i) for i in range(len(systems))]
    for i, (sp, (p, v, m)) in enumerate(systems):
        cfg = make_config(sp.num_bodies, sp.algorithm, method=method, rules=sp.rules, law=L.LAW_PLUMMER,
                          theta=sp.theta, eps=sp.epsilon, dt=sp.dt, G=sp.gravitational_constant,
                          cell_capacity=2 * sp.num_bodies + 1024)
        with NBodySim(cfg) as s:
            s.upload(p, v, m)
            s.step(steps)
            alone = s.download()
        for a, b in zip(alone, mixed[i]):
            assert a.tobytes() == b.tobytes(), f"system {i} differs"


def test_ensemble_failure_is_per_system(gpu):
    good = ic.plummer(1000, 1)
    bad_pos = good[0].copy()
    bad_pos[3, 0] = np.nan
    specs = [system_spec(1000), system_spec(1000)]
    with NBodyEnsemble(specs) as ens:
        ens.upload(0, *good)
        ens.upload(1, bad_pos, good[1], good[2])
        with pytest.raises(L.NbError):
            ens.step(1)
        st = ens.status()
        assert st[0] == L.NB_OK and st[1] == L.NB_E_NONFINITE
        p0, _, _ = ens.download(0)
        assert not np.array_equal(p0, good[0])  # the healthy system advanced
