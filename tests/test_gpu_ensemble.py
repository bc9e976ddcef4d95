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


def test_ensemble_with_empty_and_single_body_systems(gpu):
    """Systems of 0 and 1 bodies (both algorithms) sit in a batch without disturbing the others."""
    specs = [system_spec(0), system_spec(1), system_spec(0, L.BRUTE_FORCE_GLOBAL),
             system_spec(1, L.BRUTE_FORCE_GLOBAL), system_spec(2000)]
    big = ic.plummer(2000, 8)
    with NBodyEnsemble(specs, law=L.LAW_PLUMMER) as ens:
        ens.upload(0, np.zeros((0, 3), np.float32), None, np.zeros(0, np.float32))
        ens.upload(1, np.array([[0.5, 0.25, -1.0]], np.float32), np.array([[1.0, 0.0, 0.0]], np.float32),
                   np.array([1.0], np.float32))
        ens.upload(2, np.zeros((0, 3), np.float32), None, np.zeros(0, np.float32))
        ens.upload(3, np.array([[0.0, 0.0, 0.0]], np.float32), np.array([[0.0, 2.0, 0.0]], np.float32),
                   np.array([1.0], np.float32))
        ens.upload(4, *big)
        ens.step(3)
        assert (ens.status() == L.NB_OK).all()
        p1, v1, a1 = ens.download(1)
        p3, v3, a3 = ens.download(3)
        mixed = ens.download(4)
    # a lone body feels no force and drifts: x += 3 dt v
    assert not a1.any() and not a3.any()
    assert np.allclose(p1[0], [0.5 + 3e-3, 0.25, -1.0], atol=1e-6)
    assert np.allclose(p3[0], [0.0, 6e-3, 0.0], atol=1e-6)
    with NBodySim(make_config(2000, L.BARNES_HUT, rules=L.RULES_STANDARD, law=L.LAW_PLUMMER,
                              cell_capacity=2 * 2000 + 1024)) as s:
        s.upload(*big)
        s.step(3)
        alone = s.download()
    for a, b in zip(alone, mixed):
        assert a.tobytes() == b.tobytes()


def test_step_async_then_status(gpu):
    pos, vel, mass = ic.plummer(5000, 12)
    with NBodySim(make_config(5000, L.BARNES_HUT, rules=L.RULES_STANDARD)) as s:
        s.upload(pos, vel, mass)
        s.step_async(4)
        assert s.status() == L.NB_OK
        assert s.step_count() == 4
        with NBodySim(make_config(5000, L.BARNES_HUT, rules=L.RULES_STANDARD)) as t:
            t.upload(pos, vel, mass)
            t.step(4)
            for a, b in zip(s.download(), t.download()):
                assert a.tobytes() == b.tobytes()


def test_collision_per_system_in_a_batch(gpu):
    """The collision stage runs for the systems whose spec asks for it, after the batch's
    integrator; each system equals the same system stepped alone with the same setting."""
    specs = [system_spec(3000, collision=True), system_spec(2000, L.BRUTE_FORCE_GLOBAL, collision=True),
             system_spec(2500), system_spec(1500, L.BRUTE_FORCE_GLOBAL)]
    states = [ic.uniform_ball(sp.num_bodies, 70 + i, radius=0.3) for i, sp in enumerate(specs)]
    with NBodyEnsemble(specs, law=L.LAW_PLUMMER) as ens:
        for i, st in enumerate(states):
            ens.upload(i, *st)
        ens.step(4)
        mixed = [ens.download(i) for i in range(len(specs))]
    for i, (sp, st) in enumerate(zip(specs, states)):
        cfg = make_config(sp.num_bodies, sp.algorithm, rules=sp.rules, law=L.LAW_PLUMMER,
                          cell_capacity=2 * sp.num_bodies + 1024, collision=bool(sp.collision))
        with NBodySim(cfg) as s:
            s.upload(*st)
            s.step(4)
            alone = s.download()
        for a, b in zip(alone, mixed[i]):
            assert a.tobytes() == b.tobytes(), f"system {i}"
    # the setting matters: a colliding ball differs from the same ball without the stage
    with NBodySim(make_config(3000, L.BARNES_HUT, rules=L.RULES_STANDARD, law=L.LAW_PLUMMER,
                              cell_capacity=2 * 3000 + 1024)) as s:
        s.upload(*states[0])
        s.step(4)
        assert s.download()[1].tobytes() != mixed[0][1].tobytes()
