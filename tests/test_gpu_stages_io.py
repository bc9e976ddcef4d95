"""Collision stage, vertex export and checkpoints through the C ABI."""
import numpy as np
import pytest

from nbodysimulation_amd import _lib as L
from nbodysimulation_amd import ic
from nbodysimulation_amd.engine import NBodySim, make_config

pytestmark = pytest.mark.gpu


def dense_cloud(n, seed, side=0.4):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-side / 2, side / 2, (n, 3)).astype(np.float32)
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    mass = np.full(n, 1.0 / n, np.float32)
    return pos, vel, mass


@pytest.mark.parametrize("n", [2, 65, 3000])
def test_collision_stage_bit_exact(orc, gpu, n):
    pos, vel, mass = dense_cloud(n, 5 + n, side=0.4 if n > 100 else 0.015)
    with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, collision=True)) as s:
        s.upload(pos, vel, mass)
        s.collision()
        p_g, v_g, _ = s.download()
    v_o = orc.collision(pos, vel)
    assert np.array_equal(p_g, pos)
    assert (v_o != vel).any()
    assert v_g.tobytes() == v_o.tobytes()


def test_step_runs_collision_after_update(gpu):
    """nb_step with collision = brute force, Euler update, then the collision stage."""
    n = 4000
    pos, vel, mass = dense_cloud(n, 3)
    cfg = make_config(n, L.BRUTE_FORCE_GLOBAL, law=L.LAW_PLUMMER, eps=0.01, dt=1e-3, collision=True)
    with NBodySim(cfg) as s:
        s.upload(pos, vel, mass)
        s.step(2)
        stepped = s.download()
        assert s.step_count() == 2
    with NBodySim(cfg) as s:
        s.upload(pos, vel, mass)
        for _ in range(2):
            s.brute_force()
            s.update()
            s.collision()
        staged = s.download()
    for a, b in zip(stepped, staged):
        assert a.tobytes() == b.tobytes()


def test_vertex_export(gpu):
    pos, vel, mass = ic.plummer(1000, 4)
    with NBodySim(make_config(1000, L.BARNES_HUT)) as s:
        s.upload(pos, vel, mass)
        v = s.export_vertices()
    assert np.array_equal(v[:, :3], pos) and (v[:, 3] == 1.0).all()


@pytest.mark.parametrize("algo,method", [(L.BARNES_HUT, L.LEAPFROG), (L.BRUTE_FORCE_GLOBAL, L.EULER)])
def test_checkpoint_resume_is_bit_exact(gpu, tmp_path, algo, method):
    n = 5000
    pos, vel, mass = ic.plummer(n, 6)
    cfg = make_config(n, algo, method=method, rules=L.RULES_STANDARD, law=L.LAW_PLUMMER, eps=0.01, dt=1e-3)
    path = str(tmp_path / "state.nbck")
    with NBodySim(cfg) as s:
        s.upload(pos, vel, mass)
        s.step(3)
        s.save(path)
        s.step(4)
        straight = s.download()
        assert s.step_count() == 7
    with NBodySim.from_checkpoint(path) as r:
        assert r.step_count() == 3
        r.step(4)
        resumed = r.download()
        assert r.step_count() == 7
    for a, b in zip(straight, resumed):
        assert a.tobytes() == b.tobytes()


def test_checkpoint_rejects_mismatch(gpu):
    pos, vel, mass = ic.plummer(100, 1)
    with NBodySim(make_config(100, L.BARNES_HUT)) as s:
        s.upload(pos, vel, mass)
        blob = s.checkpoint()
    with NBodySim(make_config(101, L.BARNES_HUT)) as t:
        with pytest.raises(L.NbError):
            t.restore(blob)
        with pytest.raises(L.NbError):
            t.restore(b"\0" * 64)


def test_cpp_binding_on_gpu(gpu, cpp_binding_exe):
    """The reference-side C++ binding (stage calls + nbody::App) runs end to end on the device."""
    import subprocess
    out = subprocess.run([cpp_binding_exe], check=True, capture_output=True, text=True, timeout=120).stdout
    assert out.strip() == "stages=0 app=ok n=2048"
