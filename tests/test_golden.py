"""Committed regression fixtures (tests/golden/fixtures.npz, written by
tests/golden/make_golden.py from the oracle).  The CPU tests recompute every case
with the oracle and require the stored outputs bit for bit; the GPU tests run the
HIP path on the stored inputs and hold it to the parity bar: box, thresholds, tree
order, octant paths, cells and interaction counts bit-exact, accelerations within
2e-5 x sum |term| per body, 3-step trajectories within 1e-5."""
import os

import numpy as np
import pytest

from nbodysimulation_amd import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "fixtures.npz")
THETA, EPS, G, DT, VOTE = 0.5, 0.01, 1.0, 1e-3, 64
ACC_RTOL = 2e-5


def cases(prefix):
    with np.load(FIXTURES) as z:
        keys = sorted({k.split("/")[0] for k in z.files if k.startswith(prefix)})
        return [(k, {f.split("/")[1]: z[f] for f in z.files if f.split("/")[0] == k}) for k in keys]


def rules_of(name):
    return int(name.split("_")[1])


@pytest.mark.parametrize("name,c", cases("bh_"), ids=lambda x: x if isinstance(x, str) else "")
def test_oracle_tree_and_force_unchanged(orc, name, c):
    t = orc.Tree(c["pos"], c["mass"], rules_of(name))
    assert t.status == orc.OK
    ex = t.export()
    assert t.box.tobytes() == c["box"].tobytes()
    assert t.dq(THETA, EPS).tobytes() == c["dq"].tobytes()
    assert t.max_depth == int(c["max_depth"])
    w = c["paths"].shape[1]
    for k in ("sorted", "path_len", "cell_level", "cell_slot", "cell_count"):
        assert np.array_equal(ex[k], c[k]), k
    assert np.array_equal(t.force_order, c["force_order"])
    assert np.array_equal(ex["paths"][:, :w], c["paths"])
    assert (ex["paths"][:, w:] == 0xFF).all()
    assert ex["cell_com"].tobytes() == c["cell_com"].tobytes()
    acc, absacc, ninter = t.force(THETA, EPS, G, vote=VOTE)
    assert np.array_equal(ninter, c["ninter"])
    assert acc.tobytes() == c["acc"].tobytes()
    assert absacc.tobytes() == c["absacc"].tobytes()


@pytest.mark.parametrize("name,c", cases("bf_"), ids=lambda x: x if isinstance(x, str) else "")
def test_oracle_brute_force_unchanged(orc, name, c):
    acc, absacc = orc.force_bf(c["pos"], c["mass"], int(name.split("_")[1]), EPS, G)
    assert acc.tobytes() == c["acc"].tobytes()
    assert absacc.tobytes() == c["absacc"].tobytes()


@pytest.mark.parametrize("name,c", cases("step_bh_"), ids=lambda x: x if isinstance(x, str) else "")
def test_oracle_steps_unchanged(orc, name, c):
    p, v = c["pos"].copy(), c["vel"].copy()
    for _ in range(3):
        a, _, _ = orc.Tree(p, c["mass"], int(name.split("_")[2])).force(THETA, EPS, G, vote=VOTE)
        v, p = orc.euler(a, v, p, DT)
    assert p.tobytes() == c["pos3"].tobytes() and v.tobytes() == c["vel3"].tobytes()


# ------------------------------------------------------------------ the HIP path on the fixtures

@pytest.mark.gpu
@pytest.mark.parametrize("name,c", cases("bh_"), ids=lambda x: x if isinstance(x, str) else "")
def test_gpu_matches_golden_tree_and_force(gpu, name, c):
    from nbodysimulation_amd.engine import NBodySim, make_config
    n = len(c["mass"])
    print(f"case {name} n={n}", flush=True)
    with NBodySim(make_config(n, L.BARNES_HUT, rules=rules_of(name), theta=THETA, eps=EPS, G=G)) as s:
        s.upload(c["pos"], c["vel"], c["mass"])
        box = s.force_pipeline()
        info = s.tree_info()
        srt, paths, plen = s.export_bodies()
        forder = s.export_force_order()
        lv, sl, ct, com = s.export_cells()
        dq = s.export_dq()
        ninter = s.interactions()
        _, _, acc = s.download()
    assert info["status"] == L.NB_OK and info["max_depth"] == int(c["max_depth"])
    assert box.tobytes() == c["box"].tobytes()
    assert dq.tobytes() == c["dq"].tobytes()
    w = c["paths"].shape[1]
    assert np.array_equal(srt, c["sorted"]) and np.array_equal(plen, c["path_len"])
    assert np.array_equal(forder, c["force_order"])
    assert np.array_equal(paths[:, :w], c["paths"])
    assert np.array_equal(lv, c["cell_level"]) and np.array_equal(sl, c["cell_slot"])
    assert np.array_equal(ct, c["cell_count"]) and com.tobytes() == c["cell_com"].tobytes()
    assert ninter == int(c["ninter"].astype(np.int64).sum())
    assert (np.linalg.norm(acc - c["acc"], axis=1) <= ACC_RTOL * c["absacc"]).all()


@pytest.mark.gpu
@pytest.mark.parametrize("name,c", cases("bf_"), ids=lambda x: x if isinstance(x, str) else "")
def test_gpu_matches_golden_brute_force(gpu, name, c):
    from nbodysimulation_amd.engine import NBodySim, make_config
    n = len(c["mass"])
    with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, law=int(name.split("_")[1]), eps=EPS, G=G)) as s:
        s.upload(c["pos"], None, c["mass"])
        s.brute_force()
        _, _, acc = s.download()
    assert (np.linalg.norm(acc - c["acc"], axis=1) <= ACC_RTOL * c["absacc"] + 1e-30).all()


@pytest.mark.gpu
@pytest.mark.parametrize("name,c", cases("step_bh_"), ids=lambda x: x if isinstance(x, str) else "")
def test_gpu_matches_golden_steps(gpu, name, c):
    from nbodysimulation_amd.engine import NBodySim, make_config
    n = len(c["mass"])
    with NBodySim(make_config(n, L.BARNES_HUT, rules=int(name.split("_")[2]), theta=THETA, eps=EPS, G=G,
                              dt=DT)) as s:
        s.upload(c["pos"], c["vel"], c["mass"])
        s.step(3)
        p, v, _ = s.download()
    assert np.abs(p - c["pos3"]).max() <= 1e-5 * (1 + np.abs(c["pos3"]).max())
    assert np.abs(v - c["vel3"]).max() <= 1e-5 * (1 + np.abs(c["vel3"]).max())
