"""GPU parity: the HIP stages (through the C ABI) against the oracle on the same
seeded inputs.

Bar: bounding box, tree topology (per-body octant paths, cell preorder levels,
slots and counts), tree-order permutation, cell centre-of-mass/mass and opening
thresholds are bit-exact; so is every opening decision (per-body interaction
counts equal).  Accelerations differ only through the hardware reciprocal
square root and summation order, so they are held to
|a_gpu - a_oracle| <= ACC_RTOL * sum_k |term_k| per body (ACC_RTOL = 2e-5).
The oracle emulates the kernel's vote groups (VG consecutive bodies of the force order).
"""
import numpy as np
import pytest

from nbodysimulation_amd import _lib as L
from nbodysimulation_amd import ic
from nbodysimulation_amd.engine import NBodySim, make_config

pytestmark = pytest.mark.gpu

ACC_RTOL = 2e-5
VG = L.vote_group()  # bodies per opening decision of the force kernel


def say(*a):
    print(*a, flush=True)


def dist(kind, n, seed):
    r = np.random.default_rng(seed)
    if kind == "normal":
        return ic.normal_cloud(n, seed)
    if kind == "plummer":
        return ic.plummer(n, seed)
    if kind == "clusters":
        k = max(n // 3, 1)
        pos = np.concatenate([r.normal(scale=1e-3, size=(k, 3)), r.normal(size=(n - k, 3)) + 5]).astype(np.float32)
        return pos, np.zeros_like(pos), r.uniform(0.5, 1.5, n).astype(np.float32)
    if kind == "cube":
        pos = r.uniform(-1, 1, size=(n, 3)).astype(np.float32)
        return pos, np.zeros_like(pos), np.full(n, 1.0 / max(n, 1), np.float32)
    if kind == "clumps":
        # Uniform cube with three clumps chosen by body index against the build's phases
        # (every 4096th body, every 64th, the rest): 700 bodies of the strided phase in one
        # 1e-3 cube and 1500 bodies of the last phase in another, each far more than a frame
        # holds (buckets of about 710 and 1480 bodies: the queued fallback insertion), and
        # 120 pairs 2e-6 apart in a third cube, a bucket that fits a frame but whose chains
        # of cells overflow the frame's LDS cell pool (the whole frame falls back).
        pos = r.uniform(-1, 1, size=(n, 3)).astype(np.float32)
        idx = np.arange(n)
        p1 = idx[(idx % 64 == 0) & (idx % 4096 != 0)][:700]
        p2 = r.choice(idx[idx % 64 != 0], 1500, replace=False)
        pos[p1] = (np.array([0.5, 0.5, 0.5]) + r.uniform(0, 1e-3, (len(p1), 3))).astype(np.float32)
        pos[p2] = (np.array([-0.5, -0.5, 0.5]) + r.uniform(0, 1e-3, (len(p2), 3))).astype(np.float32)
        p3 = r.choice(np.setdiff1d(idx[idx % 64 != 0], p2), 240, replace=False)
        c = np.array([0.5, -0.5, -0.5]) + r.uniform(0, 1e-3, (120, 3))
        d = r.normal(size=(120, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pos[p3[:120]] = (c + 1e-6 * d).astype(np.float32)
        pos[p3[120:]] = (c - 1e-6 * d).astype(np.float32)
        return pos, np.zeros_like(pos), np.full(n, 1.0 / n, np.float32)
    raise ValueError(kind)


def gpu_tree(pos, vel, mass, rules, theta=0.5, eps=0.01, G=1.0, cap=0, accum=L.ACCUM_PLAIN):
    n = len(mass)
    s = NBodySim(make_config(n, L.BARNES_HUT, rules=rules, theta=theta, eps=eps, G=G, cell_capacity=cap,
                             accumulation=accum))
    s.upload(pos, vel, mass)
    box = s.force_pipeline()
    return s, box


def compare_tree(s, t, n):
    info = s.tree_info()
    assert info["status"] == L.NB_OK
    assert info["num_cells"] == t.ncells
    assert info["max_depth"] == t.max_depth
    ex = t.export()
    srt, paths, plen = s.export_bodies()
    assert np.array_equal(plen, ex["path_len"])
    assert np.array_equal(paths, ex["paths"])
    assert np.array_equal(srt, ex["sorted"])
    assert np.array_equal(s.export_force_order(), t.force_order)
    lv, sl, ct, com = s.export_cells()
    assert np.array_equal(lv, ex["cell_level"])
    assert np.array_equal(sl, ex["cell_slot"])
    assert np.array_equal(ct, ex["cell_count"])
    assert com.tobytes() == ex["cell_com"].tobytes()


@pytest.mark.parametrize("rules", [0, 1])
@pytest.mark.parametrize("kind,n", [("normal", 1), ("normal", 2), ("normal", 63), ("normal", 64), ("normal", 65),
                                    ("plummer", 1000), ("clusters", 3000), ("cube", 4096), ("plummer", 65536),
                                    ("clumps", 70000)])
def test_tree_and_force_parity(orc, gpu, rules, kind, n):
    say(f"case rules={rules} kind={kind} n={n}")
    pos, vel, mass = dist(kind, n, 100 + n)
    theta, eps, G = 0.5, 0.01, 1.0
    s, box = gpu_tree(pos, vel, mass, rules, theta, eps, G)
    try:
        t = orc.Tree(pos, mass, rules, cap=s.tree_info()["capacity"])
        assert t.status == orc.OK
        assert box.tobytes() == t.box.tobytes()
        compare_tree(s, t, n)
        assert s.export_dq().tobytes() == orc.dq_table(rules, box, theta, eps).tobytes()
        a_o, absacc, ni = t.force(theta, eps, G, vote=VG)
        _, _, a_g = s.download()
        assert s.interactions() == int(ni.sum())
        err = np.linalg.norm(a_g - a_o, axis=1)
        assert (err <= ACC_RTOL * absacc + 1e-30).all(), float((err / np.maximum(absacc, 1e-30)).max())
    finally:
        s.close()


def test_kahan_force_parity(orc, gpu):
    pos, vel, mass = ic.plummer(20000, 9)
    s, box = gpu_tree(pos, vel, mass, 1, accum=L.ACCUM_KAHAN)
    try:
        t = orc.Tree(pos, mass, 1, cap=s.tree_info()["capacity"])
        a_o, absacc, _ = t.force(0.5, 0.01, 1.0, vote=VG, kahan=True)
        _, _, a_g = s.download()
        assert (np.linalg.norm(a_g - a_o, axis=1) <= ACC_RTOL * absacc).all()
    finally:
        s.close()


@pytest.mark.parametrize("algo,law,accum", [(L.BRUTE_FORCE_GLOBAL, L.LAW_REFERENCE_UNIT, 0),
                                            (L.BRUTE_FORCE_LOCAL, L.LAW_REFERENCE_UNIT, 0),
                                            (L.BRUTE_FORCE_GLOBAL, L.LAW_PLUMMER, 0),
                                            (L.BRUTE_FORCE_LOCAL, L.LAW_PLUMMER, 1)])
@pytest.mark.parametrize("n", [1, 2, 100, 4096, 4097])
def test_brute_force_parity(orc, gpu, algo, law, accum, n):
    say(f"case bf algo={algo} law={law} accum={accum} n={n}")
    pos, vel, mass = ic.normal_cloud(n, n)
    eps = 0.01
    ueps = 0.1 if algo == L.BRUTE_FORCE_LOCAL else 0.01  # brute_force_update.cl:56 / :116
    with NBodySim(make_config(n, algo, law=law, accumulation=accum, eps=eps, G=1.0)) as s:
        s.upload(pos, vel, mass)
        s.brute_force()
        _, _, a_g = s.download()
    a_o, absacc = orc.force_bf(pos, mass, law, eps if law == L.LAW_PLUMMER else ueps, 1.0, kahan=bool(accum))
    err = np.linalg.norm(a_g - a_o, axis=1)
    assert (err <= ACC_RTOL * absacc + 1e-30).all()


def test_brute_force_reference_workload_100_steps(orc, gpu):
    """Plain brute force, N=4096, Normal(0,1), m~U[0.5,1.5]/N, v=0, G=1, dt=1e-3, 100 Euler
    steps, seed 0, the reference's update-kernel law (unit vectors, eps 0.01)."""
    n, steps, dt = 4096, 100, 1e-3
    pos, vel, mass = ic.normal_cloud(n, 0)
    with NBodySim(make_config(n, L.BRUTE_FORCE_GLOBAL, law=L.LAW_REFERENCE_UNIT, G=1.0, dt=dt)) as s:
        s.upload(pos, vel, mass)
        s.step(steps)
        p_g, v_g, _ = s.download()
    p, v = pos.copy(), vel.copy()
    for _ in range(steps):
        a, _ = orc.force_bf(p, mass, orc.LAW_REFERENCE_UNIT, 0.01, 1.0)
        v, p = orc.euler(a, v, p, dt)
    # the unit law is smooth: 100 steps keep the trajectories within 1e-5 of each other
    assert np.abs(v_g - v).max() <= 1e-5 * np.abs(v).max()
    assert np.abs(p_g - p).max() <= 1e-5


@pytest.mark.parametrize("rules", [0, 1])
def test_step_matches_oracle_and_is_deterministic(orc, gpu, rules):
    n = 20000
    pos, vel, mass = ic.plummer(n, 4)
    theta, eps, G, dt = 0.5, 0.01, 1.0, 1e-3
    outs = []
    for use_graph in (True, False):
        with NBodySim(make_config(n, L.BARNES_HUT, rules=rules, theta=theta, eps=eps, G=G, dt=dt,
                                  use_graph=use_graph)) as s:
            s.upload(pos, vel, mass)
            s.step(1)
            outs.append(s.download())
            s.step(4)
            outs.append(s.download())
    for a, b in zip(outs[0], outs[2]):
        assert a.tobytes() == b.tobytes()  # graph replay == stream launches, bit for bit
    for a, b in zip(outs[1], outs[3]):
        assert a.tobytes() == b.tobytes()
    t = orc.Tree(pos, mass, rules)
    a_o, absacc, _ = t.force(theta, eps, G, vote=VG)
    v_o, p_o = orc.euler(a_o, vel, pos, dt)
    p_g, v_g, a_g = outs[0]
    assert (np.linalg.norm(a_g - a_o, axis=1) <= ACC_RTOL * absacc).all()
    assert np.abs(v_g - v_o).max() <= ACC_RTOL * absacc.max() * dt * 2
    assert np.abs(p_g - p_o).max() <= 1e-6 * (1 + np.abs(p_o).max())


@pytest.mark.parametrize("rules", [0, 1])
def test_tree_exact_after_steps_and_reupload(orc, gpu, rules):
    """After full force passes the sort visits the bodies in the last force order (not index
    order); the tree, both permutations and every opening decision stay those of the oracle
    on the current positions, also after new positions of the same N are uploaded."""
    n = 30000
    pos, vel, mass = ic.plummer(n, 17)
    other = ic.uniform_ball(n, 18)
    with NBodySim(make_config(n, L.BARNES_HUT, rules=rules, law=L.LAW_PLUMMER)) as s:
        s.upload(pos, vel, mass)
        s.step(3)
        for p_now, m_now in ((s.download()[0], mass), (other[0], other[2])):
            if p_now is other[0]:
                s.upload(*other)  # same N, unrelated positions: the old visiting order still applies
            say(f"case rules={rules} n={n} after steps, reupload={p_now is other[0]}")
            s.force_pipeline()
            t = orc.Tree(p_now, m_now, rules, cap=s.tree_info()["capacity"])
            assert t.status == orc.OK
            compare_tree(s, t, n)
            _, _, ni = t.force(0.5, 0.01, 1.0, vote=VG)
            assert s.interactions() == int(ni.sum())
            s.step(2)


def test_leapfrog_kdk(orc, gpu):
    n = 4096
    pos, vel, mass = ic.plummer(n, 8)
    theta, eps, G, dt = 0.5, 0.01, 1.0, 1e-3
    with NBodySim(make_config(n, L.BARNES_HUT, rules=1, method=L.LEAPFROG, theta=theta, eps=eps, G=G,
                              dt=dt)) as s:
        s.upload(pos, vel, mass)
        s.step(1)
        p_g, v_g, a_g = s.download()
    t0 = orc.Tree(pos, mass, 1)
    a0, _, _ = t0.force(theta, eps, G, vote=VG)
    vh = orc.kick(a0, vel, 0.5 * dt)
    p1 = orc.drift(vh, pos, dt)
    t1 = orc.Tree(p1, mass, 1)
    a1, absacc, _ = t1.force(theta, eps, G, vote=VG)
    v1 = orc.kick(a1, vh, 0.5 * dt)
    assert np.abs(p_g - p1).max() <= 1e-6 * (1 + np.abs(p1).max())
    assert (np.linalg.norm(a_g - a1, axis=1) <= 1e-4 * absacc + 1e-6).all()
    assert np.abs(v_g - v1).max() <= 1e-5 * (1 + np.abs(v1).max())


def _binary(ecc, eps):
    """Two bodies of mass 0.5 at separation 1 (apocentre of an orbit of eccentricity ecc),
    Plummer-softened law, G = 1; also the circular angular velocity."""
    m = np.array([0.5, 0.5], np.float32)
    a0 = 1.0 / (1.0 + eps * eps) ** 1.5
    vrel = np.sqrt(a0) * np.sqrt(1.0 - ecc)
    p = np.array([[-0.5, 0, 0], [0.5, 0, 0]], np.float32)
    v = np.array([[0, -0.5 * vrel, 0], [0, 0.5 * vrel, 0]], np.float32)
    return p, v, m, float(np.sqrt(a0))


def _leapfrog_run(algo, p, v, m, dt, nsteps, eps=0.01):
    extra = dict(rules=L.RULES_STANDARD, theta=0.5) if algo == L.BARNES_HUT else dict(law=L.LAW_PLUMMER)
    with NBodySim(make_config(2, algo, method=L.LEAPFROG, eps=eps, G=1.0, dt=dt, **extra)) as s:
        s.upload(p, v, m)
        s.step(nsteps)
        p1, v1, _ = s.download()
        e = s.energy()["total"]
    return p1, v1, e


@pytest.mark.parametrize("algo", [L.BRUTE_FORCE_GLOBAL, L.BARNES_HUT])
def test_leapfrog_properties_on_gpu(gpu, algo):
    """The engine's LEAPFROG (brute force: fused closing kick; Barnes-Hut: k_update's opening
    kick + drift and closing kick) pinned without the oracle: a two-body orbit (Barnes-Hut
    interacts body with body there, so both paths see the exact pair force) returns to its start
    after 200 steps forward and 200 with negated velocities, and its energy and phase errors at
    t = 2 shrink 4x per halving of dt.  Semi-implicit Euler or a drift with the pre-kick velocity
    would give ratios near 2 and miss the start by 3e-2 or more."""
    eps, T = 0.01, 2.0
    p0, v0, m, _ = _binary(0.5, eps)
    p, v, _ = _leapfrog_run(algo, p0, v0, m, 0.01, 200)
    assert np.abs(p - p0).max() > 0.1
    pb, vb, _ = _leapfrog_run(algo, p, -v, m, 0.01, 200)
    say(f"leapfrog algo={algo}: reversal error {np.abs(pb - p0).max():.2e}")
    assert np.abs(pb - p0).max() < 2e-5 and np.abs(vb + v0).max() < 2e-5
    _, _, E0 = _leapfrog_run(algo, p0, v0, m, 0.01, 0)
    dts = (0.04, 0.02, 0.01)
    de = [abs(_leapfrog_run(algo, p0, v0, m, dt, int(round(T / dt)))[2] - E0) for dt in dts]
    pc, vc, mc, w = _binary(0.0, eps)
    dph = []
    for dt in dts:
        pe = _leapfrog_run(algo, pc, vc, mc, dt, int(round(T / dt)))[0].astype(np.float64)
        sep = pe[1] - pe[0]
        dph.append(abs((np.arctan2(sep[1], sep[0]) - w * T + np.pi) % (2 * np.pi) - np.pi))
    say(f"leapfrog algo={algo}: energy errors {de}, phase errors {dph}")
    for errs in (de, dph):
        for a, b in zip(errs, errs[1:]):
            assert 3.3 <= a / b <= 4.7, errs


# ---------------------------------------------------------------- failure handling (forced from the host)

def test_tree_overflow_reported(gpu):
    pos, vel, mass = ic.plummer(5000, 1)
    with NBodySim(make_config(5000, L.BARNES_HUT, cell_capacity=64)) as s:
        s.upload(pos, vel, mass)
        with pytest.raises(L.NbError) as e:
            s.step(1)
        assert e.value.status == L.NB_E_TREE_OVERFLOW
        p, v, _ = s.download()
        assert p.tobytes() == pos.tobytes()  # a failed step leaves the state untouched


def test_coincident_bodies_depth_overflow(gpu):
    pos = np.zeros((3, 3), np.float32)
    pos[2] = 1.0
    with NBodySim(make_config(3, L.BARNES_HUT, rules=1)) as s:
        s.upload(pos, None, np.ones(3, np.float32))
        with pytest.raises(L.NbError) as e:
            s.step(1)
        assert e.value.status == L.NB_E_DEPTH_OVERFLOW


def test_nonfinite_reported_and_cleared(gpu):
    pos, vel, mass = ic.plummer(1000, 2)
    bad = pos.copy()
    bad[17, 1] = np.nan
    with NBodySim(make_config(1000, L.BARNES_HUT)) as s:
        s.upload(bad, vel, mass)
        with pytest.raises(L.NbError) as e:
            s.step(1)
        assert e.value.status == L.NB_E_NONFINITE
        s.clear_status()
        s.upload(pos, vel, mass)
        s.step(2)
        assert s.status() == L.NB_OK


def test_stage_order_enforced(gpu):
    pos, vel, mass = ic.plummer(100, 2)
    with NBodySim(make_config(100, L.BARNES_HUT)) as s:
        s.upload(pos, vel, mass)
        with pytest.raises(L.NbError) as e:
            s.force()
        assert e.value.status == L.NB_E_STATE
        s.bounds()
        with pytest.raises(L.NbError):
            s.summarize()


def test_zero_bodies(gpu):
    with NBodySim(make_config(0, L.BARNES_HUT)) as s:
        s.upload(np.zeros((0, 3), np.float32), None, np.zeros(0, np.float32))
        s.step(3)


# ---------------------------------------------------------------- full size (1M Plummer)

@pytest.mark.parametrize("rules", [0, 1])
def test_full_size_tree_exact(orc, gpu, rules):
    """N = 1,048,576 Plummer (seed 1): bit-exact tree, sort and box against the oracle,
    exact interaction counts and accelerations within tolerance on 64 sampled vote groups."""
    n = 1 << 20
    say(f"case full-size rules={rules}")
    pos, vel, mass = ic.plummer(n, 1)
    s, box = gpu_tree(pos, vel, mass, rules, theta=0.5, eps=0.01, G=1.0)
    try:
        assert box.tobytes() == np.concatenate([pos.min(0), pos.max(0)]).astype(np.float32).tobytes()
        t = orc.Tree(pos, mass, rules, cap=s.tree_info()["capacity"])
        assert t.status == orc.OK
        compare_tree(s, t, n)
        groups = np.random.default_rng(0).choice(n // VG, 64, replace=False).astype(np.int32)
        a_o, absacc, ni = t.force(0.5, 0.01, 1.0, vote=VG, groups=groups)
        _, _, a_g = s.download()
        members = t.force_order[(groups[:, None] * VG + np.arange(VG)[None, :]).reshape(-1)]
        err = np.linalg.norm(a_g[members] - a_o[members], axis=1)
        assert (err <= ACC_RTOL * absacc[members]).all()
    finally:
        s.close()


TOP_CELLS = 1280  # cells of the LDS tree that the first insertion phase (every 4096th body) may use


@pytest.mark.parametrize("rules", [0, 1])
def test_first_phase_overflow_keeps_the_tree_exact(orc, gpu, rules):
    """The first insertion phase (every 4096th body) is built in LDS unless its tree needs more
    than TOP_CELLS cells; then the phase runs in global memory.  Here every odd 4096th body of a
    1M Plummer system sits a few ulps from the even one before it, so the phase's bodies need
    thousands of chain cells (counted by the oracle on the CPU below): the tree must still equal
    the oracle's bit for bit."""
    n = 1 << 20
    say(f"case first-phase overflow rules={rules}")
    pos, vel, mass = ic.plummer(n, 1)
    pos = pos.copy()
    even = np.arange(0, n, 2 * 4096)
    odd = even + 4096
    pos[odd] = np.nextafter(np.nextafter(pos[even], np.float32(np.inf)), np.float32(np.inf))
    box = np.concatenate([pos.min(0), pos.max(0)]).astype(np.float32)
    first = orc.Tree(pos[::4096], mass[::4096], rules, box=box, cap=1 << 20)
    assert first.status == orc.OK and first.ncells > TOP_CELLS  # the LDS tree overflows
    s, gbox = gpu_tree(pos, vel, mass, rules, theta=0.5, eps=0.01, G=1.0)
    try:
        assert gbox.tobytes() == box.tobytes()
        t = orc.Tree(pos, mass, rules, cap=s.tree_info()["capacity"])
        assert t.status == orc.OK
        compare_tree(s, t, n)
    finally:
        s.close()


def test_largest_served_system_tree_exact(orc, gpu):
    """N = 2,097,152 (the largest system the serving workload admits), STANDARD rules, with
    the server's pool sizing (2N + 1024 cells): bit-exact tree and sort against the oracle."""
    n = 1 << 21
    say(f"case 2M n={n}")
    pos, vel, mass = ic.plummer(n, 5)
    s, box = gpu_tree(pos, vel, mass, 1, theta=0.5, eps=0.01, G=1.0, cap=2 * n + 1024)
    try:
        t = orc.Tree(pos, mass, 1, cap=2 * n + 1024)
        assert t.status == orc.OK
        compare_tree(s, t, n)
    finally:
        s.close()


@pytest.mark.parametrize("rules", [0, 1])
def test_four_million_tree_exact(orc, gpu, rules):
    """N = 4,194,304 Plummer (seed 7): the first insertion phase (1024 bodies) exceeds the LDS
    build's 512 and runs in global memory, then the two bucketed phases; bit-exact tree and
    sort against the oracle, exact interaction counts and accelerations within tolerance on 32
    sampled vote groups."""
    n = 1 << 22
    say(f"case 4M rules={rules}")
    pos, vel, mass = ic.plummer(n, 7)
    s, box = gpu_tree(pos, vel, mass, rules, theta=0.5, eps=0.01, G=1.0)
    try:
        t = orc.Tree(pos, mass, rules, cap=s.tree_info()["capacity"])
        assert t.status == orc.OK
        compare_tree(s, t, n)
        groups = np.random.default_rng(1).choice(n // VG, 32, replace=False).astype(np.int32)
        a_o, absacc, _ = t.force(0.5, 0.01, 1.0, vote=VG, groups=groups)
        _, _, a_g = s.download()
        members = t.force_order[(groups[:, None] * VG + np.arange(VG)[None, :]).reshape(-1)]
        err = np.linalg.norm(a_g[members] - a_o[members], axis=1)
        assert (err <= ACC_RTOL * absacc[members]).all()
    finally:
        s.close()
