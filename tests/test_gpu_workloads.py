"""The batched-ensemble and serving workloads, checked against the oracle at their own shapes.

Ensemble: the exact 512-system workload of the benchmark (N log-uniform in [1024, 65536],
Plummer / Normal(0,1) / unit-ball clouds, seed 2, the autotuned crossover) takes one batched
Euler step; every system is compared with the oracle on the same initial state:
  Barnes-Hut: on up to 16 sampled vote groups per system (every body of the small ones), the
    per-body interaction counts are equal (same vote groups, same opening decisions) and
    |a_gpu - a_oracle| <= 2e-5 * sum_k |term_k| per body;
  brute force: every body within the same bound, interaction count N - 1;
  both: the stepped positions and velocities equal the Euler update of the initial state with the
    GPU accelerations, bit for bit.
Then 16 fixed systems take 5 batched steps and equal the same systems stepped alone, bit for bit.

Serving: 64 requests of the seed-3 request stream (the largest request, 2,094,576 bodies,
among them; every theta) go through a server whose HBM budget and slot count force queueing,
with 1 to 3 steps each.  Every result equals the request run alone bit for bit; the 8 one-step
requests (the largest among them) match the oracle on that step as above.
"""
import numpy as np
import pytest

from nbodysimulation_amd import _lib as L
from nbodysimulation_amd import autotune
from nbodysimulation_amd import workloads as W
from nbodysimulation_amd.engine import NBodyEnsemble, NBodyServer, NBodySim, make_config

pytestmark = pytest.mark.gpu

ACC_RTOL = 2e-5
VG = L.vote_group()
GROUPS_PER_SYSTEM = 16
GiB = 1 << 30


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def tuned(gpu):
    """The benchmark's tuning (brute-force shape and crossover); the previous tuning is restored."""
    before = L.get_tuning()
    res = autotune.load_or_tune(0.5, 0.01)
    yield int(res["crossover_n"])
    L.set_tuning(**before)


def check_against_oracle(orc, spec, state, acc, inter, stepped=None, seed=0):
    """One system's first force pass (acc, per-body interactions) and, if given, its Euler-stepped
    (pos, vel) against the oracle on the initial state.  Returns the largest relative error."""
    p0, v0, m = state
    n = spec.num_bodies
    if spec.algorithm == L.BARNES_HUT:
        t = orc.Tree(p0, m, spec.rules, cap=2 * n + 1024)
        assert t.status == orc.OK
        ngroups = (n + VG - 1) // VG
        groups = np.arange(ngroups, dtype=np.int32)
        if ngroups > GROUPS_PER_SYSTEM:
            groups = np.sort(np.random.default_rng(seed).choice(ngroups, GROUPS_PER_SYSTEM, replace=False))
            groups = groups.astype(np.int32)
        a_o, absacc, ni = t.force(spec.theta, spec.epsilon, spec.gravitational_constant, vote=VG, groups=groups)
        idx = (groups[:, None] * VG + np.arange(VG)[None, :]).reshape(-1)
        members = t.force_order[idx[idx < n]]
        assert np.array_equal(inter[members], ni[members]), "opening decisions differ"
    else:
        a_o, absacc = orc.force_bf(p0, m, orc.LAW_PLUMMER, spec.epsilon, spec.gravitational_constant)
        members = np.arange(n)
        assert (inter == n - 1).all()
    err = np.linalg.norm(acc[members] - a_o[members], axis=1)
    assert (err <= ACC_RTOL * absacc[members]).all(), float((err / absacc[members]).max())
    if stepped is not None:
        v1, p1 = orc.euler(acc, v0, p0, spec.dt)
        assert stepped[0].tobytes() == p1.tobytes() and stepped[1].tobytes() == v1.tobytes(), "Euler update"
    return float((err / np.maximum(absacc[members], 1e-30)).max())


def run_alone(spec, steps, state):
    cfg = make_config(spec.num_bodies, spec.algorithm, rules=spec.rules, law=L.LAW_PLUMMER, theta=spec.theta,
                      eps=spec.epsilon, dt=spec.dt, G=spec.gravitational_constant,
                      cell_capacity=2 * spec.num_bodies + 1024)
    with NBodySim(cfg) as s:
        s.upload(*state)
        s.step(steps)
        return s.download()


def test_ensemble_workload_against_oracle(orc, tuned):
    sizes, kinds = W.ensemble_sizes_kinds()
    specs = W.ensemble_specs(sizes, tuned)
    ics = [W.ensemble_ic(i, n, k) for i, (n, k) in enumerate(zip(sizes, kinds))]
    nbh = sum(sp.algorithm == L.BARNES_HUT for sp in specs)
    say(f"ensemble workload: {len(specs)} systems, {int(sizes.sum())} bodies, crossover {tuned}, {nbh} Barnes-Hut")
    assert 0 < nbh < len(specs)  # both algorithms in the batch
    sample = np.sort(np.random.default_rng(7).choice(len(specs), 16, replace=False))
    with NBodyEnsemble(specs, law=L.LAW_PLUMMER) as ens:
        for i, st in enumerate(ics):
            ens.upload(i, *st)
        ens.step(1)
        assert (ens.status() == L.NB_OK).all()
        worst = 0.0
        for i, (sp, st) in enumerate(zip(specs, ics)):
            p1, v1, a = ens.download(i)
            worst = max(worst, check_against_oracle(orc, sp, st, a, ens.interactions_per_body(i), (p1, v1), seed=i))
        say(f"ensemble workload: every system matches the oracle after one step (max rel err {worst:.2e})")
        ens.step(4)
        assert (ens.status() == L.NB_OK).all()
        batched = {int(i): ens.download(int(i)) for i in sample}
    for i in sample:
        alone = run_alone(specs[i], 5, ics[i])
        for x, y in zip(alone, batched[int(i)]):
            assert x.tobytes() == y.tobytes(), f"system {i} (N={specs[i].num_bodies}) differs after 5 steps"
    say(f"ensemble workload: 16 systems batched == alone after 5 steps: {sample.tolist()}")


def serving_subsample(tuned):
    """64 requests of the 2000-request stream: the first 62, the largest, and the smallest not among them,
    with 1-3 steps; the one-step requests are the oracle-checked ones."""
    rq = W.serving_requests(2000, 7.0, tuned)
    idx = list(range(62))
    for extra in (int(np.argmax(rq.sizes)), int(np.argmin(rq.sizes))):
        if extra not in idx:
            idx.append(extra)
    idx = idx[:64]
    big = int(np.argmax(rq.sizes))
    one_step = {big} | set(idx[:7])
    steps = {i: 1 if i in one_step else 2 + (i % 2) for i in idx}
    return rq, idx, steps


def test_serving_workload_against_alone_and_oracle(orc, tuned):
    rq, idx, steps = serving_subsample(tuned)
    assert len(idx) == 64 and {float(rq.thetas[i]) for i in idx} == set(W.SERVING_THETAS)
    assert max(rq.sizes[i] for i in idx) >= 1 << 20
    say(f"serving workload: {len(idx)} requests, {int(sum(rq.sizes[i] for i in idx))} bodies, "
        f"largest {int(max(rq.sizes[i] for i in idx))}")
    states = {i: W.serving_ic(i, int(rq.sizes[i])) for i in idx}
    saw_queue = False
    with NBodyServer(int(1.5 * GiB), 24) as sv:
        ids = {}
        for k, i in enumerate(idx):
            ids[sv.submit(rq.specs[i], steps[i], *states[i])] = i
            if k % 8 == 7:
                sv.tick()  # requests arrive while others run
                saw_queue |= sv.stats()["queued"] > 0
        done = {}
        for _ in range(10000):
            sv.tick()
            saw_queue |= sv.stats()["queued"] > 0
            while (r := sv.poll()) is not None:
                done[r[0]] = r[1]
            st = sv.stats()
            if st["active"] == 0 and st["queued"] == 0:
                break
        assert saw_queue, "the budget was meant to force queueing"
        assert sorted(done) == sorted(ids) and all(s == L.NB_OK for s in done.values())
        got = {ids[rid]: sv.fetch(rid) for rid in ids}
    say("serving workload: drained; comparing with runs alone")
    checked = 0
    for i in idx:
        sp = rq.specs[i]
        cfg = make_config(sp.num_bodies, sp.algorithm, rules=sp.rules, law=L.LAW_PLUMMER, theta=sp.theta,
                          eps=sp.epsilon, dt=sp.dt, G=sp.gravitational_constant,
                          cell_capacity=2 * sp.num_bodies + 1024)
        with NBodySim(cfg) as s:
            s.upload(*states[i])
            s.step(steps[i])
            alone = s.download()
            inter = s.interactions_per_body() if steps[i] == 1 else None
        for x, y in zip(alone, got[i]):
            assert x.tobytes() == y.tobytes(), f"request {i} (N={sp.num_bodies}) differs from its run alone"
        if steps[i] == 1:
            e = check_against_oracle(orc, sp, states[i], got[i][2], inter, got[i][:2], seed=i)
            checked += 1
            say(f"serving workload: request {i} N={sp.num_bodies} theta={sp.theta:.1f} matches the oracle "
                f"(max rel err {e:.2e})")
    assert checked >= 8
