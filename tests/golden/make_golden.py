"""Writes tests/golden/fixtures.npz: small regression fixtures produced by the oracle
(the project's CPU restatement; the reference's own tests hold no vectors).

    python tests/golden/make_golden.py

Inputs are stored, not regenerated, so a fixture stays valid if the initial-condition
generators change.  Each case holds the inputs and the oracle's outputs:
  bh_<rules>_<kind>_<n>: box, thresholds, tree order, force order (the vote groups'
      Hilbert order), octant paths (first max-length digits), cells in preorder (level,
      slot, count, centre of mass and mass), per-body interactions of the 64-body vote
      groups, accelerations and sum |term| per body;
  bf_<law>_<n>: all-pairs accelerations and sum |term| per body;
  step_bh_<rules>_<n>: positions and velocities after 3 Euler Barnes-Hut steps.
The fixtures pin regressions of the oracle (tests/test_golden.py, CPU) and of the HIP
path (the same file, GPU)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nbodysimulation_amd import ic  # noqa: E402  (host-side generators; no device needed)
from oracle import oracle as orc  # noqa: E402

VOTE = 64
THETA, EPS, G, DT = 0.5, 0.01, 1.0, 1e-3

BH_CASES = [("plummer", 1024, 7), ("normal", 1000, 8), ("normal", 64, 9)]
BF_CASES = [(orc.LAW_REFERENCE_UNIT, 1000, 10), (orc.LAW_PLUMMER, 1000, 11)]


def inputs(kind, n, seed):
    pos, vel, mass = (ic.plummer if kind == "plummer" else ic.normal_cloud)(n, seed)
    return np.asarray(pos, np.float32), np.asarray(vel, np.float32), np.asarray(mass, np.float32)


def bh_case(rules, kind, n, seed):
    pos, vel, mass = inputs(kind, n, seed)
    t = orc.Tree(pos, mass, rules)
    assert t.status == orc.OK
    ex = t.export()
    maxlen = int(ex["path_len"].max())
    acc, absacc, ninter = t.force(THETA, EPS, G, vote=VOTE)
    return {"pos": pos, "vel": vel, "mass": mass, "box": t.box, "dq": t.dq(THETA, EPS),
            "sorted": ex["sorted"], "force_order": t.force_order, "path_len": ex["path_len"], "paths": ex["paths"][:, :maxlen].copy(),
            "cell_level": ex["cell_level"], "cell_slot": ex["cell_slot"], "cell_count": ex["cell_count"],
            "cell_com": ex["cell_com"], "ninter": ninter.astype(np.int32), "acc": acc, "absacc": absacc,
            "max_depth": np.int32(t.max_depth)}


def bf_case(law, n, seed):
    pos, vel, mass = inputs("normal", n, seed)
    acc, absacc = orc.force_bf(pos, mass, law, EPS, G)
    return {"pos": pos, "mass": mass, "acc": acc, "absacc": absacc}


def step_case(rules, n, seed):
    pos, vel, mass = inputs("plummer", n, seed)
    p, v = pos.copy(), vel.copy()
    for _ in range(3):
        t = orc.Tree(p, mass, rules)
        a, _, _ = t.force(THETA, EPS, G, vote=VOTE)
        v, p = orc.euler(a, v, p, DT)
    return {"pos": pos, "vel": vel, "mass": mass, "pos3": p, "vel3": v}


def main():
    out = {}
    for rules in (orc.RULES_REFERENCE, orc.RULES_STANDARD):
        for kind, n, seed in BH_CASES:
            for k, v in bh_case(rules, kind, n, seed).items():
                out[f"bh_{rules}_{kind}_{n}/{k}"] = v
        for k, v in step_case(rules, 512, 12).items():
            out[f"step_bh_{rules}_512/{k}"] = v
    for law, n, seed in BF_CASES:
        for k, v in bf_case(law, n, seed).items():
            out[f"bf_{law}_{n}/{k}"] = v
    path = os.path.join(HERE, "fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
