"""GPU parity of trees rebuilt over a used cell pool.

No pass clears the cell pool between steps: every creator of a cell (the first
phase in LDS, the frame builds, the queued insertion's chains) writes all eight
of its slots before the cell is reachable, and the queued insertion reads slots
coherently.  Here one simulation is uploaded with distributions of very
different shapes in turn, so each build allocates its cells over the previous
tree's: a slot left from the earlier tree would change the topology, the
counts or the interaction counts.  The "clumps" distribution sends buckets
over the frame budget and a frame over its LDS cell pool to the queued
insertion, whose chains of cells are the ones created while other lanes
descend.  Same bar as the single-build parity tests: tree, permutations and
opening decisions bit-exact, accelerations within 2e-5 * sum |term| per body.
"""
import numpy as np
import pytest

from nbodysimulation_amd import _lib as L
from nbodysimulation_amd.engine import NBodySim, make_config

from test_gpu_parity import ACC_RTOL, VG, compare_tree, dist, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rules", [0, 1])
def test_rebuilds_over_a_used_cell_pool(orc, gpu, rules):
    n = 70000
    theta, eps, G = 0.5, 0.01, 1.0
    s = NBodySim(make_config(n, L.BARNES_HUT, rules=rules, theta=theta, eps=eps, G=G))
    try:
        for kind, seed in [("plummer", 11), ("clumps", 12), ("cube", 13), ("clumps", 14), ("plummer", 15)]:
            say(f"case rules={rules} kind={kind} seed={seed} n={n}")
            pos, vel, mass = dist(kind, n, seed)
            s.upload(pos, vel, mass)
            box = s.force_pipeline()
            t = orc.Tree(pos, mass, rules, cap=s.tree_info()["capacity"])
            assert t.status == orc.OK
            assert box.tobytes() == t.box.tobytes()
            compare_tree(s, t, n)
            a_o, absacc, ni = t.force(theta, eps, G, vote=VG)
            _, _, a_g = s.download()
            assert s.interactions() == int(ni.sum())
            err = np.linalg.norm(a_g - a_o, axis=1)
            assert (err <= ACC_RTOL * absacc + 1e-30).all(), float((err / np.maximum(absacc, 1e-30)).max())
    finally:
        s.close()
