"""Target-range force passes: one GPU's share of a force pass when the targets are
sharded over G GPUs that each hold the whole tree (rank r computes
[r*N/G, (r+1)*N/G) of the force order).  The union of the shares equals the full
pass bit for bit, wherever the cuts fall, and bodies outside every share keep
their accelerations."""
import numpy as np
import pytest

from nbodysimulation_amd import _lib as L
from nbodysimulation_amd import ic
from nbodysimulation_amd.engine import NBodySim, make_config

pytestmark = pytest.mark.gpu


def cut_sets(n, seed):
    even8 = [n * r // 8 for r in range(9)]
    rng = np.random.default_rng(seed)
    ragged = sorted({0, n, 1, 63, 64, 65, n - 1, *rng.integers(2, n - 1, 8).tolist()})
    return [even8, ragged]


def tree_up_to_sort(s):
    s.bounds()
    s.tree_reset()
    s.tree_build()
    s.summarize()
    s.sort()


@pytest.mark.parametrize("rules", [L.RULES_STANDARD, L.RULES_REFERENCE])
def test_bh_shares_equal_full_pass(gpu, rules):
    n = 150_001
    pos, vel, mass = ic.plummer(n, 21)
    with NBodySim(make_config(n, L.BARNES_HUT, rules=rules, law=L.LAW_PLUMMER)) as s:
        s.upload(pos, vel, mass)
        s.force_pipeline()
        _, _, full = s.download()
        order = s.export_force_order()
        for cuts in cut_sets(n, 5 + rules):
            pieces = list(zip(cuts[:-1], cuts[1:]))
            left_out = len(pieces) // 2
            print(f"rules={rules} cuts={cuts} left out {pieces[left_out]}", flush=True)
            s.upload(pos, vel, mass)  # clears the accelerations
            tree_up_to_sort(s)
            assert np.array_equal(s.export_force_order(), order)
            for k, (a, b) in enumerate(pieces):
                if k != left_out:
                    s.force_range(a, b - a)
            _, _, got = s.download()
            missing = np.zeros(n, bool)
            missing[order[pieces[left_out][0]:pieces[left_out][1]]] = True
            assert got[~missing].tobytes() == full[~missing].tobytes()
            assert not got[missing].any()
            # the left-out share afterwards completes the full pass
            s.force_range(pieces[left_out][0], pieces[left_out][1] - pieces[left_out][0])
            assert s.download()[2].tobytes() == full.tobytes()


@pytest.mark.parametrize("algo,law", [(L.BRUTE_FORCE_GLOBAL, L.LAW_PLUMMER), (L.BRUTE_FORCE_GLOBAL, L.LAW_REFERENCE_UNIT),
                                      (L.BRUTE_FORCE_LOCAL, L.LAW_REFERENCE_UNIT)])
def test_bf_shares_equal_full_pass(gpu, algo, law):
    n = 20_000
    pos, vel, mass = ic.normal_cloud(n, 22)
    with NBodySim(make_config(n, algo, law=law)) as s:
        s.upload(pos, vel, mass)
        s.brute_force()
        _, _, full = s.download()
        for cuts in cut_sets(n, 9):
            s.upload(pos, vel, mass)
            pieces = list(zip(cuts[:-1], cuts[1:]))
            for a, b in pieces[1:]:
                s.brute_force_range(a, b - a)
            _, _, got = s.download()
            a0, b0 = pieces[0]
            assert got[b0:].tobytes() == full[b0:].tobytes()
            assert not got[a0:b0].any()


def test_range_checks(gpu):
    n = 3000
    pos, vel, mass = ic.plummer(n, 3)
    with NBodySim(make_config(n, L.BARNES_HUT, rules=L.RULES_STANDARD)) as s:
        s.upload(pos, vel, mass)
        with pytest.raises(L.NbError) as e:
            s.force_range(0, n)  # no tree yet
        assert e.value.status == L.NB_E_STATE
        with pytest.raises(L.NbError) as e:
            s.brute_force_range(0, n)  # not a brute-force simulation
        assert e.value.status == L.NB_E_STATE
        tree_up_to_sort(s)
        for first, count in [(-1, 10), (0, n + 1), (n, 1), (10, -1), (n - 5, 6)]:
            with pytest.raises(L.NbError) as e:
                s.force_range(first, count)
            assert e.value.status == L.NB_E_INVALID_ARG
        s.force_range(n, 0)
        s.force_range(17, 0)
        assert not s.download()[2].any()
        s.force_range(0, n)
        s.update()
        assert s.status() == L.NB_OK
