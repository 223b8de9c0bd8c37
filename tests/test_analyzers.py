"""The analyzers' oracle pinned by hand-derived and closed-form cases, and the
product's host-side chain discovery against it (no GPU needed).

Reference: src/AnalyzerMsd.tpp (getLinearPolymerSystemInfo :186-288,
execute :303-421, OnlineStatistics :18-95).
"""
import math

import numpy as np
import pytest

from oracle import analyzers as ana
from oracle import oracle as orc
from pscbfm_amd import _capi, systems, updater_from_system


def rod_system(order, box=32, spacing=2):
    """One straight chain along x; order[k] is the id of the k-th monomer along it."""
    n = len(order)
    xyz = np.zeros((n, 3), dtype=np.int32)
    for k, m in enumerate(order):
        xyz[m] = (spacing * k, 4, 4)
    bonds = np.array([(order[k], order[k + 1]) for k in range(n - 1)], dtype=np.int64).reshape(-1, 2)
    return {"box": (box, box, box), "periodic": (1, 1, 1), "xyz": xyz, "bonds": bonds,
            "forbidden": orc.standard_bondset()}


def rods(n_chains, n, box=64):
    """n_chains parallel rods of n monomers (ids chain by chain), bond (2,0,0)."""
    xyz, bonds = [], []
    for c in range(n_chains):
        for k in range(n):
            xyz.append((2 * k, 4 * (c % 8), 4 * (c // 8)))
            if k:
                bonds.append((c * n + k - 1, c * n + k))
    return {"box": (box, box, box), "periodic": (1, 1, 1), "xyz": np.array(xyz, dtype=np.int32),
            "bonds": np.array(bonds, dtype=np.int64).reshape(-1, 2), "forbidden": orc.standard_bondset()}


def product_chain_table(system):
    u = updater_from_system(system, seed=1, gpu=None, validate_only=True)
    return u.chainTable()


def test_chain_walk_hand_derived():
    # chain 3-0-4-1-2, bonds listed (3,0), (0,4), (4,1), (1,2).  The walk
    # starts at the lowest id 0 (links [3, 4]), pushes 3 then 4, pops 4 -> 1
    # -> 2 (end, popped first) and then 3 (end): ends [2, 3].  Middle: 2 steps
    # from 2 -> 1 -> 4.
    s = rod_system([3, 0, 4, 1, 2])
    chains = ana.linear_chains(5, s["bonds"])
    assert chains == [[2, 1, 4, 0, 3]]
    ends = ana.chain_ends(chains)
    assert (ends["first"][0], ends["middle"][0], ends["last"][0]) == (2, 4, 3)
    t = product_chain_table(s)
    assert (t["first"][0], t["middle"][0], t["last"][0]) == (2, 4, 3)


def test_chain_walk_from_an_end():
    # ids along the chain: the lowest id is an end; it is popped first
    s = rod_system([0, 1, 2, 3, 4, 5])
    assert ana.linear_chains(6, s["bonds"]) == [[0, 1, 2, 3, 4, 5]]
    t = product_chain_table(s)
    assert (t["first"][0], t["middle"][0], t["last"][0]) == (0, 3, 5)


def test_product_chain_table_matches_oracle_on_melts():
    s = systems.linear_melt(32, 64, 16, seed=5)
    want = ana.chain_ends(ana.linear_chains(len(s["xyz"]), s["bonds"]))
    got = product_chain_table(s)
    for k in ("first", "middle", "last"):
        np.testing.assert_array_equal(got[k], want[k])


def _star():
    xyz = np.array([(8, 8, 8), (10, 8, 8), (8, 10, 8), (8, 8, 10)], dtype=np.int32)
    return {"box": (32, 32, 32), "periodic": (1, 1, 1), "xyz": xyz,
            "bonds": np.array([(0, 1), (0, 2), (0, 3)]), "forbidden": orc.standard_bondset()}


def _ring():
    xyz = np.array([(8, 8, 8), (10, 8, 8), (10, 10, 8), (8, 10, 8)], dtype=np.int32)
    return {"box": (32, 32, 32), "periodic": (1, 1, 1), "xyz": xyz,
            "bonds": np.array([(0, 1), (1, 2), (2, 3), (3, 0)]), "forbidden": orc.standard_bondset()}


@pytest.mark.parametrize("make, msg", [(_star, ana.ERR_BRANCHED), (_ring, ana.ERR_ENDS),
                                       (lambda: rod_system([0, 1]), ana.ERR_MIDDLE)])
def test_non_linear_systems_are_rejected_like_the_reference(make, msg):
    s = make()
    with pytest.raises(ValueError, match="^" + msg[:40]):
        ana.linear_chains(len(s["xyz"]), s["bonds"])
    with pytest.raises(_capi.InvalidArgument) as e:
        product_chain_table(s)
    assert str(e.value).endswith(msg)


def test_single_monomers_are_chains():
    s = {"box": (32, 32, 32), "periodic": (1, 1, 1), "xyz": np.array([(0, 0, 0), (4, 0, 0)], dtype=np.int32),
         "bonds": np.zeros((0, 2), dtype=np.int64), "forbidden": orc.standard_bondset()}
    assert ana.linear_chains(2, s["bonds"]) == [[0], [1]]
    t = product_chain_table(s)
    np.testing.assert_array_equal(t["first"], [0, 1])
    np.testing.assert_array_equal(t["middle"], [0, 1])
    np.testing.assert_array_equal(t["last"], [0, 1])


@pytest.mark.parametrize("n", [1, 3, 8, 17])
def test_rod_observables_closed_form(n):
    s = rods(5, n)
    o = ana.observables(s["xyz"], s["bonds"], len(s["xyz"]))
    assert o["n_chains"] == 5
    assert o["re2_sum"] == 5 * (2 * (n - 1)) ** 2
    # uniform spacing b: Rg^2 = b^2 (n^2 - 1) / 12
    assert math.isclose(o["rg2"], 4 * (n * n - 1) / 12, rel_tol=1e-12, abs_tol=1e-12)
    assert o["bonds"] == 5 * (n - 1)
    assert o["bond_hist"][2] == 5 * (n - 1) and o["bond_hist"].sum() == 5 * (n - 1)


def test_bond_histogram_orientation_and_sum_rule():
    # a bond is counted once, from the lower to the higher id: reversing the
    # ids of a chain mirrors every vector
    s = rod_system([2, 1, 0])
    o = ana.observables(s["xyz"], s["bonds"], 3)
    assert o["bond_hist"][(-2) & 7] == 2 and o["bond_hist"].sum() == 2
    m = systems.linear_melt(32, 64, 16, seed=9)
    o = ana.observables(m["xyz"], m["bonds"], len(m["xyz"]))
    allowed = np.nonzero(orc.standard_bondset() == 0)[0]
    assert o["bond_hist"].sum() == len(m["bonds"])
    assert set(np.nonzero(o["bond_hist"])[0]) <= set(allowed.tolist())


def test_chain_sums_reproduce_textbook_rg2():
    m = systems.linear_melt(32, 64, 16, seed=11)
    chains = ana.linear_chains(len(m["xyz"]), m["bonds"])
    sums = ana.chain_sums(m["xyz"], chains)
    rg2 = np.mean([(c["q"] - (np.array(c["s"]) ** 2).sum() / c["n"]) / c["n"] for c in sums])
    o = ana.observables(m["xyz"], m["bonds"], len(m["xyz"]))
    assert math.isclose(rg2, o["rg2"], rel_tol=1e-12)


def test_msd_bookkeeping_closed_form():
    # evaluations every MCS, a new subsequence every 2 MCS, ages 0..5:
    # subsequences start at 0, 2, 4; sample counts per lag (in chains):
    # lag 0: none (a subsequence is never evaluated when it starts), lag 1:
    # ages 1, 3, 5; lag 2: ages 2, 4; lag 3: ages 3, 5; lag 4: age 4; lag 5: age 5
    s = rods(4, 5)
    m = ana.MsdAnalyzer(len(s["xyz"]), s["bonds"], 0, 2, 1)
    for age in range(6):
        m.execute(age, s["xyz"])
    p = m.points()
    assert [q["dt"] for q in p] == [0, 1, 2, 3, 4, 5]
    assert [q["n_values"] for q in p] == [0, 12, 8, 8, 4, 4]
    assert all(math.isnan(v) for v in p[0]["mean"])
    # evaluations closer than deltaTEvaluate are skipped
    m2 = ana.MsdAnalyzer(len(s["xyz"]), s["bonds"], 0, 4, 2)
    for age in range(9):
        m2.execute(age, s["xyz"])
    assert [q["dt"] for q in m2.points()] == [0, 2, 4, 6, 8][: len(m2.points())]
    assert m2.first_mcs == [0, 4, 8]


@pytest.mark.parametrize("C", [1, 3, 6])
def test_msd_translation_closed_form(C):
    # chain 0 moves by e*d at the e-th evaluation, the others stay: with the
    # box-COM drift removed (C > 1) every chain sees a relative displacement,
    # mean over chains (C-1)/C^2 |e d|^2 for g1..g4; C = 1: |e d|^2; g5 = 0.
    s = rods(C, 7)
    n = len(s["xyz"])
    d = np.array([2, -4, 6])
    m = ana.MsdAnalyzer(n, s["bonds"], 0, 1000, 1)
    for e in range(4):
        x = s["xyz"].astype(np.int64).copy()
        x[:7] += e * d
        m.execute(e, x)
    for e, p in enumerate(m.points()):
        if e == 0:
            continue
        want = (e * e * (d * d).sum()) * ((C - 1) / C ** 2 if C > 1 else 1.0)
        for g in range(4):
            assert math.isclose(p["mean"][g], want, rel_tol=1e-12), (e, g)
        assert abs(p["mean"][4]) < 1e-18


def test_sg_column_is_the_bessel_scaled_raw_second_moment():
    # getCentralMoment(2) uses choose(3, i): 1*M2 - 3*m*m + 3*m^2 = M2, times
    # n/(n-1); a true variance would be (M2 - m^2) n/(n-1)
    st = ana.OnlineStatistics()
    xs = [1.0, 2.0, 4.0, 7.0]
    for x in xs:
        st.add_value(x)
    raw = sum(x * x for x in xs) / len(xs)
    assert math.isclose(st.sg(), math.sqrt(raw * len(xs) / (len(xs) - 1)), rel_tol=1e-12)
    assert not math.isclose(st.sg(), float(np.std(xs, ddof=1)), rel_tol=1e-3)
