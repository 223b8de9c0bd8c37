"""Device analyzers against the oracle (src/AnalyzerMsd.tpp restated in
oracle/analyzers.py) on the states of the oracle's own sweep replay.

Integer results (bond histogram, per-chain sums, the end-to-end sum, sample
counts, evaluation times) must match exactly.  Floating-point means are sums
of doubles in another order (a fixed device tree against the reference's
sequential sums), so they agree to a relative 1e-9.
"""
import math

import numpy as np
import pytest

from oracle import analyzers as ana
from oracle import oracle as orc
from pscbfm_amd import _capi, systems, updater_from_system

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def replay(system, seed, colors, ages):
    """Oracle states at the given ages (starting from age 0)."""
    osys = orc.System(system, seed, colors)
    x, age, out = system["xyz"], 0, []
    for a in ages:
        if a > age:
            x, _, _ = orc.sweep(osys, x, age, a - age, 0)
            age = a
        out.append(np.array(x, dtype=np.int64))
    return out


def compare_observables(got, want):
    assert got["bonds"] == want["bonds"]
    np.testing.assert_array_equal(got["bond_hist"], want["bond_hist"])
    assert got["n_chains"] == want["n_chains"]
    if want["n_chains"] == 0:
        assert math.isnan(got["re2"]) and math.isnan(got["rg2"])
        return
    assert got["re2_sum"] == want["re2_sum"]
    assert math.isclose(got["re2"], want["re2"], rel_tol=1e-15)
    assert math.isclose(got["rg2"], want["rg2"], rel_tol=RTOL)
    for a in range(3):
        assert math.isclose(got["box_com"][a], want["box_com"][a], rel_tol=RTOL, abs_tol=1e-9)


@pytest.mark.parametrize("sort_every", [None, 20])
def test_observables_and_chain_sums_match_oracle(gpu_ready, sort_every):
    s = systems.linear_melt(32, 128, 16, seed=21)
    n = len(s["xyz"])
    seed = 0x5EED0000000000A1
    print(f"case: melt n={n} sort_every={sort_every}", flush=True)
    u = updater_from_system(s, seed=seed, gpu=0, steps_between_sortings=sort_every)
    colors, _ = u.getColors(n)
    ages = [0, 30, 70]
    states = replay(s, seed, colors, ages)
    chains = ana.linear_chains(n, s["bonds"])
    for k, age in enumerate(ages):
        if age:
            u.runResident(age - ages[k - 1])    # no copy-back: the analyzers read the device state
        compare_observables(u.observe(), ana.observables(states[k], s["bonds"], n))
        got = u.chainSums(len(chains))
        want = ana.chain_sums(states[k], chains)
        for c, w in enumerate(want):
            assert got[c]["s"].tolist() == w["s"] and got[c]["q"] == w["q"], c
            assert got[c]["first"].tolist() == w["first"] and got[c]["middle"].tolist() == w["middle"]
            assert got[c]["last"].tolist() == w["last"] and got[c]["n"] == w["n"]
    u.cleanup()


def test_observables_of_a_network(gpu_ready):
    s = systems.diamond_network(64, (2, 2, 8), 32, seed=41)
    n = len(s["xyz"])
    seed = 0x5EED0000000000A2
    print(f"case: network n={n}", flush=True)
    u = updater_from_system(s, seed=seed, gpu=0)
    colors, _ = u.getColors(n)
    u.runResident(25)
    (state,) = replay(s, seed, colors, [25])
    compare_observables(u.observe(), ana.observables(state, s["bonds"], n))
    with pytest.raises(_capi.InvalidArgument):
        u.attachMsd(0, 10, 5, 4)
    u.cleanup()


@pytest.mark.parametrize("sort_every", [None, 30])
def test_msd_matches_oracle(gpu_ready, sort_every, tmp_path):
    s = systems.linear_melt(32, 96, 16, seed=23)
    n = len(s["xyz"])
    seed = 0x5EED0000000000A3
    print(f"case: msd n={n} sort_every={sort_every}", flush=True)
    u = updater_from_system(s, seed=seed, gpu=0, steps_between_sortings=sort_every)
    colors, _ = u.getColors(n)
    # evaluations every 10 MCS, subsequences every 20 (more than the 4
    # reserved: the snapshot store grows), runs of 5 MCS (every other call is
    # skipped by the evaluation interval)
    ages = list(range(0, 125, 5))
    states = replay(s, seed, colors, ages)
    m = ana.MsdAnalyzer(n, s["bonds"], 7, 20, 10)
    u.attachMsd(7, 20, 10, 4)
    m.execute(0, states[0])
    for k in range(1, len(ages)):
        u.runResident(5)
        u.executeMsd()
        m.execute(ages[k], states[k])
    got, want = u.msdPoints(), m.points()
    assert len(got) == len(want) == 13
    for e, (g, w) in enumerate(zip(got, want)):
        assert g["dt"] == w["dt"] and g["n_values"] == w["n_values"], e
        for j in range(5):
            for key in ("mean", "sg"):
                a, b = g[key][j], w[key][j]
                if math.isnan(b):
                    assert math.isnan(a), (e, j, key)
                else:
                    assert math.isclose(a, b, rel_tol=RTOL, abs_tol=1e-12), (e, j, key, a, b)
    path = tmp_path / "msd.dat"
    u.writeMsd(path)
    lines = path.read_text().splitlines()
    header = [l for l in lines if l.startswith("#")]
    assert header[:3] == ["# nMinMonomersPerChain = 16", "# nMaxMonomersPerChain = 16", "# nChains = 96"]
    assert header[4:10] == ["# nMcsSkipped = 7", "# deltaTRealization = 20", "# iMcsLastRealization = 120",
                            "# deltaTEvaluate = 10", "# iMcsLastEvaluation = 120", "# nSubsequences = 7"]
    rows = [l.split() for l in lines if not l.startswith("#")]
    assert len(rows) == 13 and all(len(r) == 12 for r in rows)
    for r, w in zip(rows, want):
        assert int(r[0]) == w["dt"] and float(r[1]) == w["n_values"]
        for j in range(5):
            if not math.isnan(w["mean"][j]):
                assert math.isclose(float(r[2 + 2 * j]), w["mean"][j], rel_tol=1e-5, abs_tol=1e-12)
    u.cleanup()


def test_observables_full_size_large_melt(gpu_ready):
    """512^3 melt of 16,384 chains x 512: the device reductions against numpy on
    the copied-back state (histogram and end-to-end sum exact)."""
    s = systems.linear_melt(512, 16384, 512, seed=0x5EED000000000002)
    n = len(s["xyz"])
    print(f"case: large melt n={n}", flush=True)
    u = updater_from_system(s, seed=0x5EED0000000000A4, gpu=0)
    u.runSimulationOnGPU(10)
    got = u.observe()
    x = u.getMonomerPositions(n).astype(np.int64)
    u.cleanup()
    b = np.asarray(s["bonds"], dtype=np.int64)
    lo, hi = np.minimum(b[:, 0], b[:, 1]), np.maximum(b[:, 0], b[:, 1])
    d = x[hi] - x[lo]
    idx = (d[:, 0] & 7) | ((d[:, 1] & 7) << 3) | ((d[:, 2] & 7) << 6)
    np.testing.assert_array_equal(got["bond_hist"], np.bincount(idx, minlength=512))
    L = s["chain_len"]
    c = x.reshape(-1, L, 3)
    e2e = c[:, -1] - c[:, 0]
    assert got["n_chains"] == n // L
    assert got["re2_sum"] == int((e2e * e2e).sum())
    com = c.mean(axis=1, keepdims=True)
    rg2 = float(((c - com) ** 2).sum(axis=2).mean())
    assert math.isclose(got["rg2"], rg2, rel_tol=RTOL)


def test_batch_member_observables_equal_plain(gpu_ready):
    """A batch member's device observables are those of the same system run
    alone: same state (the trajectories are bit-exact), same kernels, same
    reduction order, so every field is equal, floating point included."""
    from pscbfm_amd import Batch
    cases = [(systems.linear_melt(32, 128, 16, seed=31), 0x5EED0000000000C1),
             (systems.diamond_network(64, (2, 2, 8), 32, seed=41), 0x5EED0000000000C2),
             (systems.linear_melt(64, 300, 32, seed=33), 0x5EED0000000000C3)]
    print(f"case: batch of {len(cases)} members, 40 MCS", flush=True)
    b = Batch(0)
    ids = [b.add(updater_from_system(s, seed=seed, validate_only=True), len(s["xyz"])) for s, seed in cases]
    b.run(40, 20)
    for (s, seed), m in zip(cases, ids):
        u = updater_from_system(s, seed=seed, gpu=0)
        u.runResident(40)
        want, got = u.observe(), b.observe(m)
        u.cleanup()
        np.testing.assert_array_equal(got["bond_hist"], want["bond_hist"])
        for key in ("bonds", "n_chains", "re2_sum"):
            assert got[key] == want[key], key
        for key in ("re2", "rg2"):
            assert (math.isnan(got[key]) and math.isnan(want[key])) or got[key] == want[key], key
        np.testing.assert_array_equal(got["box_com"], want["box_com"])
