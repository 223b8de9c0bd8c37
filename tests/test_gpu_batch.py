"""Batched and served forms: every system's trajectory is bit-identical to the
oracle's serial replay of that system alone (the same trajectory as the plain
updater), whatever the batch composition, chunking, graph use or admission
order."""
import numpy as np
import pytest

from oracle import oracle as orc
from pscbfm_amd import Batch, Server, member_bytes, systems, updater_from_system
from pscbfm_amd.batch import JOB_CANCELLED, JOB_DONE, JOB_FAILED

pytestmark = pytest.mark.gpu


def mixed_systems(k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for j in range(k):
        L = int(rng.choice([16, 32, 64]))
        N = int(rng.choice([8, 16, 32]))
        phi = rng.uniform(0.3, 0.5)
        chains = max(1, int(phi * L ** 3 / (8 * N)))
        chains = min(chains, L ** 3 // (16 * N))
        out.append((systems.linear_melt(L, chains, N, seed=seed * 1000 + j), seed * 7919 + j, int(rng.integers(0, 5000))))
    out.append((systems.diamond_network(64, (2, 2, 8), 32, seed=seed), seed + 17, 123))
    return out


def oracle_final(system, seed, age, n_mcs):
    want, acc, prop = orc.sweep(orc.System(system, seed), system["xyz"], age, n_mcs, 0)
    return want, acc, prop


@pytest.mark.parametrize("graphs,xcd_ranges", [(True, True), (False, True), (True, False)])
def test_batch_members_match_oracle(gpu_ready, graphs, xcd_ranges):
    cases = mixed_systems(6, seed=3)
    print(f"case: batch of {len(cases)} systems, graphs={graphs} xcd_ranges={xcd_ranges}", flush=True)
    b = Batch(0)
    b.set_graphs(graphs)
    b.set_xcd_ranges(xcd_ranges)
    ids = []
    for s, seed, age in cases:
        u = updater_from_system(s, seed=seed, age=age, validate_only=True)
        ids.append(b.add(u, len(s["xyz"])))
    b.run(60, chunk=25)     # chunks of 25, 25, 10
    b.run(90, chunk=30)
    for m, (s, seed, age) in zip(ids, cases):
        got = b.positions(m)
        want, acc, prop = oracle_final(s, seed, age, 150)
        assert np.array_equal(got, want), f"member {m} differs"
        st = b.stats(m)
        assert (st["accepted"], st["proposals"]) == (acc, prop)
        assert b.age(m) == age + 150


@pytest.mark.parametrize("group", [2, 4, 8])
def test_batch_dilute_members_fill_every_perform_round(gpu_ready, group):
    """Members whose groups of check waves hold more survivors than a perform
    wave's first rounds: free monomers and dimers at 6 % volume fraction,
    beside a dense melt, under every perform group."""
    cases = [(systems.linear_melt(64, 2000, 1, seed=5), 0x5EED0000000000D3, 0),
             (systems.linear_melt(64, 1000, 2, seed=6), 0x5EED0000000000D4, 7),
             (systems.linear_melt(32, 128, 16, seed=7), 0x5EED0000000000D5, 3)]
    print(f"case: dilute batch members, perform group {group}", flush=True)
    b = Batch(0)
    b.set_perform_group(group)
    ids = [b.add(updater_from_system(s, seed=seed, age=age, validate_only=True), len(s["xyz"]))
           for s, seed, age in cases]
    b.run(50, chunk=20)
    for m, (s, seed, age) in zip(ids, cases):
        want, acc, prop = oracle_final(s, seed, age, 50)
        assert np.array_equal(b.positions(m), want), f"member {m} differs"
        assert (b.stats(m)["accepted"], b.stats(m)["proposals"]) == (acc, prop)


def four_melts_in_a_slab(periodic):
    """Four 16^3 melts (coordinates within [0, 14]) side by side in a 64 x 16 x 32 box."""
    s = systems.linear_melt(16, 16, 16, seed=31)
    parts, bonds = [], []
    for k, off in enumerate([(0, 0, 0), (16, 0, 0), (32, 0, 16), (48, 0, 16)]):
        parts.append(s["xyz"] + np.array(off, dtype=np.int32))
        bonds.append(s["bonds"] + k * len(s["xyz"]))
    return {"box": (64, 16, 32), "periodic": periodic, "xyz": np.concatenate(parts),
            "bonds": np.concatenate(bonds), "forbidden": s["forbidden"]}


def test_batch_kernel_classes(gpu_ready):
    """One member of every kernel class (periodic or walled, cubic box or not) and a network
    in one batch, the classes launched side by side in every slot; each member ends where the
    oracle's replay of it alone ends."""
    melt = systems.linear_melt(32, 100, 16, seed=6)
    cases = [(melt, 0x5EED0000000000E1, 0),                                   # periodic, cubic
             (dict(melt, periodic=(0, 1, 0)), 0x5EED0000000000E2, 40),         # walls, cubic
             (four_melts_in_a_slab((1, 1, 1)), 0x5EED0000000000E3, 7),          # periodic, non-cubic
             (four_melts_in_a_slab((1, 0, 1)), 0x5EED0000000000E4, 0),          # walls, non-cubic
             (systems.diamond_network(64, (2, 2, 8), 32, seed=5), 0x5EED0000000000E5, 3)]
    print(f"case: batch of the four kernel classes and a network", flush=True)
    b = Batch(0)
    ids = [b.add(updater_from_system(s, seed=seed, age=age, validate_only=True), len(s["xyz"]))
           for s, seed, age in cases]
    b.run(120, chunk=40)
    for m, (s, seed, age) in zip(ids, cases):
        want, acc, prop = oracle_final(s, seed, age, 120)
        assert np.array_equal(b.positions(m), want), f"member {m} differs"
        st = b.stats(m)
        assert (st["accepted"], st["proposals"]) == (acc, prop)
        assert st["accepted"] > 0


def test_batch_with_four_colour_members(gpu_ready):
    """Members with 4 colours (tetrahedral cluster gels), 2 colours (melts, the
    diamond network) and 4 colours with walls in one batch: in every slot a
    2-colour member idles for the colour slots it does not have, and each
    member ends where the oracle's replay of it alone ends."""
    cases = [(systems.cluster_gel(32, 0.8, 0.5, seed=51), 0x5EED0000000000D1, 0),
             (systems.cluster_gel(16, 0.9, 0.3, seed=52), 0x5EED0000000000D2, 77),
             (dict(systems.cluster_gel(64, 0.7, 0.6, seed=53), periodic=(0, 1, 1)), 0x5EED0000000000D3, 5),
             (systems.linear_melt(32, 100, 16, seed=54), 0x5EED0000000000D4, 0),
             (systems.diamond_network(64, (2, 2, 8), 32, seed=55), 0x5EED0000000000D5, 9)]
    print(f"case: batch with four-colour members, {len(cases)} members", flush=True)
    b = Batch(0)
    ids = []
    for s, seed, age in cases:
        u = updater_from_system(s, seed=seed, age=age, validate_only=True)
        colours, nc = u.getColors(len(s["xyz"]))
        own = orc.System(s, seed)
        assert nc == own.n_colors and np.array_equal(colours, own.colors)
        ids.append(b.add(u, len(s["xyz"])))
    assert [orc.System(s, 1).n_colors for s, _, _ in cases] == [4, 4, 4, 2, 2]
    b.run(90, chunk=40)     # 40, 40, 10
    for m, (s, seed, age) in zip(ids, cases):
        want, acc, prop = oracle_final(s, seed, age, 90)
        assert np.array_equal(b.positions(m), want), f"member {m} differs"
        st = b.stats(m)
        assert (st["accepted"], st["proposals"]) == (acc, prop)
        assert st["accepted"] > 0


def test_batch_members_resort_on_their_schedule(gpu_ready):
    """Members re-sorted on the device every 7, 11 and 30 MCS and one that never
    sorts in the run, together in chunks of 25 MCS: each re-sorts at the ages a
    plain updater would (the chunks are split there), keeps its trajectory and
    ends where the oracle's replay of it alone ends."""
    cases = [(systems.linear_melt(32, 100, 16, seed=71), 0x5EED0000000000F7, 0, 7),
             (systems.cluster_gel(32, 0.8, 0.5, seed=72), 0x5EED0000000000F8, 3, 11),
             (systems.diamond_network(64, (2, 2, 8), 32, seed=73), 0x5EED0000000000F9, 5, 30),
             (systems.linear_melt(64, 400, 32, seed=74), 0x5EED0000000000FA, 0, None)]
    print(f"case: batch with re-sorting members, {len(cases)} members", flush=True)
    b = Batch(0)
    ids = [b.add(updater_from_system(s, seed=seed, age=age, validate_only=True, steps_between_sortings=every),
                 len(s["xyz"])) for s, seed, age, every in cases]
    b.run(60, chunk=25)
    for m, (s, seed, age, every) in zip(ids, cases):
        want, acc, prop = oracle_final(s, seed, age, 60)
        assert np.array_equal(b.positions(m), want), f"member {m} differs"
        st = b.stats(m)
        assert (st["accepted"], st["proposals"]) == (acc, prop)
        period = every or 5000
        assert st["sorts"] == sum(1 for a in range(age + 1, age + 60) if a % period == 0)
    assert b.stats(ids[0])["sorts"] == 8 and b.stats(ids[3])["sorts"] == 0


def test_batch_subset_and_removal(gpu_ready):
    cases = mixed_systems(4, seed=9)
    b = Batch(0)
    ids = [b.add(updater_from_system(s, seed=seed, age=age, validate_only=True), len(s["xyz"]))
           for s, seed, age in cases]
    b.run_members([ids[0], ids[2]], 40, chunk=20)
    b.remove(ids[1])
    b.run(30, chunk=30)
    for k, (m, (s, seed, age)) in enumerate(zip(ids, cases)):
        if k == 1:
            continue
        mcs = 70 if k in (0, 2) else 30
        want, _, _ = oracle_final(s, seed, age, mcs)
        assert np.array_equal(b.positions(m), want)


def test_server_jobs_match_alone(gpu_ready):
    cases = mixed_systems(8, seed=21)
    mcs = [100, 300, 200, 100, 400, 200, 300, 100, 200]
    # a budget that holds only a few jobs at once forces admission waves
    bytes_each = [member_bytes(updater_from_system(s, seed=seed, age=age, validate_only=True))
                  for s, seed, age in cases]
    srv = Server(0, hbm_budget=int(2.5 * max(bytes_each)), chunk=100)
    jobs = []
    for (s, seed, age), n in zip(cases[:5], mcs):
        jobs.append(srv.submit(updater_from_system(s, seed=seed, age=age, validate_only=True), n, len(s["xyz"])))
    srv.step()
    # jobs arriving while others run join at the next round
    for (s, seed, age), n in zip(cases[5:], mcs[5:]):
        jobs.append(srv.submit(updater_from_system(s, seed=seed, age=age, validate_only=True), n, len(s["xyz"])))
    rounds = 1
    while srv.step() > 0:
        rounds += 1
        assert srv.info()["bytes_in_use"] <= srv.info()["budget"]
        assert rounds < 100
    for j, (s, seed, age), n in zip(jobs, cases, mcs):
        st, done = srv.state(j)
        assert st == JOB_DONE and done == n
        want, _, _ = oracle_final(s, seed, age, n)
        assert np.array_equal(srv.result(j), want), f"job {j} differs from running alone"
        t_sub, t_adm, t_fin = srv.times(j)
        assert 0 <= t_sub <= t_adm <= t_fin


def test_running_updater_joins_a_batch_and_a_server(gpu_ready):
    """An updater that was initialized and has run (positions copied back, layout re-sorted
    on the device meanwhile) joins a batch and a server: both continue its trajectory,
    ending where the updater alone ends."""
    s = systems.linear_melt(32, 100, 16, seed=88)
    n = len(s["xyz"])
    seed = 0x5EED0000000000E7
    print(f"case: running updater into a batch and a server, n={n}", flush=True)
    u = updater_from_system(s, seed=seed, gpu=0, steps_between_sortings=10)
    u.runSimulationOnGPU(20)
    u.runResident(10)            # no copy-back: the hand-over makes one
    assert u.getAge() == 30
    b = Batch(0)
    m = b.add(u, n)
    b.run(20, 10)
    srv = Server(0, hbm_budget=0, chunk=10)
    job = srv.submit(u, 20, n)
    rounds = 0
    while srv.step() > 0:
        rounds += 1
        assert rounds < 100
    u.runSimulationOnGPU(20)
    want = u.getMonomerPositions(n)
    u.cleanup()
    ref, _, _ = oracle_final(s, seed, 0, 50)
    assert np.array_equal(want, ref)
    assert b.age(m) == 50
    assert np.array_equal(b.positions(m), want)
    assert srv.state(job)[0] == JOB_DONE
    assert np.array_equal(srv.result(job), want)


def test_server_cancel_and_oversized(gpu_ready):
    cases = mixed_systems(3, seed=33)
    srv = Server(0, hbm_budget=0, chunk=50)
    jobs = [srv.submit(updater_from_system(s, seed=seed, age=age, validate_only=True), 200, len(s["xyz"]))
            for s, seed, age in cases]
    srv.step()
    srv.step()                       # 100 of 200 MCS done
    srv.cancel(jobs[1])
    assert srv.state(jobs[1])[0] == JOB_CANCELLED
    while srv.step() > 0:
        pass
    for k in (0, 2, 3):
        s, seed, age = cases[k]
        want, _, _ = oracle_final(s, seed, age, 200)
        assert np.array_equal(srv.result(jobs[k]), want)
    small = Server(0, hbm_budget=4096, chunk=50)
    s, seed, age = cases[0]
    j = small.submit(updater_from_system(s, seed=seed, age=age, validate_only=True), 100, len(s["xyz"]))
    assert small.state(j)[0] == JOB_FAILED
    assert "budget" in small.error(j)


def test_batch_autotune_keeps_the_trajectories(gpu_ready):
    """Autotuning runs the members under every launch configuration; each one
    ends where it would alone after the same number of MCS."""
    cases = [(systems.linear_melt(32, 128, 16, seed=85), 0x5EED0000000000F4),
             (systems.diamond_network(64, (2, 2, 8), 32, seed=41), 0x5EED0000000000F5),
             (systems.linear_melt(64, 400, 32, seed=86), 0x5EED0000000000F6)]
    print(f"case: batch autotune, {len(cases)} members", flush=True)
    b = Batch(0)
    ids = [b.add(updater_from_system(s, seed=seed, validate_only=True), len(s["xyz"])) for s, seed in cases]
    cfg = b.autotune(5, 5)
    assert cfg["perform_group"] in (2, 4, 8)
    b.run(20, 10)
    for (s, seed), m in zip(cases, ids):
        assert b.age(m) == 55       # one untimed probe + 6 configurations, 5 MCS each, then 20
        u = updater_from_system(s, seed=seed, gpu=0)
        u.runSimulationOnGPU(55)
        assert np.array_equal(b.positions(m), u.getMonomerPositions(len(s["xyz"])))
        u.cleanup()


def test_mixed_batch_full_size_matches_plain(gpu_ready):
    """The 512-system mixed workload at full size (20.3 M monomers; box edges 32, 64 and 128,
    chains of 16 to 128, volume fraction 0.3 to 0.5): 20 MCS as one batch, in chunks of 10,
    end where each system ends after 20 MCS alone."""
    from pscbfm_amd import workloads
    specs = workloads.mixed_batch(512, seed=0xBA7C4)
    built = [workloads.build(sp) for sp in specs]
    print(f"case: mixed batch at full size, {len(built)} systems, "
          f"{sum(len(x['xyz']) for x in built)} monomers", flush=True)
    b = Batch(0)
    ids = [b.add(updater_from_system(x, seed=sp["seed"], validate_only=True), len(x["xyz"]))
           for x, sp in zip(built, specs)]
    b.run(20, 10)
    b.synchronize()
    bad = []
    for x, sp, m in zip(built, specs, ids):
        u = updater_from_system(x, seed=sp["seed"], gpu=0)
        u.runSimulationOnGPU(20)
        if not np.array_equal(b.positions(m), u.getMonomerPositions(len(x["xyz"]))):
            bad.append(m)
        u.cleanup()
    assert not bad, f"{len(bad)} of {len(ids)} members differ from their plain runs"


def test_mixed_batch_full_run_1000_mcs(gpu_ready):
    """The 512-system mixed workload at its full size and length (20.3 M monomers, 10^3 MCS
    each, chunks of 100 MCS, 2.0e10 attempts; the members in working-set groups): every
    member ends where it ends alone, and the three smallest members where the oracle's
    replay ends."""
    from pscbfm_amd import workloads
    specs = workloads.mixed_batch(512, seed=0xBA7C4)
    built = [workloads.build(sp) for sp in specs]
    print(f"case: mixed batch at full size, {len(built)} systems, "
          f"{sum(len(x['xyz']) for x in built)} monomers, 1000 MCS in chunks of 100", flush=True)
    b = Batch(0)
    ids = [b.add(updater_from_system(x, seed=sp["seed"], validate_only=True), len(x["xyz"]))
           for x, sp in zip(built, specs)]
    # about 0.48 GB of hot arrays: run() sweeps 4 groups of members in turn,
    # each for all 1000 MCS
    assert b.working_set_groups() == 4
    b.run(1000, 100)
    b.synchronize()
    bad = []
    for x, sp, m in zip(built, specs, ids):
        u = updater_from_system(x, seed=sp["seed"], gpu=0)
        u.runResident(1000)
        u.copyBack()
        if not np.array_equal(b.positions(m), u.getMonomerPositions(len(x["xyz"]))):
            bad.append(m)
        u.cleanup()
    assert not bad, f"{len(bad)} of {len(ids)} members differ from their plain runs"
    for j in sorted(range(len(built)), key=lambda j: len(built[j]["xyz"]))[:3]:
        x, sp = built[j], specs[j]
        print(f"case: member {j} ({len(x['xyz'])} monomers) against the oracle", flush=True)
        want, _, _ = orc.sweep(orc.System(x, sp["seed"]), x["xyz"], 0, 1000, 0)
        assert np.array_equal(b.positions(ids[j]), want), f"member {j} differs from the oracle"
