"""Bit-exact parity of the HIP colour sweep with the CPU oracle's serial replay.

The reference's only correctness test is an end-to-end golden-file diff of a
simulated system (src/pscbfm/testCorrectness.sh:219-237); these tests do the
same against the oracle's replay of the same sweep with the same RNG stream:
final coordinates and the accepted / proposed move counts must match exactly.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from pscbfm_amd import systems, updater_from_system

pytestmark = pytest.mark.gpu


def run_pair(system, seed, chunks, age0=0, xcd_ranges=True, spatial=True, rng=orc.RNG_INT_HASH, sort_every=None,
             perform_group=None):
    n = len(system["xyz"])
    print(f"case: n={n} box={system['box']} periodic={system.get('periodic')} seed={seed:#x} "
          f"chunks={chunks} age0={age0} xcd_ranges={xcd_ranges} spatial={spatial} rng={rng} "
          f"sort_every={sort_every} perform_group={perform_group}", flush=True)
    u = updater_from_system(system, seed=seed, gpu=0, age=age0, spatial_sorting=spatial, rng=rng,
                            steps_between_sortings=sort_every)
    u.setXcdRanges(xcd_ranges)
    if perform_group is not None:
        u.setPerformGroup(perform_group)
    for c in chunks:
        u.runSimulationOnGPU(c)
    got = u.getMonomerPositions(n)
    colors, nc = u.getColors(n)
    stats = u.getStats()
    osys = orc.System(system, seed, colors, rng=rng)
    want, acc, prop = orc.sweep(osys, system["xyz"], age0, sum(chunks), 0)
    u.cleanup()
    return got, want, stats, acc, prop, osys


def assert_parity(system, seed, chunks, age0=0, xcd_ranges=True, spatial=True, rng=orc.RNG_INT_HASH,
                  sort_every=None, perform_group=None):
    got, want, stats, acc, prop, osys = run_pair(system, seed, chunks, age0, xcd_ranges, spatial, rng, sort_every,
                                                 perform_group)
    diff = np.nonzero((got != want).any(axis=1))[0]
    assert diff.size == 0, f"{diff.size} monomers differ, first {diff[:5]}: gpu {got[diff[:3]]} oracle {want[diff[:3]]}"
    assert stats["accepted"] == acc
    assert stats["proposals"] == prop
    assert stats["attempts"] == sum(chunks) * len(system["xyz"])
    rc, msg = orc.check_system(osys, got)
    assert rc == 0, msg
    return stats


def test_tiny_melt(gpu_ready):
    s = systems.linear_melt(16, 16, 16, seed=11)
    st = assert_parity(s, seed=0x1234, chunks=[200])
    assert st["accepted"] > 0


def test_small_melt_1000_mcs(gpu_ready):
    # the plain small-melt workload (64^3, 512 x N=32, 2 colours) for 1000 MCS
    s = systems.linear_melt(64, 512, 32, seed=0x5EED000000000001)
    assert_parity(s, seed=0x5EED000000000001, chunks=[1000])


def test_small_melt_full_run_10000_mcs(gpu_ready):
    # the small-melt configuration at its full length (10^4 MCS, 1.6e8
    # attempts), in the drop-in's save intervals; the oracle's reference is
    # its OpenMP colour sweep, equal to the serial replay (tests/test_oracle.py)
    s = systems.linear_melt(64, 512, 32, seed=0x5EED000000000001)
    n = len(s["xyz"])
    seed = 0x5EED000000000001
    print(f"case: small melt n={n} seed={seed:#x} 10 x 1000 MCS", flush=True)
    u = updater_from_system(s, seed=seed, gpu=0)
    for _ in range(10):
        u.runSimulationOnGPU(1000)
    got = u.getMonomerPositions(n)
    colors, _ = u.getColors(n)
    stats = u.getStats()
    u.cleanup()
    osys = orc.System(s, seed, colors)
    want, acc, prop = orc.sweep_parallel(osys, s["xyz"], 0, 10000, 0)
    diff = np.nonzero((got != want).any(axis=1))[0]
    assert diff.size == 0, f"{diff.size} monomers differ, first {diff[:5]}"
    assert stats["accepted"] == acc and stats["proposals"] == prop
    assert stats["attempts"] == 10000 * n
    rc, msg = orc.check_system(osys, got)
    assert rc == 0, msg


def test_round_robin_block_placement(gpu_ready):
    # the XCD placement changes which hardware runs a block, never the result
    s = systems.linear_melt(64, 512, 32, seed=21)
    assert_parity(s, seed=77, chunks=[300], xcd_ranges=False)


def test_id_order_layout(gpu_ready):
    # monomers stored in id order instead of Morton order: same trajectory
    s = systems.linear_melt(64, 512, 32, seed=22)
    assert_parity(s, seed=78, chunks=[300], spatial=False)


def test_periodic_resort_keeps_the_trajectory(gpu_ready):
    # on-device re-sort every 7 MCS (ages 7, 14, ..., 56) between and inside chunks
    s = systems.linear_melt(64, 512, 32, seed=24)
    st = assert_parity(s, seed=79, chunks=[25, 35], sort_every=7)
    assert st["sorts"] == 8


def test_periodic_resort_network_non_cubic(gpu_ready):
    # three colours, branch points (4 partners) and a non-cubic box: every
    # segment is re-sorted and every partner index remapped
    s = systems.diamond_network(64, (2, 2, 8), 32, seed=43)
    st = assert_parity(s, seed=44, chunks=[60], sort_every=5, age0=3)
    assert st["sorts"] == 12
    big = non_cubic_box_system()
    st = assert_parity(big, seed=45, chunks=[40], sort_every=10)
    assert st["sorts"] == 3


def test_resume_is_continuous(gpu_ready):
    # counter RNG keyed on the age: split runs equal one long run
    s = systems.linear_melt(32, 64, 16, seed=5)
    assert_parity(s, seed=99, chunks=[37, 63], age0=1000)


def test_philox_generator(gpu_ready):
    # the reference's Rng::Philox slot: Philox4x32-10 keyed on (seed; id, MCS)
    s = systems.linear_melt(64, 512, 32, seed=23)
    assert_parity(s, seed=0xC0FFEE, chunks=[200, 100], age0=77, rng=orc.RNG_PHILOX)


def test_hard_walls(gpu_ready):
    s = systems.linear_melt(32, 100, 16, seed=6)
    s = dict(s, periodic=(0, 1, 0))
    assert_parity(s, seed=7, chunks=[300])
    s = dict(s, periodic=(0, 0, 0))
    assert_parity(s, seed=8, chunks=[300])


def test_low_density_and_long_chains(gpu_ready):
    s = systems.linear_melt(64, 40, 128, seed=21)
    assert_parity(s, seed=22, chunks=[400])


def non_cubic_box_system():
    # four 16^3 melts side by side in a 64 x 16 x 32 box
    s = systems.linear_melt(16, 16, 16, seed=31)
    xyz = s["xyz"].copy()
    big = {"box": (64, 16, 32), "periodic": (1, 1, 1), "forbidden": s["forbidden"]}
    parts, bonds = [], []
    for k, off in enumerate([(0, 0, 0), (16, 0, 0), (32, 0, 16), (48, 0, 16)]):
        parts.append(xyz + np.array(off, dtype=np.int32))
        bonds.append(s["bonds"] + k * len(xyz))
    big["xyz"] = np.concatenate(parts)
    big["bonds"] = np.concatenate(bonds)
    return big


def tiled_system(box, periodic, axis, seed):
    """16^3 melts (coordinates within [0, 14]) tiled along one axis of a
    larger box, every tile its own seed."""
    parts, bonds, off = [], [], 0
    forbidden = None
    for k in range(box[axis] // 16):
        s = systems.linear_melt(16, 16, 16, seed=seed + k)
        shift = np.zeros(3, dtype=np.int32)
        shift[axis] = 16 * k
        parts.append(s["xyz"] + shift)
        bonds.append(s["bonds"] + off)
        off += len(s["xyz"])
        forbidden = s["forbidden"]
    return {"box": box, "periodic": periodic, "xyz": np.concatenate(parts), "bonds": np.concatenate(bonds),
            "forbidden": forbidden}


def test_largest_box_edge_periodic(gpu_ready):
    # 1024 sites along x, the largest edge (10-bit packed coordinates): chains
    # wrap from 1023 to 0 and their periodic images change
    s = tiled_system((1024, 16, 16), (1, 1, 1), 0, seed=61)
    st = assert_parity(s, seed=0x5EED0000000000D1, chunks=[150, 150])
    assert st["accepted"] > 0


def test_largest_box_edge_walls(gpu_ready):
    # walls on y with corners up to 1022 = L - 2 (the last legal corner)
    s = tiled_system((16, 1024, 16), (1, 0, 1), 1, seed=71)
    assert s["xyz"][:, 1].max() >= 1008   # the last tile touches the far wall
    assert_parity(s, seed=0x5EED0000000000D2, chunks=[300])


def test_smallest_box_and_unbonded_monomers(gpu_ready):
    # one monomer in the smallest box (8^3), then a gas of monomers without
    # bonds (no partner columns at all)
    forbidden = systems.standard_forbidden()
    one = {"box": (8, 8, 8), "periodic": (1, 1, 1), "xyz": np.array([[3, 4, 5]], dtype=np.int32),
           "bonds": np.zeros((0, 2), dtype=np.int64), "forbidden": forbidden}
    assert_parity(one, seed=0x5EED0000000000D3, chunks=[500])
    grid = np.array([(x, y, z) for x in range(0, 32, 4) for y in range(0, 32, 4) for z in range(0, 32, 4)],
                    dtype=np.int32)
    gas = {"box": (32, 32, 32), "periodic": (1, 1, 1), "xyz": grid, "bonds": np.zeros((0, 2), dtype=np.int64),
           "forbidden": forbidden}
    st = assert_parity(gas, seed=0x5EED0000000000D4, chunks=[200])
    assert st["accepted"] > 0


@pytest.mark.parametrize("group", [2, 8])
def test_perform_groups(gpu_ready, group):
    # check waves per perform wave: a launch parameter, never the trajectory
    s = systems.linear_melt(32, 128, 16, seed=81)
    assert_parity(s, seed=0x5EED0000000000F1, chunks=[150], perform_group=group)
    assert_parity(diamond_small(), seed=0x5EED0000000000F2, chunks=[60], perform_group=group)


def diamond_small():
    return systems.diamond_network(64, (2, 2, 8), 32, seed=41)


def test_autotune_keeps_the_trajectory(gpu_ready):
    s = systems.linear_melt(64, 512, 32, seed=83)
    n = len(s["xyz"])
    seed = 0x5EED0000000000F3
    print(f"case: autotune n={n}", flush=True)
    u = updater_from_system(s, seed=seed, gpu=0)
    # one untimed probe, 4 layout probes, 18 configurations, and 4 more when
    # the pick is not the default (compared by wall time): 5 MCS each
    cfg = u.autotune(5, n_colors=2)
    assert cfg["perform_group"] in (2, 4, 8)
    assert all(b in (64, 128, 256, 512, 1024) for b in cfg["check_block"])
    assert all(i in (1, 2) for i in cfg["check_items"])
    assert all(g in (2, 4, 8) for g in cfg["perform_groups"])
    age = u.getAge()
    assert age in (115, 135)
    u.runSimulationOnGPU(150 - age)
    got = u.getMonomerPositions(n)
    colors, _ = u.getColors(n)
    st = u.getStats()
    u.cleanup()
    want, acc, prop = orc.sweep(orc.System(s, seed, colors), s["xyz"], 0, 150, 0)
    assert np.array_equal(got, want)
    assert (st["accepted"], st["proposals"]) == (acc, prop)


@pytest.mark.parametrize("system", ["melt", "network", "walls"])
def test_every_launch_candidate_keeps_the_trajectory(gpu_ready, system):
    """Every candidate the autotuner chooses among -- check workgroup size 64,
    128, 256, 512 or 1024 monomers with one or two monomers per lane, perform
    group 2, 4 or 8, XCD placement on or off, Morton or id order within
    colours -- and different configurations per colour segment give the
    oracle's trajectory bit-exactly."""
    if system == "melt":
        s = systems.linear_melt(32, 128, 16, seed=91)
    elif system == "network":
        s = diamond_small()
    else:
        s = dict(systems.cluster_gel(32, 0.8, 0.5, seed=92), periodic=(1, 0, 1))
    n = len(s["xyz"])
    seed = 0x5EED0000000000E9
    mcs = 12
    want, acc, prop = orc.sweep(orc.System(s, seed), s["xyz"], 0, mcs, 0)
    configs = [(b, i, g, x, m) for b in (64, 128, 256, 512, 1024) for i in (1, 2) for g in (2, 4, 8)
               for x in (True, False) for m in (True, False) if b > 64 or i == 1]
    print(f"case: {system}, n={n}, {len(configs)} launch configurations + mixed per-segment ones", flush=True)
    for b, i, g, x, m in configs:
        u = updater_from_system(s, seed=seed, gpu=0, spatial_sorting=m)
        u.setCheckBlock(b)
        u.setCheckItems(i)
        u.setPerformGroup(g)
        u.setXcdRanges(x)
        u.runSimulationOnGPU(mcs)
        st = u.getStats()
        got = u.getMonomerPositions(n)
        assert u.launchConfig(0)["check_items"] == i
        u.cleanup()
        assert np.array_equal(got, want), (b, i, g, x, m)
        assert (st["accepted"], st["proposals"]) == (acc, prop), (b, i, g, x, m)
    u = updater_from_system(s, seed=seed, gpu=0)
    _, nc = u.getColors(n)
    blocks, groups, items = (1024, 64, 512, 128), (8, 2, 4, 2), (2, 1, 1, 2)
    for c in range(nc):
        u.setSegmentConfig(c, blocks[c % 4], groups[c % 4], items[c % 4])
    assert [u.launchConfig(c)["check_block"] for c in range(nc)] == [blocks[c % 4] for c in range(nc)]
    assert [u.launchConfig(c)["check_items"] for c in range(nc)] == [items[c % 4] for c in range(nc)]
    assert [u.launchConfig(c)["perform_group"] for c in range(nc)] == [groups[c % 4] for c in range(nc)]
    u.runSimulationOnGPU(mcs)
    assert np.array_equal(u.getMonomerPositions(n), want)
    u.cleanup()


@pytest.mark.parametrize("group", [2, 4, 8])
def test_dilute_systems_fill_every_perform_round(gpu_ready, group):
    """Free monomers (93 % of proposals survive) and dimers (53 %) at 6 %
    volume fraction: a group of g check waves holds far more survivors than
    the perform wave's first rounds (64 for groups of 2 and 4, 128 for 8), so
    the rounds after them run too."""
    free = systems.linear_melt(64, 2000, 1, seed=5)
    st = assert_parity(free, seed=0x5EED0000000000D1, chunks=[40, 25], perform_group=group)
    assert st["proposals"] > 0.9 * st["attempts"]
    dimers = systems.linear_melt(64, 1000, 2, seed=6)
    st = assert_parity(dimers, seed=0x5EED0000000000D2, chunks=[40, 25], perform_group=group)
    assert st["proposals"] > 0.5 * st["attempts"]


def test_non_cubic_box(gpu_ready):
    # compact Morton order for a 64 x 16 x 32 box
    assert_parity(non_cubic_box_system(), seed=32, chunks=[250])


def test_large_melt_full_size_10_mcs(gpu_ready):
    # the large-melt workload (512^3, 16384 x N=512) at full size, 10 MCS
    s = systems.linear_melt(512, 16384, 512, seed=0x5EED000000000002)
    assert_parity(s, seed=0x5EED000000000002, chunks=[10])


def test_large_melt_full_run_1000_mcs(gpu_ready):
    # the large-melt configuration at its full length (10^3 MCS, 8.4e9
    # attempts) in 100-MCS drop-in calls, against the oracle's OpenMP colour
    # sweep (equal to the serial replay, tests/test_oracle.py); the oracle
    # takes about 45 s on 16 host threads
    s = systems.linear_melt(512, 16384, 512, seed=0x5EED000000000002)
    n = len(s["xyz"])
    seed = 0x5EED000000000002
    print(f"case: large melt n={n} seed={seed:#x} 10 x 100 MCS", flush=True)
    u = updater_from_system(s, seed=seed, gpu=0)
    for _ in range(10):
        u.runSimulationOnGPU(100)
    got = u.getMonomerPositions(n)
    colors, _ = u.getColors(n)
    stats = u.getStats()
    u.cleanup()
    print("running the oracle's OpenMP sweep for 1000 MCS", flush=True)
    osys = orc.System(s, seed, colors)
    want, acc, prop = orc.sweep_parallel(osys, s["xyz"], 0, 1000, 0)
    diff = np.nonzero((got != want).any(axis=1))[0]
    assert diff.size == 0, f"{diff.size} monomers differ, first {diff[:5]}"
    assert stats["accepted"] == acc and stats["proposals"] == prop
    rc, msg = orc.check_system(osys, got)
    assert rc == 0, msg


def test_network_three_colours(gpu_ready):
    s = systems.diamond_network(64, (2, 2, 8), 32, seed=41)
    st = assert_parity(s, seed=42, chunks=[150, 150])
    assert st["accepted"] > 0


@pytest.mark.parametrize("pattern", ["forward", "alternating"])
def test_one_way_links(gpu_ready, pattern):
    """Links given one way only (setConnectivity is directed): only the listing
    monomer tests the bond, the other side keeps a mirror entry current.
    The oracle replays the same directed links with the product's colours;
    re-sorts every 20 MCS move the entries."""
    s = systems.linear_melt(32, 128, 16, seed=55)
    b = np.asarray(s["bonds"], dtype=np.int64)
    if pattern == "forward":
        links = b
    else:
        links = np.where((np.arange(len(b)) % 2 == 0)[:, None], b, b[:, ::-1])
    s = dict(s, links=links)
    n = len(s["xyz"])
    seed = 0x5EED0000000000B1
    print(f"case: one-way links {pattern} n={n}", flush=True)
    u = updater_from_system(s, seed=seed, gpu=0, steps_between_sortings=20)
    u.runSimulationOnGPU(60)
    u.runSimulationOnGPU(90)
    got = u.getMonomerPositions(n)
    colors, _ = u.getColors(n)
    stats = u.getStats()
    u.cleanup()
    osys = orc.System(s, seed, colors)
    want, acc, prop = orc.sweep(osys, s["xyz"], 0, 150, 1)
    diff = np.nonzero((got != want).any(axis=1))[0]
    assert diff.size == 0, f"{diff.size} monomers differ, first {diff[:5]}"
    assert (stats["accepted"], stats["proposals"]) == (acc, prop)
    assert stats["sorts"] > 0


def test_network_full_size_20_mcs(gpu_ready):
    # the end-linked network workload (256^3, 16384 crosslinkers, 32768 strands of N=32)
    s = systems.diamond_network(256, (8, 8, 32), 32, seed=0x5EED000000000003)
    assert len(s["xyz"]) == 1064960
    assert_parity(s, seed=0x5EED000000000003, chunks=[20])


def test_network_full_run_1000_mcs(gpu_ready):
    # the network configuration at its full length (10^3 MCS, 1.06e9
    # attempts) in 100-MCS drop-in calls, against the oracle's OpenMP colour
    # sweep (equal to the serial replay, tests/test_oracle.py)
    s = systems.diamond_network(256, (8, 8, 32), 32, seed=0x5EED000000000003)
    n = len(s["xyz"])
    seed = 0x5EED000000000003
    print(f"case: network n={n} seed={seed:#x} 10 x 100 MCS", flush=True)
    u = updater_from_system(s, seed=seed, gpu=0)
    for _ in range(10):
        u.runSimulationOnGPU(100)
    got = u.getMonomerPositions(n)
    colors, nc = u.getColors(n)
    stats = u.getStats()
    u.cleanup()
    assert nc == 2
    osys = orc.System(s, seed, colors)
    want, acc, prop = orc.sweep_parallel(osys, s["xyz"], 0, 1000, 0)
    diff = np.nonzero((got != want).any(axis=1))[0]
    assert diff.size == 0, f"{diff.size} monomers differ, first {diff[:5]}"
    assert stats["accepted"] == acc and stats["proposals"] == prop
    rc, msg = orc.check_system(osys, got)
    assert rc == 0, msg


@pytest.mark.parametrize("rng", [orc.RNG_INT_HASH, orc.RNG_PHILOX])
def test_four_colour_cluster_gel(gpu_ready, rng):
    """Tetrahedral clusters (every pair of a cluster bonded: the bond graph
    needs 4 colours) mixed with chains of four; the rebalance spreads the
    chains over all 4 colours.  Every MCS sweeps 4 colours in a permutation of
    4, 8 kernels per MCS."""
    s = systems.cluster_gel(32, 0.8, 0.5, seed=41)
    n = len(s["xyz"])
    u = updater_from_system(s, seed=1, validate_only=True)
    colors, nc = u.getColors(n)
    own = orc.System(s, 1)
    assert nc == own.n_colors == 4
    assert np.array_equal(colors, own.colors)
    st = assert_parity(s, seed=0x5EED0000000000C4, chunks=[150, 250], rng=rng)
    assert st["accepted"] > 0
    assert st["substeps"] == 4 * 400


def test_four_colour_gel_with_resort_and_walls(gpu_ready):
    """A denser cluster gel between walls on one axis, re-sorted on the device
    every 9 MCS: still the oracle's trajectory."""
    s = dict(systems.cluster_gel(64, 0.9, 0.7, seed=42), periodic=(1, 0, 1))
    st = assert_parity(s, seed=0x5EED0000000000C5, chunks=[60, 40], sort_every=9)
    assert st["accepted"] > 0 and st["sorts"] >= 10
