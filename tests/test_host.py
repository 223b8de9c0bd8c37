"""Host logic of the product library, no GPU needed: the C ABI loads and
exports every declared symbol, the setters and initialize's checks raise what
the reference raises, colouring and the generators are right."""
import re
from pathlib import Path

import numpy as np
import pytest

import pscbfm_amd
from oracle import oracle as orc
from pscbfm_amd import BfmError, InvalidArgument, systems, updater_from_system
from pscbfm_amd import _capi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bfm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = pscbfm_amd.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} declared in include/ but not exported"
    assert names <= set(_capi.SIGNATURES), "binding table misses a declared function"


def test_last_error_is_thread_local_message():
    u = pscbfm_amd.UpdaterGPUScBFM_AB_Type()
    with pytest.raises(InvalidArgument) as e:
        u.setRng(5)
    assert "not available" in str(e.value)


def fresh():
    return pscbfm_amd.UpdaterGPUScBFM_AB_Type()


def test_setter_errors_mirror_reference():
    u = fresh()
    with pytest.raises(BfmError, match="power of 2"):
        u.setLatticeSize(48, 64, 64)                 # runtime_error, .cu:2673
    u.setNrOfAllMonomers(3)
    with pytest.raises(BfmError, match="already set"):
        u.setNrOfAllMonomers(3)                      # runtime_error, .cu:2550-2556
    with pytest.raises(InvalidArgument):
        u.setMonomerCoordinates(3, 0, 0, 0)          # index out of range
    for _ in range(8):
        u.setConnectivity(0, 1)
    with pytest.raises(InvalidArgument, match="maximum amount of bonds"):
        u.setConnectivity(0, 2)                      # invalid_argument, .cu:2625-2633
    with pytest.raises(InvalidArgument):
        u.copyBondSet(4, 0, 0, True)                 # outside [-4,3]
    with pytest.raises(InvalidArgument):
        u.setStepsBetweenSortings(0)


def test_getters_return_what_was_set():
    """getRng / getStepsBetweenSortings (UpdaterGPUScBFM_AB_Type.h:509,514):
    the defaults (IntHash, .cu:1407; 5000 MCS, .cu:1379), then the set values."""
    u = pscbfm_amd.UpdaterGPUScBFM_AB_Type()
    assert u.getRng() == 1
    assert u.getStepsBetweenSortings() == 5000
    u.setRng(3)
    u.setStepsBetweenSortings(77)
    assert u.getRng() == 3
    assert u.getStepsBetweenSortings() == 77
    with pytest.raises(InvalidArgument):
        u.setRng(2)
    assert u.getRng() == 3


def valid_melt():
    return systems.linear_melt(16, 16, 16, seed=3)


def test_validate_accepts_melt_and_colours_like_oracle():
    s = systems.linear_melt(32, 64, 16, seed=4)
    u = updater_from_system(s, seed=1, validate_only=True)
    col, nc = u.getColors(len(s["xyz"]))
    o = orc.System(s, 1)
    assert nc == o.n_colors == 2
    assert np.array_equal(col, o.colors)


def test_product_colouring_matches_reference_vectors():
    """The product's colouring (host_system.cpp graph_colouring, run by
    tests/cpp/device_helpers_check) equals the reference's own graphColoring.tpp
    on every golden graph, with and without the uniform rebalance."""
    import json
    import subprocess
    exe = Path(__file__).resolve().parent / "cpp" / "device_helpers_check"
    assert exe.exists(), "built by __graft_entry__.build()"
    graphs = json.loads((Path(__file__).resolve().parent / "golden" / "coloring_vectors.json").read_text())["graphs"]
    text = "".join(f"{g['uniform']} {len(g['lists'])}\n" +
                   "".join(f"{len(l)} " + " ".join(map(str, l)) + "\n" for l in g["lists"]) for g in graphs)
    out = subprocess.run([str(exe), "colour"], input=text, capture_output=True, text=True, timeout=60, check=True)
    rows = out.stdout.strip().split("\n")
    assert len(rows) == len(graphs)
    for g, row in zip(graphs, rows):
        assert [int(v) for v in row.split()] == [g["n_colors"]] + g["colors"], (g["name"], g["uniform"])


@pytest.mark.parametrize("kind", ["network", "cluster_gel", "odd_chains", "one_way"])
def test_validate_colours_like_oracle(kind):
    """initialize()'s colouring equals the oracle's on generated systems: a
    diamond network (2 colours: the bond graph is bipartite), tetrahedral
    clusters with chains (4 colours, rebalanced), chains of odd length and
    one-way links (each link keeps both ends apart)."""
    if kind == "network":
        s, want = systems.diamond_network(64, (2, 2, 8), 32, seed=3), 2
    elif kind == "cluster_gel":
        s, want = systems.cluster_gel(32, 0.8, 0.5, seed=4), 4
    elif kind == "odd_chains":
        s, want = systems.linear_melt(32, 60, 7, seed=6), 2
    else:
        s = systems.linear_melt(32, 40, 8, seed=7)
        b = s["bonds"]
        s = dict(s, links=np.concatenate([b[0::2], b[1::2, ::-1]]))    # every bond listed once
        want = 2
    u = updater_from_system(s, seed=1, validate_only=True)
    col, nc = u.getColors(len(s["xyz"]))
    o = orc.System(s, 1)
    assert nc == o.n_colors == want
    assert np.array_equal(col, o.colors)


def test_validate_rejects_wrong_bond_set():
    s = valid_melt()
    f = s["forbidden"].copy()
    f[(2 & 7) | ((2 & 7) << 3)] = 0                   # allow (2,2,0): 109 vectors
    with pytest.raises(BfmError, match="Expected 108 allowed bonds, but got 109"):
        updater_from_system(dict(s, forbidden=f), seed=1, validate_only=True)


def test_validate_rejects_overlap_and_bad_bond():
    s = valid_melt()
    x = s["xyz"].copy()
    x[7] = x[100] + np.array([0, 1, 1])
    with pytest.raises(BfmError, match="overlaps"):
        updater_from_system(dict(s, xyz=x), seed=1, validate_only=True)
    x = s["xyz"].copy()
    # move the chain end 0 so that its bond to 1 becomes (2,2,0)
    x[0] = x[1] + np.array([2, 2, 0])
    occupied = {tuple(v) for v in x[1:]}
    if tuple(x[0]) not in occupied:
        with pytest.raises(BfmError):
            updater_from_system(dict(s, xyz=x), seed=1, validate_only=True)


def test_validate_rejects_walls_and_box():
    s = valid_melt()
    x = s["xyz"].copy()
    x[:, 2] += 1                                     # cubes reach z = L
    with pytest.raises(BfmError, match="walls"):
        updater_from_system(dict(s, xyz=x, periodic=(1, 1, 0)), seed=1, validate_only=True)
    with pytest.raises(BfmError, match="power of two"):
        tiny = {"box": (4, 4, 4), "xyz": np.zeros((1, 3), np.int32), "bonds": np.zeros((0, 2), np.uint32),
                "forbidden": s["forbidden"]}
        updater_from_system(tiny, seed=1, validate_only=True)


def test_initialize_without_device_fails_loudly():
    if pscbfm_amd.device_count() > 0:
        pytest.skip("a HIP device is visible")
    u = fresh()
    with pytest.raises(pscbfm_amd.NoDeviceError):
        u.setGpu(0)
    s = valid_melt()
    with pytest.raises(pscbfm_amd.NoDeviceError):
        updater_from_system(s, seed=1, gpu=None)


@pytest.mark.parametrize("box,chains,n", [(16, 16, 16), (32, 64, 32), (64, 100, 64), (64, 8, 1), (32, 30, 7)])
def test_generator_produces_valid_seeded_melts(box, chains, n):
    a = systems.linear_melt(box, chains, n, seed=10)
    b = systems.linear_melt(box, chains, n, seed=10)
    c = systems.linear_melt(box, chains, n, seed=11)
    assert np.array_equal(a["xyz"], b["xyz"])
    assert not np.array_equal(a["xyz"], c["xyz"]) or chains * n == 1
    o = orc.System(a, 1)
    rc, msg = orc.check_system(o, a["xyz"])
    assert rc == 0, msg
    assert len(a["bonds"]) == chains * (n - 1)


def test_generator_rejects_overfull_box():
    with pytest.raises(InvalidArgument):
        systems.linear_melt(16, 17, 16, seed=1)      # 17 chains of 16 in 16^3 exceeds 0.5


def test_large_melt_shape():
    s = systems.linear_melt(128, 256, 512, seed=2)
    assert systems.melt_volume_fraction(s) == 0.5
    o = orc.System(s, 1)
    assert orc.check_system(o, s["xyz"])[0] == 0


@pytest.mark.parametrize("box,cells,n", [(64, (2, 2, 8), 32), (64, (4, 4, 4), 13), (32, (2, 2, 4), 8)])
def test_network_generator_valid(box, cells, n):
    a = systems.diamond_network(box, cells, n, seed=5)
    b = systems.diamond_network(box, cells, n, seed=5)
    assert np.array_equal(a["xyz"], b["xyz"])
    k = 8 * cells[0] * cells[1] * cells[2]
    assert len(a["xyz"]) == k + 2 * k * n
    assert len(a["bonds"]) == 2 * k * (n + 1)
    deg = np.bincount(a["bonds"].ravel(), minlength=len(a["xyz"]))
    assert (deg[:k] == 4).all() and (deg[k:] == 2).all()     # 4-functional crosslinkers, linear strands
    o = orc.System(a, 1)
    rc, msg = orc.check_system(o, a["xyz"])
    assert rc == 0, msg
    u = updater_from_system(a, seed=1, validate_only=True)
    col, nc = u.getColors(len(a["xyz"]))
    assert nc == o.n_colors >= 2
    assert np.array_equal(col, o.colors)


def test_network_generator_rejects_parity():
    with pytest.raises(InvalidArgument, match="parity"):
        systems.diamond_network(64, (2, 2, 8), 31, seed=1)   # 32 steps vs spacing 4+4+1


def test_batch_and_server_need_a_device():
    if pscbfm_amd.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(pscbfm_amd.NoDeviceError):
        pscbfm_amd.Batch(0)
    with pytest.raises(pscbfm_amd.NoDeviceError):
        pscbfm_amd.Server(0)


def test_member_bytes_cover_the_layout():
    s = systems.linear_melt(64, 512, 32, seed=1)
    u = updater_from_system(s, seed=1, validate_only=True)
    b = pscbfm_amd.member_bytes(u)
    n = len(s["xyz"])
    assert b >= 16 * n + 4 * 8 * n + 4 * 64 ** 3 // 8   # positions, partner slots, four bit lattices
    assert b < 200 * n + 8 * 64 ** 3


def test_kernel_timing_interval_must_not_be_negative():
    u = pscbfm_amd.UpdaterGPUScBFM_AB_Type()
    with pytest.raises(InvalidArgument):
        u.setKernelTiming(-1)
    u.setKernelTiming(8)


def test_device_helpers_on_host():
    """The kernels' 9-site front (at most 4 lattice words and their bit masks,
    plus the old and moved corner sites) equals a direct enumeration with the
    lattice's site function, exhaustively for boxes up to 16384 sites and on
    random corners of larger and non-cubic boxes; the batched kernels' colour
    lookup equals the full colour permutation (tests/cpp/device_helpers_check.cpp)."""
    import subprocess
    exe = Path(__file__).resolve().parent / "cpp" / "device_helpers_check"
    assert exe.exists(), "built by __graft_entry__.build()"
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok ")


def test_host_code_under_sanitizers(tmp_path):
    """The generators, initialize()'s planning and checks, the .bfm reader and
    writer and the oracle's sweeps run clean under AddressSanitizer (leaks
    included) and UndefinedBehaviorSanitizer (tests/cpp/host_sanitizer_check.cpp)."""
    import os
    import subprocess
    exe = Path(__file__).resolve().parent / "cpp" / "host_sanitizer_check"
    assert exe.exists(), "built by __graft_entry__.build()"
    # verify_asan_link_order=0: a library the environment preloads may come
    # before the sanitizer runtime; that is not an error of the code under test
    env = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0, out.stdout + out.stderr[-4000:]
    assert out.stdout.startswith("ok ")


def test_perform_group_must_be_compiled():
    from pscbfm_amd import _capi, systems, updater_from_system
    u = updater_from_system(systems.linear_melt(16, 8, 8, seed=1), seed=1, gpu=None, validate_only=True)
    for g in (2, 4, 8):
        u.setPerformGroup(g)
    for g in (0, 1, 3, 16):
        with pytest.raises(_capi.InvalidArgument):
            u.setPerformGroup(g)


def test_check_block_must_be_compiled():
    # the check kernel is built for 64 ... 1024 monomers per workgroup, one or
    # two per lane
    from pscbfm_amd import _capi, systems, updater_from_system
    u = updater_from_system(systems.linear_melt(16, 8, 8, seed=1), seed=1, gpu=None, validate_only=True)
    for b in (64, 128, 256, 512, 1024):
        u.setCheckBlock(b)
    for b in (0, 32, 96, 192, 2048):
        with pytest.raises(_capi.InvalidArgument):
            u.setCheckBlock(b)
    for i in (1, 2):
        u.setCheckItems(i)
    for i in (0, 3, 4):
        with pytest.raises(_capi.InvalidArgument):
            u.setCheckItems(i)
