"""The header-only LeMonADE adapter (include/pscbfm_amd_lemonade.hpp) driven by
a C++ host with a stand-in Ingredients: what a LeMonADE TaskManager does with
GPUScBFM_AB_Type (src/pscbfm/GPUScBFM_AB_Type.h:120-242)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc
from pscbfm_amd import device_count, systems

ROOT = Path(__file__).resolve().parents[1]
DEMO = ROOT / "tests" / "cpp" / "lemonade_adapter_demo"


def write_input(path, s, seed, steps, calls, age):
    xyz, bonds = s["xyz"], s["bonds"]
    with open(path, "w") as f:
        f.write(f"{s['box'][0]} {len(xyz)} {len(bonds)} {seed} {steps} {calls} {age}\n")
        np.savetxt(f, xyz, fmt="%d")
        np.savetxt(f, bonds, fmt="%d")


def run_demo(tmp_path, s, seed, steps, calls, age, rng=None):
    if not DEMO.exists():
        pytest.fail(f"{DEMO} not built: run __graft_entry__.build()")
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    write_input(inp, s, seed, steps, calls, age)
    extra = [] if rng is None else [str(rng)]
    r = subprocess.run([str(DEMO), str(inp), str(out), *extra], capture_output=True, text=True, timeout=120)
    return r, out


def test_adapter_reports_missing_device(tmp_path):
    if device_count() > 0:
        pytest.skip("a HIP device is visible")
    s = systems.linear_melt(16, 8, 16, seed=1)
    r, _ = run_demo(tmp_path, s, 5, 10, 1, 0)
    assert r.returncode == 3
    assert "no HIP device" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("rng", [None, orc.RNG_INT_HASH, orc.RNG_PHILOX])
def test_adapter_matches_oracle(tmp_path, gpu_ready, rng):
    """initialize / execute x4 / cleanup through the adapter, with the default
    generator and each one selected by setRng, bit-exact against the oracle."""
    s = systems.linear_melt(32, 64, 16, seed=2)
    seed, steps, calls, age = 0xABCDEF, 25, 4, 500
    r, out = run_demo(tmp_path, s, seed, steps, calls, age, rng)
    assert r.returncode == 0, r.stderr
    lines = out.read_text().split("\n")
    assert int(lines[0]) == age + steps * calls
    got = np.loadtxt(lines[1:], dtype=np.int32).reshape(-1, 3)
    osys = orc.System(s, seed, rng=orc.RNG_INT_HASH if rng is None else rng)
    want, _, _ = orc.sweep(osys, s["xyz"], age, steps * calls, 0)
    assert np.array_equal(got, want)


def test_adapter_rejects_unknown_rng(tmp_path):
    """setRng with a generator this build does not have fails in initialize
    with the updater's invalid_argument (exit status 3 of the demo)."""
    s = systems.linear_melt(16, 8, 16, seed=1)
    r, _ = run_demo(tmp_path, s, 5, 10, 1, 0, rng=2)
    assert r.returncode == 3
    assert "not available" in r.stderr or "no HIP device" in r.stderr
