"""bench.py's command-line contract and workload table (no GPU needed)."""
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("name,monomers,bonds", [
    ("small_melt", 16384, 512 * 31),          # 512 chains of 32
    ("network", 1064960, 32768 * 33),         # 32,768 strands of 32: 31 inner bonds + 2 to crosslinkers
])
def test_workload_sizes(name, monomers, bonds):
    s = bench.build_workload(bench.WORKLOADS[name])
    assert len(s["xyz"]) == monomers
    assert len(s["bonds"]) == bonds
    cfg = bench.workload_config(name, bench.WORKLOADS[name])
    assert cfg["workload"] == name and cfg["box"] == s["box"][0]


def test_large_melt_is_the_default_workload():
    w = bench.WORKLOADS["large_melt"]
    assert (w["box"], w["n_chains"], w["chain_len"]) == (512, 16384, 512)
