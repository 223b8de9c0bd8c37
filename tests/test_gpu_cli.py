"""The command-line simulator (pscbfm_amd/bin/pscbfm_amd_sim, the reference's
SimulatorCUDAGPUScBFM_AB_Type flags -i -o -m -s -g -e -r): a BFM file in, the
appended conformations out, the last one bit-exact against the oracle."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc
from pscbfm_amd import UpdaterGPUScBFM_AB_Type, systems, updater_from_system

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parent.parent / "pscbfm_amd" / "bin" / "pscbfm_amd_sim"


@pytest.mark.parametrize("rng", [orc.RNG_INT_HASH, orc.RNG_PHILOX])
def test_cli_continues_a_bfm_file(gpu_ready, tmp_path, rng):
    assert EXE.exists(), "built by __graft_entry__.build()"
    s = systems.linear_melt(32, 64, 16, seed=11)
    n = len(s["xyz"])
    init = tmp_path / "init.bfm"
    updater_from_system(s, seed=0, validate_only=True, age=40).writeBfmFile(str(init))
    seeds = tmp_path / "run.seeds"
    seeds.write_text("305419896\n2596069104\n7\n")
    seed = (305419896 << 32) | 2596069104
    out = tmp_path / "out.bfm"
    cmd = [str(EXE), "-i", str(init), "-o", str(out), "-m", "250", "-s", "100", "-g", "0", "-e", str(seeds),
           "-r", str(rng)]
    print("case: " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    text = out.read_bytes().decode("latin-1")
    # the initial state, then one conformation per completed save interval
    assert [line for line in text.splitlines() if line.startswith("!mcs=")] == ["!mcs=40", "!mcs=140", "!mcs=240"]
    v = UpdaterGPUScBFM_AB_Type()
    v.readBfmFile(str(out))
    assert v.getAge() == 240
    want, _, _ = orc.sweep(orc.System(s, seed, rng=rng), s["xyz"], 40, 200, 1)
    assert np.array_equal(v.getMonomerPositions(n), want)
