"""The entry point that allocates nothing, bfm_run_simulation_on_gpu_noalloc,
against the existing bfm_run_simulation_on_gpu: the same final positions,
counters and age, bit for bit, on calls one after another that share the
device scratch -- a larger system, then a smaller one, then the larger one
again (re-sorted on the device in between)."""
import numpy as np
import pytest

from pscbfm_amd import InvalidArgument, UpdaterGPUScBFM_AB_Type, _capi, systems, updater_from_system


def test_noalloc_rejects_systems_beyond_its_scratch():
    cap = _capi.load().bfm_noalloc_max_monomers()
    assert cap == (32 << 20) // 12
    u = UpdaterGPUScBFM_AB_Type()
    u.setNrOfAllMonomers(cap + 1)
    with pytest.raises(InvalidArgument, match="at most 2796202 monomers"):
        u.runSimulationOnGPUNoAlloc(1)


@pytest.mark.gpu
def test_noalloc_matches_run_simulation_on_gpu(gpu_ready):
    big = systems.linear_melt(64, 512, 32, seed=0x5EED0000000000A1)      # 16,384 monomers, 2 colours
    small = systems.cluster_gel(32, 0.8, 0.5, seed=0xA2)               # 1,680 monomers, 4 colours
    cases = [(big, 0x5EED0000000000A3), (small, 0x5EED0000000000A4)]
    pairs = []
    for s, seed in cases:
        mine = updater_from_system(s, seed=seed, gpu=0, steps_between_sortings=30)
        ref = updater_from_system(s, seed=seed, gpu=0, steps_between_sortings=30)
        pairs.append((s, mine, ref))
    for k, (case, mcs) in enumerate([(0, 40), (1, 25), (0, 35)]):
        s, mine, ref = pairs[case]
        n = len(s["xyz"])
        print(f"case: call {k}, system {case} ({n} monomers), {mcs} MCS", flush=True)
        mine.runSimulationOnGPUNoAlloc(mcs)
        ref.runSimulationOnGPU(mcs)
        assert np.array_equal(mine.getMonomerPositions(n), ref.getMonomerPositions(n)), f"call {k} differs"
        assert mine.getAge() == ref.getAge()
        a, b = mine.getStats(), ref.getStats()
        assert (a["accepted"], a["proposals"], a["sorts"]) == (b["accepted"], b["proposals"], b["sorts"])
    assert pairs[0][1].getStats()["sorts"] >= 2
    for _, mine, ref in pairs:
        mine.cleanup()
        ref.cleanup()
