"""The served form at its own job distribution: jobs drawn from the mixed-batch
size distribution (box edges 32, 64, 128; chains of 16 to 128; volume fraction
0.3 to 0.5), each asking for 100 to 2000 MCS, arriving while others run, under
an HBM budget that admits them in waves.  Every job's result is bit-identical
to the same system run alone with the plain updater, and three of them
(the largest among them) to the oracle's replay."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from pscbfm_amd import Server, member_bytes, updater_from_system, workloads
from pscbfm_amd.batch import JOB_CANCELLED, JOB_DONE

pytestmark = pytest.mark.gpu


def job_stream_sample():
    jobs = workloads.job_stream(24, 1.0, seed=0xC5C5)
    # one of the largest jobs the distribution holds: box 128, chains of 128
    jobs[4].update(box=128, chain_len=128, phi=0.45, chains=int(0.45 * 128 ** 3 / (8 * 128)), mcs=1000)
    return jobs


def test_served_job_stream_matches_plain_and_oracle(gpu_ready):
    jobs = job_stream_sample()
    built = [workloads.build(j) for j in jobs]
    sizes = [(j["box"], j["chain_len"]) for j in jobs]
    assert (128, 128) in sizes and min(j["mcs"] for j in jobs) >= 100 and max(j["mcs"] for j in jobs) <= 2000
    print(f"case: served stream of {len(jobs)} jobs, {sum(len(x['xyz']) for x in built)} monomers, "
          f"MCS {min(j['mcs'] for j in jobs)}..{max(j['mcs'] for j in jobs)}", flush=True)
    upd = [updater_from_system(x, seed=j["seed"], validate_only=True) for x, j in zip(built, jobs)]
    budget = int(3 * max(member_bytes(u) for u in upd))
    srv = Server(0, hbm_budget=budget, chunk=100)
    assert srv.host_memory() == {"host_bytes": 0, "jobs": 0, "members": 0}
    # arrivals replayed in order, two per scheduling round on average
    order = np.argsort([j["arrival"] for j in jobs], kind="stable")
    round_of = {int(k): int(r) // 2 for r, k in enumerate(order)}
    ids, results, running_seen = {}, {}, []
    rnd = 0
    while True:
        for k in [k for k, r in round_of.items() if r == rnd]:
            ids[k] = srv.submit(upd[k], jobs[k]["mcs"], len(built[k]["xyz"]))
        pending = srv.step()
        info = srv.info()
        assert info["bytes_in_use"] <= budget
        running_seen.append(info["running"])
        for k, jid in ids.items():
            if k not in results and srv.state(jid)[0] == JOB_DONE:
                results[k] = srv.result(jid)      # taken as a client would, between rounds
        rnd += 1
        assert rnd < 2000
        if pending == 0 and len(ids) == len(jobs):
            break
    assert max(running_seen) < len(jobs), "the budget should admit the jobs in waves"
    assert len(results) == len(jobs)
    for k, jid in ids.items():
        assert srv.state(jid) == (JOB_DONE, jobs[k]["mcs"])
    # every job against the same system alone
    bad = []
    for k, (x, j) in enumerate(zip(built, jobs)):
        u = updater_from_system(x, seed=j["seed"], gpu=0)
        u.runSimulationOnGPU(j["mcs"])
        if not np.array_equal(u.getMonomerPositions(len(x["xyz"])), results[k]):
            bad.append(k)
        u.cleanup()
    assert not bad, f"jobs {bad} differ from their plain runs"
    # three against the oracle: the box-128 job and the two smallest others
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    small = sorted((len(x["xyz"]) * j["mcs"], k) for k, (x, j) in enumerate(zip(built, jobs)) if k != 4)
    for k in [4, small[0][1], small[1][1]]:
        x, j = built[k], jobs[k]
        want, _, _ = orc.sweep_parallel(orc.System(x, j["seed"]), x["xyz"], 0, j["mcs"], min(threads, 16))
        assert np.array_equal(results[k], want), f"job {k} differs from the oracle"
    for jid in ids.values():
        srv.release(jid)
    assert srv.host_memory() == {"host_bytes": 0, "jobs": 0, "members": 0}


def test_server_host_memory_returns_to_baseline(gpu_ready):
    """Jobs released with or without their result taken, and a job cancelled
    while running, leave no host arrays behind."""
    specs = workloads.mixed_batch(6, seed=0x4E7)
    built = [workloads.build(sp) for sp in specs]
    srv = Server(0, hbm_budget=0, chunk=50)
    jobs = [srv.submit(updater_from_system(x, seed=sp["seed"], validate_only=True), 100 + 50 * (k % 3), len(x["xyz"]))
            for k, (x, sp) in enumerate(zip(built, specs))]
    held = srv.host_memory()
    assert held["jobs"] == 6 and held["host_bytes"] > 0
    srv.step()
    srv.cancel(jobs[5])
    assert srv.state(jobs[5])[0] == JOB_CANCELLED
    while srv.step() > 0:
        pass
    for k in (0, 1, 2):
        srv.result(jobs[k])
    held = srv.host_memory()
    # retired jobs keep only their results (pinned), not their systems
    assert held["members"] == 0 and held["jobs"] == 6
    assert held["host_bytes"] >= sum(12 * len(x["xyz"]) for x in built[:5])
    for j in jobs:
        srv.release(j)
    assert srv.host_memory() == {"host_bytes": 0, "jobs": 0, "members": 0}


def test_submit_take_hands_the_system_over(gpu_ready):
    """bfm_server_submit_take moves a validated updater's system into the job
    (the updater is left empty) and gives the same result as a copied one."""
    from pscbfm_amd import systems
    s = systems.linear_melt(32, 100, 16, seed=61)
    n = len(s["xyz"])
    srv = Server(0, hbm_budget=0, chunk=50)
    u_copy = updater_from_system(s, seed=0x5EED0000000000B1, age=3, validate_only=True)
    u_take = updater_from_system(s, seed=0x5EED0000000000B1, age=3, validate_only=True)
    a = srv.submit(u_copy, 150, n)
    b = srv.submit(u_take, 150, n, take=True)
    assert u_take.getAge() == 3
    with pytest.raises(Exception):
        u_take.getColors(n)                  # nothing left in the updater
    u_copy.getColors(n)                      # a copied updater keeps its system
    while srv.step() > 0:
        pass
    assert np.array_equal(srv.result(a), srv.result(b))
    want, _, _ = orc.sweep(orc.System(s, 0x5EED0000000000B1), s["xyz"], 3, 150, 0)
    assert np.array_equal(srv.result(b), want)


def test_served_jobs_resort_on_their_schedule(gpu_ready):
    """A served job whose updater re-sorts every 40 MCS re-sorts inside its
    run (chunks of 100 split at ages 40, 80, 120) and ends where the oracle
    ends; a job beside it that never reaches a sort age is not given a
    re-sort workspace."""
    from pscbfm_amd import systems
    a = systems.linear_melt(32, 100, 16, seed=62)
    b = systems.cluster_gel(32, 0.8, 0.5, seed=63)
    srv = Server(0, hbm_budget=0, chunk=100)
    ja = srv.submit(updater_from_system(a, seed=0x5EED0000000000B2, validate_only=True, steps_between_sortings=40),
                    150, len(a["xyz"]), take=True)
    jb = srv.submit(updater_from_system(b, seed=0x5EED0000000000B3, age=7, validate_only=True), 150, len(b["xyz"]))
    while srv.step() > 0:
        pass
    for j, s, seed, age in ((ja, a, 0x5EED0000000000B2, 0), (jb, b, 0x5EED0000000000B3, 7)):
        want, _, _ = orc.sweep(orc.System(s, seed), s["xyz"], age, 150, 0)
        assert np.array_equal(srv.result(j), want)
