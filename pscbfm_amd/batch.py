"""Batched and served forms over the C ABI (include/pscbfm_amd.h, bfm_batch_* and
bfm_server_*): many independent systems per launch, and continuous batching of
simulation jobs admitted by free HBM."""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_int, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

from . import _capi
from .updater import UpdaterGPUScBFM_AB_Type

JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED, JOB_CANCELLED = range(5)


def member_bytes(u: UpdaterGPUScBFM_AB_Type) -> int:
    v = c_uint64()
    _capi.check(_capi.load().bfm_member_bytes(u._h, byref(v)))
    return v.value


class Batch:
    """Independent systems swept together; each member's trajectory equals the
    plain updater's for the same system, seed and age."""

    def __init__(self, gpu: int = 0):
        self._lib = _capi.load()
        h = c_void_p()
        _capi.check(self._lib.bfm_batch_create(byref(h), gpu))
        self._h = h
        self._n = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.bfm_batch_destroy(h)
            self._h = None

    def add(self, u: UpdaterGPUScBFM_AB_Type, n_monomers: int) -> int:
        m = c_uint32()
        _capi.check(self._lib.bfm_batch_add(self._h, u._h, byref(m)))
        self._n[m.value] = n_monomers
        return m.value

    def run(self, n_mcs: int, chunk: int = 100) -> None:
        _capi.check(self._lib.bfm_batch_run(self._h, n_mcs, chunk))

    def run_members(self, members, n_mcs: int, chunk: int = 100) -> None:
        ids = np.ascontiguousarray(members, dtype=np.uint32)
        _capi.check(self._lib.bfm_batch_run_members(self._h, ids.ctypes.data, ids.size, n_mcs, chunk))

    def positions(self, member: int) -> np.ndarray:
        n = self._n[member]
        out = np.empty((n, 3), dtype=np.int32)
        _capi.check(self._lib.bfm_batch_get_positions(self._h, member, out.ctypes.data, n))
        return out

    def set_perform_group(self, group: int) -> None:
        _capi.check(self._lib.bfm_batch_set_perform_group(self._h, int(group)))

    def autotune(self, probe_mcs: int, chunk: int = 100) -> dict:
        """Times probe_mcs MCS of every member under each launch configuration
        and keeps the fastest (members advance by 7 * probe_mcs MCS: one untimed
        probe, then one per configuration)."""
        g, x = c_uint32(), ctypes.c_int()
        _capi.check(self._lib.bfm_batch_autotune(self._h, int(probe_mcs), int(chunk), byref(g), byref(x)))
        return {"perform_group": g.value, "xcd_ranges": bool(x.value)}

    def observe(self, member: int) -> dict:
        """Device observables of a member's resident state (as
        UpdaterGPUScBFM_AB_Type.observe)."""
        from .updater import observables_dict
        o = _capi.BfmObservables()
        _capi.check(self._lib.bfm_batch_observe(self._h, member, ctypes.byref(o)))
        return observables_dict(o)

    def age(self, member: int) -> int:
        v = c_int64()
        _capi.check(self._lib.bfm_batch_get_age(self._h, member, byref(v)))
        return v.value

    def stats(self, member: int) -> dict:
        st = _capi.BfmStats()
        _capi.check(self._lib.bfm_batch_get_stats(self._h, member, byref(st)))
        return {name: getattr(st, name) for name, _ in _capi.BfmStats._fields_}

    def remove(self, member: int) -> None:
        _capi.check(self._lib.bfm_batch_remove(self._h, member))

    def set_graphs(self, on: bool) -> None:
        _capi.check(self._lib.bfm_batch_set_graphs(self._h, int(bool(on))))

    def set_working_set_bytes(self, nbytes: int) -> None:
        """Working-set budget of run() (0: never split the members into groups)."""
        _capi.check(self._lib.bfm_batch_set_working_set_bytes(self._h, int(nbytes)))

    def working_set_groups(self) -> int:
        """Number of member groups run() would use now."""
        g = c_uint32()
        _capi.check(self._lib.bfm_batch_get_working_set_groups(self._h, byref(g)))
        return g.value

    def set_xcd_ranges(self, on: bool) -> None:
        _capi.check(self._lib.bfm_batch_set_xcd_ranges(self._h, int(bool(on))))

    def synchronize(self) -> None:
        _capi.check(self._lib.bfm_batch_synchronize(self._h))


class Server:
    """Continuous batching: submit jobs at any time, call step() to run one
    chunk on every admitted job; finished jobs free their device memory."""

    def __init__(self, gpu: int = 0, hbm_budget: int = 0, chunk: int = 100):
        self._lib = _capi.load()
        h = c_void_p()
        _capi.check(self._lib.bfm_server_create(byref(h), gpu, hbm_budget, chunk))
        self._h = h
        self._n = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.bfm_server_destroy(h)
            self._h = None

    def submit(self, u: UpdaterGPUScBFM_AB_Type, n_mcs: int, n_monomers: int, take: bool = False) -> int:
        """Queues the updater's system for n_mcs MCS.  take=True hands a
        not-initialized updater's validated system over without a copy (the
        updater is left as after cleanup)."""
        j = c_uint64()
        fn = self._lib.bfm_server_submit_take if take else self._lib.bfm_server_submit
        _capi.check(fn(self._h, u._h, n_mcs, byref(j)))
        self._n[j.value] = n_monomers
        return j.value

    def step(self) -> int:
        p = c_uint32()
        _capi.check(self._lib.bfm_server_step(self._h, byref(p)))
        return p.value

    def state(self, job: int) -> tuple[int, int]:
        st, done = c_int(), c_uint32()
        _capi.check(self._lib.bfm_server_job_state(self._h, job, byref(st), byref(done)))
        return st.value, done.value

    def times(self, job: int) -> tuple[float, float, float]:
        a, b, c = c_double(), c_double(), c_double()
        _capi.check(self._lib.bfm_server_job_times(self._h, job, byref(a), byref(b), byref(c)))
        return a.value, b.value, c.value

    def result(self, job: int) -> np.ndarray:
        n = self._n[job]
        out = np.empty((n, 3), dtype=np.int32)
        _capi.check(self._lib.bfm_server_job_result(self._h, job, out.ctypes.data, n))
        return out

    def error(self, job: int) -> str:
        buf = ctypes.create_string_buffer(1024)
        _capi.check(self._lib.bfm_server_job_error(self._h, job, buf, 1024))
        return buf.value.decode()

    def cancel(self, job: int) -> None:
        _capi.check(self._lib.bfm_server_cancel(self._h, job))

    def release(self, job: int) -> None:
        _capi.check(self._lib.bfm_server_release(self._h, job))
        self._n.pop(job, None)

    def info(self) -> dict:
        q, r = c_uint32(), c_uint32()
        used, budget, att = c_uint64(), c_uint64(), c_uint64()
        _capi.check(self._lib.bfm_server_info(self._h, byref(q), byref(r), byref(used), byref(budget), byref(att)))
        return {"queued": q.value, "running": r.value, "bytes_in_use": used.value, "budget": budget.value,
                "attempts": att.value}

    def host_memory(self) -> dict:
        """Host bytes held for jobs, jobs held, batch members still tracked."""
        b, j, m = c_uint64(), c_uint32(), c_uint32()
        _capi.check(self._lib.bfm_server_host_bytes(self._h, byref(b), byref(j), byref(m)))
        return {"host_bytes": b.value, "jobs": j.value, "members": m.value}

    def graph_stats(self) -> dict:
        """Chunk graphs launched, updated in place, instantiated."""
        v = [c_uint64() for _ in range(3)]
        _capi.check(self._lib.bfm_server_graph_stats(self._h, *[byref(x) for x in v]))
        return dict(zip(("launches", "updates", "instantiations"), (x.value for x in v)))
