"""Synthetic workloads of the batched and served forms.

mixed_batch: independent linear melts with box edge from {32, 64, 128}, chain
length from {16, 32, 64, 128}, volume fraction uniform in [0.3, 0.5], chains =
floor(phi * L^3 / (8 N)), each with its own seed.

job_stream: Poisson arrivals of jobs whose systems follow the same
distribution, each asking for 100..2000 MCS in steps of 100.

uniform_batch: n melts of one box edge, chain length and volume fraction
(chains as above), each with its own seed: the compositions that separate
member size from density in the batched sweep's cost.
"""
from __future__ import annotations

import numpy as np

from . import systems

BOXES = (32, 64, 128)
CHAIN_LENGTHS = (16, 32, 64, 128)


def _spec(L: int, N: int, phi: float, seed: int) -> dict:
    chains = int(phi * L ** 3 / (8 * N))
    chains = max(1, min(chains, L ** 3 // (16 * N)))
    return {"box": L, "chain_len": N, "phi": phi, "chains": chains, "seed": seed}


def mixed_system_spec(rng: np.random.Generator) -> dict:
    L = int(rng.choice(BOXES))
    N = int(rng.choice(CHAIN_LENGTHS))
    phi = float(rng.uniform(0.3, 0.5))
    return _spec(L, N, phi, int(rng.integers(1, 2 ** 63)))


def uniform_batch(n_systems: int, box: int, chain_len: int, phi: float, seed: int) -> list[dict]:
    rng = np.random.default_rng(seed)
    return [_spec(box, chain_len, phi, int(rng.integers(1, 2 ** 63))) for _ in range(n_systems)]


def build(spec: dict) -> dict:
    return systems.linear_melt(spec["box"], spec["chains"], spec["chain_len"], spec["seed"])


def mixed_batch(n_systems: int, seed: int) -> list[dict]:
    rng = np.random.default_rng(seed)
    return [mixed_system_spec(rng) for _ in range(n_systems)]


def job_stream(n_jobs: int, rate_hz: float, seed: int) -> list[dict]:
    """Jobs with arrival times (s), system specs and requested MCS."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for _ in range(n_jobs):
        t += float(rng.exponential(1.0 / rate_hz))
        spec = mixed_system_spec(rng)
        spec["arrival"] = t
        spec["mcs"] = int(rng.integers(1, 21)) * 100
        jobs.append(spec)
    return jobs
