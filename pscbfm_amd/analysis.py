"""Observables of linear-chain BFM melts (numpy, host side).

The mean-square displacements follow the reference's analyzer
(src/AnalyzerMsd.tpp:303-421), which numbers them
    g1 = first monomer, g2 = middle monomer, g3 = last monomer,
    g4 = chain centre of mass, g5 = middle monomer relative to the chain COM,
with the displacement of the whole box's centre of mass removed from g1..g4.
Conventional polymer notation calls the middle-monomer MSD g1 and the COM MSD
g3; `msds` returns the reference's g1..g5 and names the conventional pair
explicitly (middle, com).
"""
from __future__ import annotations

import numpy as np

BOND_LENGTHS2 = (4, 5, 6, 9, 10)   # squared lengths of the 108 BFM bond vectors


def chains_view(xyz: np.ndarray, chain_len: int) -> np.ndarray:
    """(n_chains, chain_len, 3) view of unwrapped coordinates with ids running chain by chain."""
    x = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    return x.reshape(-1, chain_len, 3)


def bond_length2_histogram(xyz: np.ndarray, chain_len: int) -> np.ndarray:
    """Fraction of bonds with |b|^2 = 4, 5, 6, 9, 10."""
    c = chains_view(xyz, chain_len)
    b2 = ((c[:, 1:] - c[:, :-1]) ** 2).sum(axis=2).ravel()
    return np.array([(b2 == v).mean() for v in BOND_LENGTHS2])


def re2(xyz: np.ndarray, chain_len: int) -> float:
    c = chains_view(xyz, chain_len)
    return float(((c[:, -1] - c[:, 0]) ** 2).sum(axis=1).mean())


def rg2(xyz: np.ndarray, chain_len: int) -> float:
    c = chains_view(xyz, chain_len)
    com = c.mean(axis=1, keepdims=True)
    return float(((c - com) ** 2).sum(axis=2).mean())


def msds(traj: list, chain_len: int, lag: int) -> dict:
    """MSDs over all time origins of a trajectory (list of xyz frames, equally
    spaced) at a lag of `lag` frames: the reference's g1..g5, plus
    'middle' (= g2) and 'com' (= g4)."""
    frames = [chains_view(f, chain_len) for f in traj]
    mid = chain_len // 2
    acc = np.zeros(5)
    count = 0
    for t0 in range(len(frames) - lag):
        a, b = frames[t0], frames[t0 + lag]
        com_a, com_b = a.mean(axis=1), b.mean(axis=1)
        drift = com_b.mean(axis=0) - com_a.mean(axis=0)       # box COM displacement
        d1 = b[:, 0] - a[:, 0] - drift
        d2 = b[:, mid] - a[:, mid] - drift
        d3 = b[:, -1] - a[:, -1] - drift
        d4 = com_b - com_a - drift
        d5 = (b[:, mid] - com_b) - (a[:, mid] - com_a)
        acc += [float((d ** 2).sum(axis=1).mean()) for d in (d1, d2, d3, d4, d5)]
        count += 1
    g = acc / max(count, 1)
    return {"g1": g[0], "g2": g[1], "g3": g[2], "g4": g[3], "g5": g[4], "middle": g[1], "com": g[3]}


def block_mean(values, n_blocks: int = 8) -> tuple[float, float]:
    """Mean and standard error from block averages of a time series."""
    v = np.asarray(values, dtype=np.float64)
    k = len(v) // n_blocks
    blocks = v[: k * n_blocks].reshape(n_blocks, k).mean(axis=1)
    return float(blocks.mean()), float(blocks.std(ddof=1) / np.sqrt(n_blocks))
