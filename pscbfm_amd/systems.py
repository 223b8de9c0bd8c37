"""Seeded synthetic systems (linear melts) as plain dicts:
{"box": (L, L, L), "periodic": (1, 1, 1), "xyz": int32 [n, 3],
 "bonds": uint32 [m, 2] undirected, "forbidden": uint8 [512]}.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import _capi


def standard_forbidden() -> np.ndarray:
    """512-entry table of the standard 108-vector BFM bond set (1 = forbidden),
    index (dx&7)|(dy&7)<<3|(dz&7)<<6 as UpdaterGPUScBFM_AB_Type.cu:886-888."""
    f = np.ones(512, dtype=np.uint8)
    allowed = {(0, 0, 2), (0, 1, 2), (1, 1, 2), (1, 2, 2), (0, 0, 3), (0, 1, 3)}
    for d in itertools.product(range(-3, 4), repeat=3):
        if tuple(sorted(abs(v) for v in d)) in allowed:
            f[(d[0] & 7) | ((d[1] & 7) << 3) | ((d[2] & 7) << 6)] = 0
    return f


def linear_melt(box: int, n_chains: int, chain_len: int, seed: int) -> dict:
    """Linear melt grown by bfm_gen_linear_melt (include/pscbfm_amd_systems.h)."""
    lib = _capi.load()
    n = n_chains * chain_len
    xyz = np.empty((n, 3), dtype=np.int32)
    bonds = np.empty((n_chains * max(chain_len - 1, 0), 2), dtype=np.uint32)
    _capi.check(lib.bfm_gen_linear_melt(box, n_chains, chain_len, seed & 0xFFFFFFFFFFFFFFFF,
                                        xyz.ctypes.data, bonds.ctypes.data if bonds.size else None))
    return {"box": (box, box, box), "periodic": (1, 1, 1), "xyz": xyz, "bonds": bonds,
            "forbidden": standard_forbidden(), "chain_len": chain_len}


def diamond_network(box: int, cells: tuple, strand_len: int, seed: int) -> dict:
    """End-linked diamond network grown by bfm_gen_diamond_network
    (include/pscbfm_amd_systems.h)."""
    import ctypes
    lib = _capi.load()
    nm, nb = ctypes.c_uint32(), ctypes.c_uint32()
    cx, cy, cz = cells
    _capi.check(lib.bfm_gen_diamond_network_size(box, cx, cy, cz, strand_len, ctypes.byref(nm), ctypes.byref(nb)))
    xyz = np.empty((nm.value, 3), dtype=np.int32)
    bonds = np.empty((nb.value, 2), dtype=np.uint32)
    _capi.check(lib.bfm_gen_diamond_network(box, cx, cy, cz, strand_len, seed & 0xFFFFFFFFFFFFFFFF,
                                            xyz.ctypes.data, bonds.ctypes.data))
    return {"box": (box, box, box), "periodic": (1, 1, 1), "xyz": xyz, "bonds": bonds,
            "forbidden": standard_forbidden(), "strand_len": strand_len,
            "crosslinkers": 8 * cx * cy * cz}


def melt_volume_fraction(system: dict) -> float:
    L = system["box"]
    return 8.0 * len(system["xyz"]) / float(L[0] * L[1] * L[2])


# Corners of a tetrahedral cluster in a 4^3 cell: every pair of corners is an
# allowed bond ((2,0,0), (1,2,0), (1,1,2), (-1,2,0), (-1,1,2), (0,-1,2) up to
# sign) and no two cubes overlap, so the four monomers are mutually bonded
# (a K4 bond graph, which needs 4 colours).
K4_CORNERS = np.array([(0, 0, 0), (2, 0, 0), (1, 2, 0), (1, 1, 2)], dtype=np.int32)
# A chain of four monomers around a square in the same cell (bonds (2,0,0)).
SQUARE_CHAIN = np.array([(0, 0, 0), (2, 0, 0), (2, 2, 0), (0, 2, 0)], dtype=np.int32)


def cluster_gel(box: int, fill: float, k4_fraction: float, seed: int) -> dict:
    """Tetrahedral clusters and short chains in a periodic cubic box.

    The box is cut into cells of edge 4; each cell holds, with probability
    `fill`, four monomers: with probability `k4_fraction` a K4 cluster (six
    bonds, four colours), else a chain of four (three bonds).  Each unit is
    reflected and its axes permuted at random inside its cell, so cubes never
    leave their cell.  Volume fraction = fill / 2.  Ids run cell by cell."""
    if box % 4 or box < 8:
        raise ValueError("box must be a multiple of 4 and at least 8")
    rng = np.random.default_rng(seed)
    g = box // 4
    xyz, bonds = [], []
    for cz in range(g):
        for cy in range(g):
            for cx in range(g):
                if rng.random() >= fill:
                    continue
                k4 = rng.random() < k4_fraction
                unit = (K4_CORNERS if k4 else SQUARE_CHAIN).copy()
                unit = unit[:, rng.permutation(3)]
                flip = rng.random(3) < 0.5
                unit[:, flip] = 2 - unit[:, flip]
                base = len(xyz)
                xyz.extend((unit + 4 * np.array([cx, cy, cz], dtype=np.int32)).tolist())
                pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)] if k4 else [(0, 1), (1, 2), (2, 3)]
                bonds.extend((base + a, base + b) for a, b in pairs)
    return {"box": (box, box, box), "periodic": (1, 1, 1), "xyz": np.array(xyz, dtype=np.int32).reshape(-1, 3),
            "bonds": np.array(bonds, dtype=np.uint32).reshape(-1, 2), "forbidden": standard_forbidden()}
