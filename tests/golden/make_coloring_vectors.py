"""Writes coloring_vectors.json: colourings that the reference's own graph
colouring (src/pscbfm/graphColoring.tpp, the call of
UpdaterGPUScBFM_AB_Type.cu:1561-1566), built from its source into
oracle/_ref/coloring_ref by oracle/ref.mk, gives for a set of bond graphs:
paths and rings with shuffled ids, complete graphs, stars, trees, random
graphs of at most 8 partners per monomer with shuffled link order, tetrahedral
clusters mixed with chains (the uniform rebalance at work) and a small diamond
network.  Every graph is stored as per-monomer link lists in call order (both
directions of every bond, as the LeMonADE adapter sets them), with and
without the uniform rebalance.

Run in a tree where oracle/_ref/coloring_ref exists:
    python tests/golden/make_coloring_vectors.py
"""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
EXE = ROOT / "oracle" / "_ref" / "coloring_ref"
sys.path.insert(0, str(ROOT))


def lists_from_bonds(n, bonds, rng=None):
    """Per-monomer lists in call order: for each bond (a, b) in turn, a -> b
    then b -> a; with rng, each monomer's list is shuffled afterwards."""
    lists = [[] for _ in range(n)]
    for a, b in bonds:
        lists[a].append(int(b))
        lists[b].append(int(a))
    if rng is not None:
        for l in lists:
            rng.shuffle(l)
    return lists


def relabel(n, bonds, rng):
    perm = rng.permutation(n)
    return [(int(perm[a]), int(perm[b])) for a, b in bonds]


def graphs():
    rng = np.random.default_rng(20261204)
    out = []
    path = [(i, i + 1) for i in range(11)]
    out.append(("path12", 12, lists_from_bonds(12, path)))
    out.append(("path12_relabelled", 12, lists_from_bonds(12, relabel(12, path, rng))))
    for k in (8, 9):
        ring = [(i, (i + 1) % k) for i in range(k)]
        out.append((f"ring{k}", k, lists_from_bonds(k, relabel(k, ring, rng))))
    for k in (4, 5, 9):
        comp = [(a, b) for a in range(k) for b in range(a + 1, k)]
        out.append((f"complete{k}", k, lists_from_bonds(k, comp, rng)))
    out.append(("star9", 9, lists_from_bonds(9, [(0, i) for i in range(1, 9)])))
    tree = [(i, int(rng.integers(0, i))) for i in range(1, 40)]
    deg = np.zeros(40, dtype=int)
    tree_ok = []
    for a, b in tree:
        if deg[a] < 8 and deg[b] < 8:
            tree_ok.append((a, b))
            deg[a] += 1
            deg[b] += 1
    out.append(("tree40", 40, lists_from_bonds(40, tree_ok, rng)))
    for g in range(6):
        n = int(rng.integers(20, 90))
        deg = np.zeros(n, dtype=int)
        seen, bonds = set(), []
        for _ in range(int(rng.integers(n, 3 * n))):
            a, b = (int(v) for v in rng.integers(0, n, 2))
            if a == b or (min(a, b), max(a, b)) in seen or deg[a] >= 8 or deg[b] >= 8:
                continue
            seen.add((min(a, b), max(a, b)))
            bonds.append((a, b))
            deg[a] += 1
            deg[b] += 1
        out.append((f"random{g}_n{n}", n, lists_from_bonds(n, bonds, rng)))
    from pscbfm_amd import systems
    for seed, (fill, k4) in enumerate([(0.7, 0.5), (0.9, 0.2), (0.5, 0.8)]):
        s = systems.cluster_gel(16, fill, k4, seed=11 + seed)
        n = len(s["xyz"])
        out.append((f"cluster_gel_{fill}_{k4}", n, lists_from_bonds(n, s["bonds"].tolist())))
    d = systems.diamond_network(32, (1, 1, 4), 8, seed=5)
    n = len(d["xyz"])
    out.append(("diamond_32_1x1x4_strand8", n, lists_from_bonds(n, d["bonds"].tolist())))
    m = systems.linear_melt(16, 8, 7, seed=9)             # odd chains: unequal colour counts
    n = len(m["xyz"])
    out.append(("melt_16_8x7", n, lists_from_bonds(n, m["bonds"].tolist())))
    return out


def main() -> None:
    vec = []
    for name, n, lists in graphs():
        for uniform in (1, 0):
            text = f"{uniform} {n}\n" + "".join(f"{len(l)} " + " ".join(map(str, l)) + "\n" for l in lists)
            res = subprocess.run([str(EXE)], input=text, capture_output=True, text=True, check=True).stdout.split()
            vec.append({"name": name, "uniform": uniform, "lists": lists, "n_colors": int(res[0]),
                        "colors": [int(c) for c in res[1:]]})
    (HERE / "coloring_vectors.json").write_text(json.dumps(
        {"source": "src/pscbfm/graphColoring.tpp:103-439 via UpdaterGPUScBFM_AB_Type.cu:1561-1566",
         "graphs": vec}) + "\n")


if __name__ == "__main__":
    main()
