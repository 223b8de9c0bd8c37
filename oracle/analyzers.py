"""CPU restatement of the reference's analyzers, for checking the device ones.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
package (pscbfm_amd) never imports it.

Follows src/AnalyzerMsd.tpp line by line in plain Python loops:
  * OnlineStatistics (:18-95): value-by-value moments, getCentralMoment with
    choose(number of stored moments, i) (:79);
  * getLinearPolymerSystemInfo (:186-288): stack walk, ends in pop order,
    middle monomer floor(n/2) links from the first end;
  * execute (:303-421): chain COMs summed monomer by monomer in walk order,
    box COM as the mean of chain COMs, subsequences every
    deltaTRealization, evaluation every deltaTEvaluate, box-COM drift removed
    from g1..g4 when there is more than one chain;
  * cleanup (:423-468): the output file.
Re^2, Rg^2 and the bond-vector histogram have no analyzer in the reference;
they are restated from their textbook definitions:
  Re^2 = |r_last - r_first|^2, Rg^2 = (1/n) sum |r_i - r_com|^2,
  bond index = (dx&7) | (dy&7)<<3 | (dz&7)<<6 (UpdaterGPUScBFM_AB_Type.cu:871-889).
"""
from __future__ import annotations

import math

import numpy as np

ERR_BRANCHED = ("Some monomers have more than two bonds, indicating that this polymer system does not only "
                "consist of linear chains. This Analyzer module only works with linear chains!")
ERR_ENDS = ("Found not exactly 2 end monomers! The polymer system might be weird, e.g. end monomers with have "
            "loops or ring polymers.")
ERR_MIDDLE = "Unexpected end monomer encountered when iterating over the polymer to find the middle monomer."


def link_lists(n: int, bonds) -> list[list[int]]:
    """Links per monomer in the order the bonds list them (each partner once):
    what the LeMonADE adapter's setConnectivity calls produce."""
    links: list[list[int]] = [[] for _ in range(n)]
    for a, b in np.asarray(bonds, dtype=np.int64).reshape(-1, 2):
        a, b = int(a), int(b)
        if b not in links[a]:
            links[a].append(b)
        if a not in links[b]:
            links[b].append(a)
    return links


def _step(links, cur, prev):
    """AnalyzerMsd.tpp:331-341: link 0 unless we came from there, else link 1."""
    l = links[cur]
    if l and l[0] != prev:
        return l[0]
    if len(l) >= 2:
        return l[1]
    return None


def linear_chains(n: int, bonds) -> list[list[int]]:
    """Chains as member lists walked from the first end (AnalyzerMsd.tpp:186-288).
    Raises ValueError with the reference's message like its std::invalid_argument
    / std::runtime_error."""
    return _linear_chains(n, bonds, True)


def _chains_without_middle_check(n: int, bonds) -> list[list[int]]:
    return _linear_chains(n, bonds, False)


def _linear_chains(n: int, bonds, middle_check: bool) -> list[list[int]]:
    links = link_lists(n, bonds)
    visited = [False] * n
    chains = []
    for start in range(n):
        if visited[start]:
            continue
        ends = []
        todo = [start]
        visited[start] = True
        size = 1
        while todo:
            m = todo.pop()
            if len(links[m]) > 2:
                raise ValueError(ERR_BRANCHED)
            if len(links[m]) <= 1:
                ends.append(m)
            for p in links[m]:
                if not visited[p]:
                    size += 1
                    visited[p] = True
                    todo.append(p)
        if size == 1 and len(ends) == 1:
            ends.append(ends[0])
        if len(ends) != 2:
            raise ValueError(ERR_ENDS)
        mid = prev = ends[0]
        for _ in range(size // 2):
            nxt = _step(links, mid, prev)
            prev, mid = mid, nxt
            if middle_check and len(links[mid]) < 2:
                raise ValueError(ERR_MIDDLE)
        members = []
        cur = prev = ends[0]
        while True:
            members.append(cur)
            nxt = _step(links, cur, prev)
            if nxt is None or len(members) == size:
                break
            prev, cur = cur, nxt
        chains.append(members)
    return chains


def chain_ends(chains) -> dict:
    return {"first": np.array([c[0] for c in chains], dtype=np.uint32),
            "middle": np.array([c[len(c) // 2] for c in chains], dtype=np.uint32),
            "last": np.array([c[-1] for c in chains], dtype=np.uint32)}


def chain_sums(xyz, chains) -> list[dict]:
    """Per chain, relative to the first member: sum of r - r_first, sum of
    |r - r_first|^2, first / middle / last positions, member count."""
    x = np.asarray(xyz, dtype=np.int64).reshape(-1, 3)
    out = []
    for c in chains:
        f = x[c[0]]
        d = x[c] - f
        out.append({"s": d.sum(axis=0).tolist(), "q": int((d * d).sum()), "first": x[c[0]].tolist(),
                    "middle": x[c[len(c) // 2]].tolist(), "last": x[c[-1]].tolist(), "n": len(c)})
    return out


def chain_com(xyz, members) -> list[float]:
    """AnalyzerMsd.tpp:319-344: positions summed in doubles in walk order, / n."""
    com = [0.0, 0.0, 0.0]
    for m in members:
        for a in range(3):
            com[a] += float(xyz[m][a])
    return [v / float(len(members)) for v in com]


def observables(xyz, bonds, n: int) -> dict:
    x = np.asarray(xyz, dtype=np.int64).reshape(-1, 3)
    hist = np.zeros(512, dtype=np.uint64)
    pairs = sorted({(min(int(a), int(b)), max(int(a), int(b))) for a, b in np.asarray(bonds).reshape(-1, 2)})
    for a, b in pairs:
        d = x[b] - x[a]
        hist[(int(d[0]) & 7) | ((int(d[1]) & 7) << 3) | ((int(d[2]) & 7) << 6)] += 1
    out = {"bonds": len(pairs), "bond_hist": hist}
    try:
        chains = linear_chains(n, bonds)
    except ValueError as e:
        # two-monomer chains only fail the middle-monomer walk of the MSD
        # analyzer; Re^2 and Rg^2 are defined for them
        chains = _chains_without_middle_check(n, bonds) if str(e) == ERR_MIDDLE else None
    if chains is None:
        out.update(n_chains=0, re2_sum=0, re2=math.nan, rg2=math.nan, box_com=[math.nan] * 3)
        return out
    re2_sum = 0
    rg2 = 0.0
    box = [0.0, 0.0, 0.0]
    for c in chains:
        d = x[c[-1]] - x[c[0]]
        re2_sum += int((d * d).sum())
        com = chain_com(x, c)
        rg2 += sum(float(((x[m] - np.array(com)) ** 2).sum()) for m in c) / len(c)
        for a in range(3):
            box[a] += com[a]
    C = len(chains)
    out.update(n_chains=C, re2_sum=re2_sum, re2=re2_sum / C, rg2=rg2 / C, box_com=[v / C for v in box])
    return out


class OnlineStatistics:
    """AnalyzerMsd.tpp:18-95 with two moments."""

    def __init__(self, n_max_moment: int = 2):
        self.moments = [0.0] * (1 + n_max_moment)

    @staticmethod
    def choose(total: int, k: int) -> float:
        if k > total // 2:
            k = total - k
        r = 1.0
        for i in range(k):
            r *= float(total - i) / float(i + 1)
        return r

    def add_value(self, x: float) -> None:
        p = 1.0
        for i in range(len(self.moments)):
            self.moments[i] += p
            p *= x

    def moment(self, i: int) -> float:
        m0 = self.moments[0]
        return self.moments[i] / m0 if m0 else math.nan

    def central_moment(self, k: int) -> float:
        cmom = 0.0
        for i in range(k + 1):
            cmom += (self.choose(len(self.moments), i) * self.moment(k - i) * (1 if i % 2 == 0 else -1)
                     * math.pow(self.moment(1), i))
        m0 = self.moments[0]
        return cmom * (m0 / (m0 - 1)) if m0 != 1 else math.copysign(math.inf, cmom) if cmom else math.nan

    def mean(self) -> float:
        return self.moment(1)

    def sg(self) -> float:
        v = self.central_moment(2)
        return math.sqrt(v) if v >= 0 else math.nan

    @property
    def n(self) -> float:
        return self.moments[0]


class MsdAnalyzer:
    """AnalyzerMsd (:99-468) over unwrapped positions given per call.  As in the
    device analyzer, the first evaluation starts the first subsequence (the
    reference leaves that time uninitialised, :107)."""

    N_MSDS = 5

    def __init__(self, n: int, bonds, mcs_to_skip: int, dt_realization: int, dt_evaluate: int):
        self.chains = linear_chains(n, bonds)
        self.mcs_to_skip = mcs_to_skip
        self.dt_real = dt_realization
        self.dt_eval = dt_evaluate
        self.last_eval = (-dt_evaluate) % (1 << 64)
        self.last_real = 0
        self.any_real = False
        self.first_values = []      # [sub][g][chain] -> vector
        self.first_box = []
        self.first_mcs = []
        self.msds = []              # [eval][g] -> OnlineStatistics
        self.eval_times = []

    def execute(self, age: int, xyz) -> None:
        x = np.asarray(xyz, dtype=np.int64).reshape(-1, 3)
        if (age - self.last_eval) % (1 << 64) < self.dt_eval:
            return
        self.last_eval = age
        C = len(self.chains)
        coms = [chain_com(x, c) for c in self.chains]
        box = [0.0, 0.0, 0.0]
        for com in coms:
            for a in range(3):
                box[a] += com[a]
        box = [v / float(C) for v in box]
        n_sub = len(self.first_values)
        if not self.any_real or age - self.last_real >= self.dt_real:
            self.any_real = True
            self.last_real = age
            fv = [[None] * C for _ in range(self.N_MSDS)]
            for k, c in enumerate(self.chains):
                mid = c[len(c) // 2]
                fv[0][k] = [float(v) for v in x[c[0]]]
                fv[1][k] = [float(v) for v in x[mid]]
                fv[2][k] = [float(v) for v in x[c[-1]]]
                fv[3][k] = list(coms[k])
                fv[4][k] = [fv[1][k][a] - coms[k][a] for a in range(3)]
            self.first_values.append(fv)
            self.first_box.append(box)
            self.first_mcs.append(age)
        for s in range(n_sub):
            e = (age - self.first_mcs[s]) // self.dt_eval
            if e >= len(self.msds):
                while e >= len(self.msds):
                    self.msds.append([OnlineStatistics() for _ in range(self.N_MSDS)])
                while e >= len(self.eval_times):
                    self.eval_times.append(0)
                self.eval_times[e] = age - self.first_mcs[s]

        def dot(d):
            return (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]

        for s in range(n_sub):
            e = (age - self.first_mcs[s]) // self.dt_eval
            dbox = [box[a] - self.first_box[s][a] for a in range(3)]
            fv = self.first_values[s]
            for k, c in enumerate(self.chains):
                picks = (c[0], c[len(c) // 2], c[-1])
                for j in range(3):
                    d = [float(x[picks[j]][a]) - fv[j][k][a] for a in range(3)]
                    if C > 1:
                        d = [d[a] - dbox[a] for a in range(3)]
                    self.msds[e][j].add_value(dot(d))
                d = [coms[k][a] - fv[3][k][a] for a in range(3)]
                if C > 1:
                    d = [d[a] - dbox[a] for a in range(3)]
                self.msds[e][3].add_value(dot(d))
                d = [(float(x[picks[1]][a]) - coms[k][a]) - fv[4][k][a] for a in range(3)]
                self.msds[e][4].add_value(dot(d))

    def points(self) -> list[dict]:
        out = []
        for e, row in enumerate(self.msds):
            out.append({"dt": self.eval_times[e], "n_values": row[0].n,
                        "mean": [st.mean() for st in row], "sg": [st.sg() for st in row]})
        return out
