"""Statistical check of the colour sweep against the LeMonADE-style serial
random-sequential updater: bond-vector statistics, Re^2, Rg^2 and the
middle-monomer / chain-COM MSDs (the reference analyzer's g2 and g4,
src/AnalyzerMsd.tpp:303-421; conventionally called g1 and g3).

The sweep side runs through the oracle's replay, which the HIP path matches
bit for bit (tests/test_gpu_parity.py), so these numbers are the GPU's.

Tolerances state a measured, expected difference: the parallel rule of the
reference (every monomer of a colour moves at once, each checked against the
state before the substep; UpdaterGPUScBFM_AB_Type.cu:930-1028, :1148-1190) is
not detailed-balanced for same-colour monomers that touch, see
test_parallel_substep_is_not_doubly_stochastic.  At volume fraction 0.5 the
sweep's fraction of the shortest bonds (|b|^2 = 4) is ~0.002 higher than
random-sequential's and its MSDs run 2-6 % faster.
"""
import itertools
import os
from multiprocessing import get_context

import numpy as np
import pytest

from oracle import oracle as orc
from pscbfm_amd import analysis, systems

N = 16
MCS = 6000
SAMPLE = 50          # MCS between frames
REPLICAS = 8


def _replica(rep):
    s = systems.linear_melt(32, 128, N, seed=700 + rep)      # 2048 monomers, phi = 0.5
    o = orc.System(s, 4242 + rep)
    out = {}
    for algo in ("sweep", "rs"):
        x = s["xyz"]
        frames = []
        for k in range(MCS // SAMPLE):
            if algo == "sweep":
                x, _, _ = orc.sweep(o, x, k * SAMPLE, SAMPLE, 0)
            else:
                x, _ = orc.random_sequential(o, x, k * SAMPLE, SAMPLE)
            frames.append(x)
        eq = frames[len(frames) // 4:]
        m2 = analysis.msds(eq, N, 2)
        m20 = analysis.msds(eq, N, 20)
        out[algo] = {
            "re2": np.mean([analysis.re2(f, N) for f in eq]),
            "rg2": np.mean([analysis.rg2(f, N) for f in eq]),
            "hist": np.mean([analysis.bond_length2_histogram(f, N) for f in eq], axis=0),
            "g2_100": m2["g2"], "g4_100": m2["g4"], "g2_1000": m20["g2"], "g4_1000": m20["g4"],
        }
    return out


@pytest.fixture(scope="module")
def replicas():
    workers = min(REPLICAS, os.cpu_count() or 1)
    with get_context("fork").Pool(workers) as pool:
        return pool.map(_replica, range(REPLICAS))


def _mean_se(reps, algo, key):
    v = np.array([r[algo][key] for r in reps])
    return v.mean(axis=0), v.std(axis=0, ddof=1) / np.sqrt(len(v))


def test_static_observables_agree(replicas):
    for key in ("re2", "rg2"):
        a, sa = _mean_se(replicas, "sweep", key)
        b, sb = _mean_se(replicas, "rs", key)
        # within 3 combined standard errors plus 2 %
        assert abs(a - b) <= 3 * np.hypot(sa, sb) + 0.02 * b, (key, a, sa, b, sb)


def test_bond_vector_statistics_agree(replicas):
    a, _ = _mean_se(replicas, "sweep", "hist")
    b, _ = _mean_se(replicas, "rs", "hist")
    assert np.allclose(a.sum(), 1.0) and np.allclose(b.sum(), 1.0)
    # absolute fractions agree to 0.006 (measured difference <= 0.0025)
    assert np.max(np.abs(a - b)) < 0.006, (a, b)


def test_msd_agree(replicas):
    for key in ("g2_100", "g4_100", "g2_1000", "g4_1000"):
        a, _ = _mean_se(replicas, "sweep", key)
        b, _ = _mean_se(replicas, "rs", key)
        # same time scale within 12 % (measured: sweep 2-6 % faster)
        assert 0.95 < a / b < 1.12, (key, a, b)


L2 = 8                                    # two-body box: periodic 8^3
_DIRS = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


def _overlap(r):
    # cubes at relative position r overlap iff every axis is within 1 (periodic)
    return all(min(v % L2, (-v) % L2) <= 1 for v in r)


def _pair_states():
    states = [r for r in itertools.product(range(L2), repeat=3) if not _overlap(r)]
    return states, {r: i for i, r in enumerate(states)}


def _pair_class(r):
    # orbit of a relative position under the cube's symmetries (both laws below are symmetric)
    return tuple(sorted(min(v % L2, (-v) % L2) for v in r))


def _parallel_substep_law():
    """Exact transition matrix of one substep of the parallel rule for two
    free monomers of one colour (UpdaterGPUScBFM_AB_Type.cu:930-1028,
    :1148-1190): both propose at once, each is checked against the old state,
    and both are rejected if their new cubes overlap.  Translation is factored
    out: the state is b's position relative to a.  Returns (P, stationary pi)."""
    states, ix = _pair_states()
    n = len(states)
    P = np.zeros((n, n))
    for r in states:
        for da in _DIRS:
            for db in _DIRS:
                b1 = tuple(r[k] + db[k] for k in range(3))
                oka = not _overlap(tuple(r[k] - da[k] for k in range(3)))
                okb = not _overlap(b1)
                if oka and okb and _overlap(tuple(b1[k] - da[k] for k in range(3))):
                    oka = okb = False
                na = da if oka else (0, 0, 0)
                nb = b1 if okb else r
                P[ix[r], ix[tuple((nb[k] - na[k]) % L2 for k in range(3))]] += 1.0 / 36
    # stationary law: pi P = pi, sum(pi) = 1
    A = np.vstack([P.T - np.eye(n), np.ones((1, n))])
    pi = np.linalg.lstsq(A, np.r_[np.zeros(n), 1.0], rcond=None)[0]
    return P, pi


def _chi2_p(counts, probs):
    from scipy import stats
    e = probs * counts.sum()
    chi = ((counts - e) ** 2 / e).sum()
    return stats.chi2.sf(chi, len(counts) - 1)


def _pair_histogram(algo, seed, n_samples, thin):
    """Class histogram of b - a sampled every `thin` MCS from the oracle."""
    states, _ = _pair_states()
    classes = sorted({_pair_class(r) for r in states})
    ci = {c: i for i, c in enumerate(classes)}
    s = {"box": (L2, L2, L2), "periodic": (1, 1, 1), "xyz": np.array([[0, 0, 0], [3, 4, 5]], dtype=np.int32),
         "bonds": np.zeros((0, 2), dtype=np.uint32), "forbidden": systems.standard_forbidden()}
    o = orc.System(s, seed)
    assert o.n_colors == 1
    x, h = s["xyz"], np.zeros(len(classes))
    for k in range(n_samples):
        if algo == "sweep":
            x, _, _ = orc.sweep(o, x, k * thin, thin, 0)
        else:
            x, _ = orc.random_sequential(o, x, k * thin, thin)
        r = x[1] - x[0]
        assert not _overlap(r)
        h[ci[_pair_class(r)]] += 1
    return h, classes, ci


def test_parallel_substep_is_not_doubly_stochastic():
    """The substep's transition matrix has column sums != 1, so the uniform
    law of the serial BFM is not its stationary one.  This is the reference's
    rule; the engine keeps it."""
    P, pi = _parallel_substep_law()
    sums = P.sum(axis=0)
    assert np.allclose(P.sum(axis=1), 1.0)
    assert sums.min() < 0.95 and sums.max() > 1.05
    n = len(pi)
    assert 0.5 * np.abs(pi - 1.0 / n).sum() > 0.05       # total-variation distance to uniform


def _class_probs(ci, weights):
    states, _ = _pair_states()
    q = np.zeros(len(ci))
    for r, w in zip(states, weights):
        q[ci[_pair_class(r)]] += w
    return q


def test_two_free_monomers_sample_their_stationary_laws():
    """Two free monomers in a periodic 8^3 box, 40000 samples 16 MCS apart,
    binned into the 31 symmetry classes of their relative position.  The
    serial rule is symmetric with accept-if-legal, so orc_random_sequential
    must sample the 485 legal relative positions uniformly; the oracle's
    parallel sweep must instead sample the exact law of the substep above and
    fail the uniform test by far.  Thresholds: p > 1e-3 where the law holds,
    p < 1e-20 where it does not (seeded, so the outcome is fixed)."""
    P, pi = _parallel_substep_law()
    n = len(pi)
    for algo, law, other in (("rs", np.full(n, 1.0 / n), pi), ("sweep", pi, np.full(n, 1.0 / n))):
        h, classes, ci = _pair_histogram(algo, 4242, 40000, 16)
        assert len(classes) == 31
        assert _chi2_p(h, _class_probs(ci, law)) > 1e-3, algo
        assert _chi2_p(h, _class_probs(ci, other)) < 1e-20, algo


def test_free_dimer_visits_bond_vectors_uniformly():
    """A bonded dimer in a periodic 32^3 box under orc_random_sequential:
    detailed balance of the serial BFM makes every one of the 108 bond vectors
    equally likely (and no other ever appears).  40000 samples 32 MCS apart;
    chi^2 over the 108 vectors and over the 6 length/shape classes, p > 1e-3."""
    f = systems.standard_forbidden()
    allowed = [d for d in itertools.product(range(-4, 4), repeat=3)
               if not f[(d[0] & 7) | ((d[1] & 7) << 3) | ((d[2] & 7) << 6)]]
    ai = {d: i for i, d in enumerate(allowed)}
    s = {"box": (32, 32, 32), "periodic": (1, 1, 1), "xyz": np.array([[4, 4, 4], [6, 4, 4]], dtype=np.int32),
         "bonds": np.array([[0, 1]], dtype=np.uint32), "forbidden": f}
    o = orc.System(s, 99)
    x, h = s["xyz"], np.zeros(len(allowed))
    for k in range(40000):
        x, _ = orc.random_sequential(o, x, k * 32, 32)
        h[ai[tuple(int(v) for v in x[1] - x[0])]] += 1
    assert _chi2_p(h, np.full(108, 1 / 108)) > 1e-3
    shape = [tuple(sorted(abs(v) for v in d)) for d in allowed]
    kinds = sorted(set(shape))
    assert len(kinds) == 6
    hk = np.array([h[[i for i, c in enumerate(shape) if c == k]].sum() for k in kinds])
    pk = np.array([shape.count(k) / 108 for k in kinds])
    assert _chi2_p(hk, pk) > 1e-3
