"""Degenerate 'X' reference bases (multi-path cells, impl/pairwiseAligner.c:501-692):
the oracle pinned by the reference's path KATs and by the mixture identity, and the
GPU path held to the oracle.
"""
import itertools
import math

import numpy as np
import pytest

import oracle
from signalalign_amd import models


def test_path_kmers_match_the_reference_kats():
    """tests/signalPairwiseAlignerTest.c:556-581 (test_hdCellConstruct,
    test_hdCellConstructWorstCase): ATGXAX over CEO has 9 paths, the first ATGCAC and
    the last ATGOAO; XXXXXX has 729, CCCCCC to OOOOOO."""
    p = oracle.path_kmers("ATGXAX", "CEO")
    assert len(p) == 9 and p[0] == "ATGCAC" and p[8] == "ATGOAO"
    p = oracle.path_kmers("XXXXXX", "CEO")
    assert len(p) == 729 and p[0] == "CCCCCC" and p[728] == "OOOOOO"
    assert oracle.path_kmers("ACGTAC", "CEO") == ["ACGTAC"]
    # the first X varies slowest
    assert oracle.path_kmers("XAX", "CE") == ["CAC", "CAE", "EAC", "EAE"]


def test_path_legality_matches_the_reference_kat():
    """tests/variableOrderPairwiseAlignerTests.c:107-123 (test_pathLegalTransitions)."""
    assert oracle.path_legal(None, "AAETTT")
    assert oracle.path_legal("AAETTT", None)
    assert oracle.path_legal("AAETTT", "AETTTC")
    assert not oracle.path_legal("AAETTT", "ACTTTC")


def _tiny(seed, k=3):
    m = models.synthetic_gaussian_model("ACEGOT", k, seed=seed)
    rng = np.random.default_rng(seed)
    ref = list(rng.choice(list("ACGT"), 9))
    ref[3] = "X"
    ref[6] = "X"
    ref = "".join(ref)
    kids = models.kmer_ids(ref.replace("X", "C"), m.alphabet, k)
    ev = np.zeros((6, 4))
    for j in range(6):
        t = m.match[kids[min(j, len(kids) - 1)]]
        ev[j, 0] = t[0] + rng.normal(0, t[1])
        ev[j, 1] = rng.wald(t[2], t[4])
    return m, ref, ev


@pytest.mark.parametrize("ragged", [(False, False), (True, True)])
@pytest.mark.parametrize("options", ["CEO", "CE"])
def test_degenerate_cells_are_the_mixture_over_site_options(ragged, options, oracle_lib):
    """A path through the matrix keeps one option per 'X' site (legal transitions
    agree on the overlap), so with the exact log-sum-exp the degenerate run's total is
    the log-sum-exp of the totals of every substituted reference, and a pair's
    posterior on path p is the weighted sum of the substituted runs' posteriors over
    the substitutions that path p agrees with (weights exp(total_a - total))."""
    m, ref, ev = _tiny(5)
    k = m.kmer_length
    om = oracle.Model(oracle.GAUSS, m.alphabet, k, m.transitions, m.match, m.gapy, m.gapx)
    params = oracle.Params(threshold=0.0, diagonal_expansion=20)
    sites = [i for i, c in enumerate(ref) if c == "X"]
    oracle.set_exact_logadd(True)
    try:
        multi, st = oracle.get_aligned_pairs(om, ref, ev, [], params, *ragged, options=options)
        runs = []
        for a in itertools.product(options, repeat=len(sites)):
            r = list(ref)
            for s_, o in zip(sites, a):
                r[s_] = o
            pairs, sa = oracle.get_aligned_pairs(om, "".join(r), ev, [], params, *ragged)
            runs.append((a, sa["log_likelihood"], {(int(x), int(y)): p / 1e7 for p, x, y in pairs}))
    finally:
        oracle.set_exact_logadd(False)
    lls = np.array([ll for _, ll, _ in runs])
    mx = lls.max()
    total = mx + math.log(np.exp(lls - mx).sum())
    assert abs(st["log_likelihood"] - total) < 1e-9
    got = {}
    for p, x, y, path in multi:
        got[(int(x), int(y), int(path))] = p / 1e7
    want = {}
    for x in range(len(ref) - k + 1):
        kmer_sites = [s_ for s_ in sites if x <= s_ < x + k]
        paths = oracle.path_kmers(ref[x:x + k], options)
        for pi, pk in enumerate(paths):
            opts = {s_: pk[s_ - x] for s_ in kmer_sites}
            for a, ll, post in runs:
                if all(a[sites.index(s_)] == o for s_, o in opts.items()):
                    for (xx, yy), pv in post.items():
                        if xx == x:
                            want[(x, yy, pi)] = want.get((x, yy, pi), 0.0) + math.exp(ll - total) * pv
    assert set(got) == set(want)
    for key, pv in want.items():
        assert abs(got[key] - min(pv, 1.0)) < 2e-6 * len(runs), (key, got[key], pv)


def test_degenerate_without_x_is_the_plain_run(oracle_lib):
    """Options change nothing on a reference with no 'X'."""
    m, ref, ev = _tiny(7)
    ref = ref.replace("X", "G")
    om = oracle.Model(oracle.GAUSS, m.alphabet, m.kmer_length, m.transitions, m.match, m.gapy, m.gapx)
    p1, s1 = oracle.get_aligned_pairs(om, ref, ev, [], oracle.Params(threshold=0.01, diagonal_expansion=20))
    p2, s2 = oracle.get_aligned_pairs(om, ref, ev, [], oracle.Params(threshold=0.01, diagonal_expansion=20),
                                      options="CEO")
    assert np.array_equal(p1, p2[:, :3]) and np.all(p2[:, 3] == 0)
    assert s1["log_likelihood"] == s2["log_likelihood"]
