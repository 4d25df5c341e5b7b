"""Pins the FP64 oracle against the reference's own known-answer tests, closed
forms and invariants (CPU tests).

Each test names the reference test or fact it reproduces.  A plausible slip in
the restatement (a dropped transition, a wrong index, a swapped table) breaks at
least one of these.
"""
import itertools
import math
import os

import numpy as np
import pytest
from scipy import stats

import oracle
from signalalign_amd import models

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# tests/stateMachineTests.c:565-692 (test_sm3_5merDiagonalDPCalculations)
KAT5_REF = "ACGATATGGACAT"
KAT5_EVENTS = np.array([
    [70.0423375640843, 2.1070814631739, 0.0571, 0.0],
    [73.7087073662952, 1.90162684687837, 0.0571, 0.1],
    [105.375581864011, 2.87252862011704, 0.0571, 0.2],
    [82.9620934477158, 2.38320603353748, 0.0571, 0.3],
    [84.6977645711335, 3.08486975249442, 0.0571, 0.4],
    [58.0551144225027, 2.52297561817531, 0.0571, 0.5],
    [94.337668063878, 1.9731952395105, 0.0571, 0.67],
])
KAT5_PAIRS = [(0, 0), (1, 1), (2, 2), (3, 3), (5, 4), (6, 5), (8, 6)]


@pytest.fixture(scope="module")
def m5(oracle_lib):
    m = models.load_model_file(os.path.join(GOLDEN, "testModelR9_5mer_acgt_template.model"))
    return m, oracle.Model(oracle.GAUSS, m.alphabet, m.kmer_length, m.transitions, m.match, m.gapy, m.gapx)


# --------------------------------------------------------------------------
# band and split geometry
# --------------------------------------------------------------------------

def test_band_kat(oracle_lib):
    """tests/signalPairwiseAlignerTest.c:447-510 (test_bands)."""
    L, R = oracle.band([(1, 0), (2, 1), (3, 3)], 6, 5, 2)
    want = [(0, 0), (-1, 1), (-2, 2), (-1, 3), (-2, 4), (-1, 3), (-2, 4), (-3, 3), (-2, 2), (-1, 3), (0, 2), (1, 1)]
    assert list(zip(L.tolist(), R.tolist())) == want


def test_band_closed_form_without_anchors(oracle_lib):
    """With no anchors and expansion >= the sequence lengths the band is the whole
    lX x lY rectangle: xmy ranges over max(-d, d-2lY) .. min(d, 2lX-d)."""
    lX, lY = 7, 5
    L, R = oracle.band([], lX, lY, 20)
    for d in range(lX + lY + 1):
        assert L[d] == max(-d, d - 2 * lY)
        assert R[d] == min(d, 2 * lX - d)


def test_band_rejects_bad_anchors(oracle_lib):
    with pytest.raises(ValueError):
        oracle.band([(3, 3), (2, 4)], 6, 5, 2)
    with pytest.raises(ValueError):
        oracle.band([(1, 1)], 6, 5, 3)  # odd expansion


def test_split_points_kat(oracle_lib):
    """tests/signalPairwiseAlignerTest.c:376-445 (test_getSplitPoints)."""
    ms = 2000 * 2000
    sp = lambda a, lX, lY, rl, rr: oracle.split_points(a, lX, lY, ms, rl, rr).tolist()
    assert sp([], 3000, 1000, 0, 0) == [[0, 0, 3000, 1000]]
    assert sp([], 20000, 25000, 1, 1) == []
    assert sp([], 20000, 25000, 1, 0) == [[18000, 23000, 20000, 25000]]
    assert sp([], 20000, 25000, 0, 1) == [[0, 0, 2000, 2000]]
    assert sp([], 20000, 25000, 0, 0) == [[0, 0, 2000, 2000], [18000, 23000, 20000, 25000]]
    anchors = [(2000, 2000), (4002, 4001), (5000, 5000), (8000, 6000), (9000, 9000), (10000, 14000),
               (15000, 15000), (16000, 16000)]
    assert sp(anchors, 20000, 25000, 0, 0) == [[0, 0, 3001, 3001], [3002, 3001, 9500, 11001],
                                                [9501, 12000, 12001, 14500], [13000, 14501, 18000, 18001],
                                                [18001, 23000, 20000, 25000]]


# --------------------------------------------------------------------------
# logAdd and emissions
# --------------------------------------------------------------------------

def test_logadd_reference_tolerance(oracle_lib):
    """tests/signalPairwiseAlignerTest.c:121-131: |exp(logAdd(log i, log j)) - (i+j)| < 1e-3."""
    rng = np.random.default_rng(5)
    for i, j in rng.random((20000, 2)):
        assert abs(math.exp(oracle.logadd(math.log(i), math.log(j))) - (i + j)) < 1e-3


def test_logadd_closed_form(oracle_lib):
    """Pieces: cubic in |x-y| with the reference's coefficients; exact max beyond 7.5;
    -inf operands pass the other through."""
    assert oracle.logadd(-math.inf, 3.0) == 3.0
    assert oracle.logadd(3.0, -math.inf) == 3.0
    assert oracle.logadd(-math.inf, -math.inf) == -math.inf
    assert oracle.logadd(10.0, 2.5) == 10.0
    assert oracle.logadd(2.0, 2.0) == 2.0 + oracle.lookup(0.0)
    c = [(-0.009350833524763, 0.130659527668286, 0.498799810682272, 0.693203116424741),
         (-0.014532321752540, 0.139942324101744, 0.495635523139337, 0.692140569840976),
         (-0.004605031767994, 0.063427417320019, 0.695956496475118, 0.514272634594009),
         (-0.000458661602210, 0.009695946122598, 0.930734667215156, 0.168037164329057)]
    f32 = lambda v: float(np.float32(v))
    for d in np.linspace(0.0, 7.49, 300):
        piece = 0 if d <= 1.0 else 1 if d <= 2.5 else 2 if d <= 4.5 else 3
        a, b, cc, e = (f32(v) for v in c[piece])
        assert oracle.lookup(d) == ((a * d + b) * d + cc) * d + e
        assert abs(oracle.lookup(d) - math.log1p(math.exp(d))) < 6e-4
    # the piecewise approximation is monotone above max(x, y), so chains never decrease
    for d in np.linspace(0.0, 7.4999, 2000):
        assert oracle.lookup(d) > d


def test_exact_logadd_switch(oracle_lib):
    oracle.set_exact_logadd(True)
    try:
        assert abs(oracle.logadd(1.0, 2.0) - math.log(math.e + math.e ** 2)) < 1e-14
        assert oracle.logadd(10.0, 0.0) > 10.0
    finally:
        oracle.set_exact_logadd(False)
    assert oracle.logadd(10.0, 0.0) == 10.0


def test_lookup_logadd_knot_sensitivity(oracle_lib):
    """Why an FP32 recurrence is held to a looser posterior bar than FP64 (CPU experiment).

    The reference's lookup logAdd (impl/pairwiseAligner.c:234-254) jumps at a difference
    of 7.5: just below it a term still adds lookup(7.5) - 7.5 = 4.47e-4 to the log sum,
    at 7.5 and beyond it adds nothing. A total refresh (impl/pairwiseAligner.c:1367-1379)
    with a term that close to the knot moves by 4.47e-4 when the term moves by any amount
    that crosses it, and every posterior of the following diagonals scales with it.
    Read 6 of the batched workload (4644 nt, seed 2024): rounding its event means to FP32
    (a relative change below 6e-8, the least any FP32 recurrence must make) moves the FP64
    oracle's posteriors by more than 3e-4; the same rounding moves them by under 1e-5
    when the oracle uses the exact log-sum-exp. So the 1e-4 posterior bar cannot hold on
    every pair for arithmetic narrower than FP64, and the excess comes from the knots,
    not from the rounding drift itself."""
    from signalalign_amd import native, synth
    from tests import _parity as P

    m = synth.gaussian_model("ACEGT", 6, seed=1)
    r = synth.batch_reads(m, 7, 8000, seed=2024, methyl_cpg=0.3)[6]
    assert len(r["ref"]) == 4644
    r32 = dict(r)
    r32["events"] = np.asarray(r["events"], dtype=np.float32).astype(np.float64)
    assert np.max(np.abs(r32["events"][:, 0] / r["events"][:, 0] - 1)) < 6e-8
    params = native.Params(diagonal_expansion=50)

    def max_dp(exact):
        oracle.set_exact_logadd(exact)
        try:
            a, sa = P.run_oracle(m, r, params)
            b, sb = P.run_oracle(m, r32, params)
        finally:
            oracle.set_exact_logadd(False)
        ga = {(x, y): p / 1e7 for p, x, y in a}
        gb = {(x, y): p / 1e7 for p, x, y in b}
        common = [k for k in ga if k in gb]
        assert len(common) > 5000
        return max(abs(ga[k] - gb[k]) for k in common), P.rel(sb["log_likelihood"], sa["log_likelihood"])

    dp_lookup, ll_lookup = max_dp(False)
    dp_exact, _ = max_dp(True)
    assert dp_lookup > 3e-4, dp_lookup
    assert dp_exact < 1e-5, dp_exact
    assert ll_lookup < 1e-8  # the log-likelihood keeps the FP64 bar


def test_emission_pdfs_closed_form(oracle_lib):
    """tests/signalPairwiseAlignerTest.c:81-111 at 1e-3, here against scipy at 1e-12."""
    rng = np.random.default_rng(9)
    for _ in range(200):
        mu, sd, x = rng.uniform(50, 120), rng.uniform(0.5, 4), rng.uniform(40, 130)
        assert abs(oracle.log_gauss_pdf(x, mu, sd) - stats.norm.logpdf(x, mu, sd)) < 1e-12
        nmu, lam, n = rng.uniform(0.5, 3), rng.uniform(0.5, 5), rng.uniform(0.2, 4)
        assert abs(oracle.log_inv_gauss_pdf(n, nmu, lam) - stats.invgauss.logpdf(n, nmu / lam, scale=lam)) < 1e-10
    assert oracle.log_gauss_pdf(1.0, 0.0, 0.0) == -math.inf


def test_match_emission_descaling(oracle_lib):
    """impl/stateMachine.c:498-545: the event mean is descaled with
    (x + var*mu - scale*mu - shift)/var before the Gaussian; the noise term is
    the inverse Gaussian; gapY reads the second table."""
    m = models.synthetic_gaussian_model("ACGT", 6, seed=2)
    om = oracle.Model(oracle.GAUSS_DESCALED, m.alphabet, 6, m.transitions, m.match, m.gapy, m.gapx, 1.03, -2.0, 1.1)
    kmer = "ACGTTG"
    k = oracle.kmer_id(kmer, "ACGT")
    mu, sd, nmu, _, lam = m.match[k]
    ev = (81.0, 1.3)
    z = (ev[0] + 1.1 * mu - 1.03 * mu + 2.0) / 1.1
    want = stats.norm.logpdf(z, mu, sd) + stats.invgauss.logpdf(ev[1], nmu / lam, scale=lam)
    assert abs(oracle.emission(om, kmer, ev, True) - want) < 1e-10
    sdy = m.gapy[k][1]
    assert abs(sdy - 1.75 * sd) < 1e-12
    want_y = stats.norm.logpdf(z, mu, sdy) + stats.invgauss.logpdf(ev[1], nmu / lam, scale=lam)
    assert abs(oracle.emission(om, kmer, ev, False) - want_y) < 1e-10
    assert oracle.emission(om, None, ev, True) == -math.inf


def test_kmer_id_kats(oracle_lib):
    """tests/nanoporeHdpTests.c:102-108 (test_kmer_id) and
    tests/variableOrderPairwiseAlignerTests.c:158-167 (test_getKmerIndex: the i-th
    k-mer of path_listPotentialKmers over ACEGOT has id i)."""
    assert oracle.kmer_id("AAAC", "ACGT") == 1
    assert oracle.kmer_id("AAAT", "ACGT") == 3
    assert oracle.kmer_id("AAAT", "ACT") == 2
    assert oracle.kmer_id("GGGG", "ABCDEFG") == 7 ** 4 - 1
    assert oracle.kmer_id("AACAA", "ACGT") == 16
    assert oracle.kmer_id("AANAAA", "ACGT") == -1
    allk = ["".join(p) for p in itertools.product("ACEGOT", repeat=6)]
    for i in range(0, len(allk), 997):
        assert oracle.kmer_id(allk[i], "ACEGOT") == i
    ids = models.kmer_ids("".join(allk[i] for i in range(0, 46656, 6)), "ACEGOT", 6)
    assert np.array_equal(ids[::6], np.arange(0, 46656, 6))


def test_pore_model_parse(tmp_path, oracle_lib):
    """tests/stateMachineTests.c:235-293 (test_poreModel): values load in order and
    the gapY level_sd is the match level_sd x 1.75."""
    for k, alpha in [(6, "ACGT"), (5, "ACEGT"), (3, "ACEGOT")]:
        n = len(alpha) ** k
        f = tmp_path / f"m{k}{alpha}.model"
        with open(f, "w") as fh:
            fh.write(f"3\t{len(alpha)}\t{alpha}\t{k}\n")
            fh.write("\t".join(str(i) for i in range(10)) + "\t\n")
            fh.write("\t".join(str(i) for i in range(5 * n)) + "\t\n")
        m = models.load_model_file(str(f))
        assert m.alphabet == "".join(sorted(alpha)) and m.kmer_length == k
        assert np.array_equal(m.match.ravel(), np.arange(5 * n, dtype=float))
        assert np.allclose(m.gapy[:, 1], m.match[:, 1] * 1.75, rtol=0, atol=0)
        # transitions: log of tokens 0,1,2,3,4,6,7,8; gapY->gapX goes to SWITCH_TO_Y;
        # SWITCH_TO_X stays log(0) (impl/stateMachine.c:1086, :1090-1146)
        t = m.transitions
        assert t[models.T_MATCH_CONTINUE] == -math.inf  # log(0)
        assert t[models.T_GAP_OPEN_X] == math.log(1)
        assert t[models.T_GAP_OPEN_Y] == math.log(2)
        assert t[models.T_MATCH_FROM_GAP_X] == math.log(3)
        assert t[models.T_GAP_EXTEND_X] == math.log(4)
        assert t[models.T_MATCH_FROM_GAP_Y] == math.log(6)
        assert t[models.T_GAP_SWITCH_TO_Y] == math.log(7)
        assert t[models.T_GAP_EXTEND_Y] == math.log(8)
        assert t[models.T_GAP_SWITCH_TO_X] == -math.inf


# --------------------------------------------------------------------------
# the DP itself
# --------------------------------------------------------------------------

def test_reference_5mer_forward_backward_kat(m5):
    """tests/stateMachineTests.c:565-692: forward total == backward total (1e-3), every
    diagonal's total probability within 1e-2 of it, and exactly 7 pairs at threshold 0.2."""
    m, om = m5
    pairs, tf, tb, diag = oracle.full_forward_backward(om, KAT5_REF, KAT5_EVENTS, [], 2, 0.2)
    assert abs(tf - tb) < 1e-3
    assert np.all(np.abs(diag - tf) < 1e-2)
    assert sorted((int(x), int(y)) for _, x, y in pairs) == KAT5_PAIRS
    assert len(pairs) == 7


def test_reference_5mer_through_banded_path(m5):
    """The same inputs through getAlignedPairsUsingAnchors (one window, totals
    refreshed at diagonals 16 and 6) give the same 7 pairs."""
    m, om = m5
    pairs, st = oracle.get_aligned_pairs(om, KAT5_REF, KAT5_EVENTS, [], oracle.Params(threshold=0.2,
                                                                                       diagonal_expansion=2),
                                         False, False)
    assert sorted((int(x), int(y)) for _, x, y in pairs) == KAT5_PAIRS
    assert st["n_windows"] == 1


def _brute_force(om, m, ref, ev, ragged_left, ragged_right, exact_emission):
    """Enumerate every 3-state path from (0,0) to (lX,lY): match steps (x+1,y+1) emit
    (k-mer x, event y), gapX steps (x+1,y) emit log 0.1, gapY steps (x,y+1) emit
    (k-mer x-1, event y) and need x >= 1.  Returns (log total, match posteriors)."""
    k = m.kmer_length
    lX, lY = len(ref) - k + 1, len(ev)
    t = m.transitions
    trans = {(0, 0): t[0], (1, 0): t[1], (2, 0): t[2], (0, 1): t[3], (0, 2): t[4], (1, 1): t[5], (2, 2): t[6],
             (2, 1): t[7], (1, 2): -math.inf}
    start = [-math.inf, 0.0, 0.0] if ragged_left else [0.0, -math.inf, -math.inf]
    end = ([(t[3] + t[4]) / 2, t[5], t[6]] if ragged_right else [t[0], t[1], t[2]])
    em = {}
    for x in range(lX):
        for y in range(lY):
            em[(x, y, 1)] = exact_emission(ref[x:x + k], ev[y], True)
            em[(x, y, 0)] = exact_emission(ref[x:x + k], ev[y], False)
    scores = []
    post = {}

    def rec(x, y, s, score, path):
        if x == lX and y == lY:
            scores.append((score + end[s], list(path)))
            return
        moves = []
        if x < lX and y < lY:
            moves.append((x + 1, y + 1, 0, em[(x, y, 1)]))
        if x < lX:
            moves.append((x + 1, y, 1, -2.3025850929940455))
        if y < lY and x >= 1:
            moves.append((x, y + 1, 2, em[(x - 1, y, 0)]))
        for nx, ny, ns, e in moves:
            tp = trans[(s, ns)]
            if tp == -math.inf:
                continue
            if ns == 0:
                path.append((x, y))
            rec(nx, ny, ns, score + e + tp, path)
            if ns == 0:
                path.pop()

    for s0 in range(3):
        if start[s0] > -math.inf:
            rec(0, 0, s0, start[s0], [])
    sc = np.array([s for s, _ in scores])
    mx = sc.max()
    total = mx + math.log(np.exp(sc - mx).sum())
    for s, path in scores:
        for xy in path:
            post[xy] = post.get(xy, 0.0) + math.exp(s - total)
    return total, post


@pytest.mark.parametrize("ragged", [(False, False), (True, True), (True, False)])
def test_brute_force_enumeration(ragged, oracle_lib):
    """Exact log-sum-exp forward/backward equals explicit path enumeration (1e-9),
    for the total and for every match posterior; the lookup logAdd stays within
    the reference's 1e-3 tolerance of it."""
    m = models.synthetic_gaussian_model("ACGT", 3, seed=4)
    rng = np.random.default_rng(12)
    ref = "".join(rng.choice(list("ACGT"), 7))
    kids = models.kmer_ids(ref, "ACGT", 3)
    ev = np.zeros((4, 4))
    for j in range(4):
        t = m.match[kids[min(j, len(kids) - 1)]]
        ev[j, 0] = t[0] + rng.normal(0, t[1])
        ev[j, 1] = rng.wald(t[2], t[4])
    om = oracle.Model(oracle.GAUSS, m.alphabet, 3, m.transitions, m.match, m.gapy, m.gapx)
    exact_em = lambda kmer, e, match: oracle.emission(om, kmer, e, match)
    total, post = _brute_force(om, m, ref, ev, ragged[0], ragged[1], exact_em)
    oracle.set_exact_logadd(True)
    try:
        pairs, st = oracle.get_aligned_pairs(om, ref, ev, [], oracle.Params(threshold=0.0, diagonal_expansion=20),
                                             *ragged)
    finally:
        oracle.set_exact_logadd(False)
    assert abs(st["log_likelihood"] - total) < 1e-9
    got = {(int(x), int(y)): p / 1e7 for p, x, y in pairs}
    for (x, y), p in post.items():
        if x >= 1 and y >= 1:
            pass
    # posterior of a match emitting (k-mer x, event y) is reported at pair (x, y)
    for (x, y), p in post.items():
        assert abs(got.get((x, y), 0.0) - min(p, 1.0)) < 2e-7
    pairs2, st2 = oracle.get_aligned_pairs(om, ref, ev, [], oracle.Params(threshold=0.0, diagonal_expansion=20),
                                           *ragged)
    assert abs(st2["log_likelihood"] - total) < 1e-2


def _window_rule(L, R, N, expansion, min_diags, traceback):
    """The traceback schedule of getPosteriorProbsWithBanding (impl/pairwiseAligner.c:1326-1331,
    1350): a window ends at the first diagonal d >= tracedBackTo + minDiagsBetweenTraceBack
    whose band width is at most 2 * expansion + 1, or at the last diagonal; its posterior
    diagonals are (tracedBackTo, d - traceBackDiagonals - 1], or up to d at the end."""
    wins, T = [], 0
    for d in range(1, N + 1):
        at_end = d == N
        width = (R[d] - L[d]) // 2 + 1
        if at_end or (d >= T + min_diags and width <= 2 * expansion + 1):
            TF = d if at_end else d - traceback - 1
            wins.append((T, d, TF, int(at_end)))
            T = TF
    return wins


def _windowed_brute_force(m, ref, ev, anchors, expansion, min_diags, traceback, ragged_left, ragged_right,
                          exact_emission):
    """Explicit path enumeration of the windowed traceback.  Window (T, D, TF) seeds its
    backward pass on diagonal D with the end prior (the ragged-end prior only on the
    last diagonal of a ragged-right alignment), so with exact arithmetic the posterior
    of a match cell on a diagonal in (T, TF] is

        sum over band paths from the start that end on diagonal D and enter the cell
        as a match, each weighted by start x transitions x emissions x end(state on D)
        / the same sum over every band path ending on D,

    and the total probability of every posterior diagonal of the window is that
    denominator.  Returns (windows, {(x-1, y-1): posterior}, log of the last
    window's denominator)."""
    k = m.kmer_length
    lX, lY = len(ref) - k + 1, len(ev)
    N = lX + lY
    L, R = oracle.band(anchors, lX, lY, expansion)
    wins = _window_rule(L, R, N, expansion, min_diags, traceback)
    t = m.transitions
    trans = {(0, 0): t[0], (1, 0): t[1], (2, 0): t[2], (0, 1): t[3], (0, 2): t[4], (1, 1): t[5], (2, 2): t[6],
             (2, 1): t[7], (1, 2): -math.inf}
    start = [-math.inf, 0.0, 0.0] if ragged_left else [0.0, -math.inf, -math.inf]
    end = [t[0], t[1], t[2]]
    rend = [(t[3] + t[4]) / 2, t[5], t[6]]
    em = {}
    for x in range(lX):
        for y in range(lY):
            em[(x, y, 1)] = exact_emission(ref[x:x + k], ev[y], True)
            em[(x, y, 0)] = exact_emission(ref[x:x + k], ev[y], False)

    def in_band(x, y):
        return L[x + y] <= x - y <= R[x + y]

    top = {D: i for i, (_, D, _, _) in enumerate(wins)}
    recs = [[] for _ in wins]

    def rec(x, y, s, score, matches):
        d = x + y
        if d in top:
            w = top[d]
            e = rend if (wins[w][3] and ragged_right) else end
            if e[s] > -math.inf:
                recs[w].append((score + e[s], matches))
        if d == N:
            return
        moves = []
        if x < lX and y < lY:
            moves.append((x + 1, y + 1, 0, em[(x, y, 1)]))
        if x < lX:
            moves.append((x + 1, y, 1, -2.3025850929940455))
        if y < lY and x >= 1:
            moves.append((x, y + 1, 2, em[(x - 1, y, 0)]))
        for nx, ny, ns, e in moves:
            tp = trans[(s, ns)]
            if tp == -math.inf or not in_band(nx, ny):
                continue
            rec(nx, ny, ns, score + e + tp, matches + ((x, y),) if ns == 0 else matches)

    for s0 in range(3):
        if start[s0] > -math.inf:
            rec(0, 0, s0, start[s0], ())
    post, logz = {}, None
    for (T, D, TF, _), rs in zip(wins, recs):
        sc = np.array([s for s, _ in rs])
        mx = sc.max()
        logz = mx + math.log(np.exp(sc - mx).sum())
        for s, matches in rs:
            w = math.exp(s - logz)
            for (px, py) in matches:
                if T < px + py + 2 <= TF:
                    post[(px, py)] = post.get((px, py), 0.0) + w
        # every band match cell of the window's posterior diagonals, reached or not
        for d in range(T + 1, TF + 1):
            for xmy in range(L[d], R[d] + 1, 2):
                x, y = (d + xmy) // 2, (d - xmy) // 2
                if x >= 1 and y >= 1:
                    post.setdefault((x - 1, y - 1), 0.0)
    return wins, post, logz


_WINDOW_CASES = [
    # (min_diags, traceback, anchors, expansion, ragged)
    (3, 1, [], 20, (False, False)),
    (4, 2, [], 20, (True, True)),
    (5, 1, [], 20, (False, True)),
    (5, 3, [], 20, (True, False)),
    (3, 1, [(2, 1), (4, 3)], 2, (False, False)),
    (4, 1, [(1, 1), (3, 2), (5, 4)], 2, (True, True)),
    (4, 1, [], 2, (False, True)),   # widths 6, 7, 7 defer a window from diagonal 6 to 9
    (3, 1, [], 2, (True, False)),   # and from 5 to 9
]


@pytest.mark.parametrize("case", _WINDOW_CASES)
def test_windowed_traceback_brute_force(case, oracle_lib):
    """The windowed traceback of getPosteriorProbsWithBanding (impl/pairwiseAligner.c:1319-1410)
    against explicit path enumeration, with exact log-sum-exp: the oracle's windows
    follow the schedule rule (:1326-1331, 1350); every band match cell of every
    posterior diagonal is reported once, with the posterior of the paths that end on
    its window's top diagonal weighted by the end prior (:1335-1337) -- the
    ragged-end prior only at the end of the matrix -- over their sum (1e-9 before
    the reference's floor(p * 1e7)); the total is the last window's sum.  Small
    minDiagsBetweenTraceBack / traceBackDiagonals put several windows in a
    13-diagonal matrix, full-band and banded (where the width rule defers a window)."""
    min_diags, tb, anchors, expansion, ragged = case
    m = models.synthetic_gaussian_model("ACGT", 3, seed=6)
    rng = np.random.default_rng(31 + min_diags * 7 + tb)
    ref = "".join(rng.choice(list("ACGT"), 9))
    kids = models.kmer_ids(ref, "ACGT", 3)
    ev = np.zeros((6, 4))
    for j in range(6):
        tt = m.match[kids[min(j, len(kids) - 1)]]
        ev[j, 0] = tt[0] + rng.normal(0, tt[1])
        ev[j, 1] = rng.wald(tt[2], tt[4])
    om = oracle.Model(oracle.GAUSS, m.alphabet, 3, m.transitions, m.match, m.gapy, m.gapx)
    exact_em = lambda kmer, e, match: oracle.emission(om, kmer, e, match)
    wins, post, logz = _windowed_brute_force(m, ref, ev, anchors, expansion, min_diags, tb, ragged[0], ragged[1],
                                             exact_em)
    assert len(wins) >= 4
    if expansion == 2 and not anchors:
        assert any(D - T > min_diags for T, D, _, _ in wins)  # the width rule deferred a window
    params = oracle.Params(threshold=0.0, diagonal_expansion=expansion, min_diags_between_traceback=min_diags,
                           traceback_diagonals=tb)
    oracle.set_exact_logadd(True)
    try:
        (pairs, st), rec = oracle.record_windows(oracle.get_aligned_pairs, om, ref, ev, anchors, params, *ragged)
    finally:
        oracle.set_exact_logadd(False)
    assert [tuple(int(v) for v in w) for w in rec] == wins
    assert st["n_windows"] == len(wins)
    assert abs(st["log_likelihood"] - logz) < 1e-9
    got = [((int(x), int(y)), p / 1e7) for p, x, y in pairs]
    assert len(got) == len(set(c for c, _ in got)) == len(post)
    for c, p in got:
        want = min(post[c], 1.0)
        assert want - 1e-7 - 1e-9 <= p <= want + 1e-9, (c, p, want)


def test_window_schedule_invariants(oracle_lib):
    """Multi-window read: the oracle's schedule runs (asserting every diagonal gets
    exactly one posterior, impl/pairwiseAligner.c:1411), windows follow the
    minDiags/width rule, and posteriors agree with a single-window run away from the
    window edges."""
    m = models.synthetic_gaussian_model("ACGT", 6, seed=1)
    from signalalign_amd import synth
    r = synth.single_read(m, 1500, seed=3)
    om = oracle.Model(oracle.GAUSS_DESCALED, m.alphabet, 6, m.transitions, m.match, m.gapy, m.gapx,
                      r["scale"], r["shift"], 1.0)
    p1, s1 = oracle.get_aligned_pairs(om, r["ref"], r["events"], r["anchors"], oracle.Params(diagonal_expansion=20))
    N = len(r["ref"]) - 5 + len(r["events"])
    assert s1["n_windows"] == N // 1000 + 1 or s1["n_windows"] >= 2
    p2, s2 = oracle.get_aligned_pairs(om, r["ref"], r["events"], r["anchors"],
                                      oracle.Params(diagonal_expansion=20, min_diags_between_traceback=10 ** 9))
    assert s2["n_windows"] == 1
    a = {(x, y): p for p, x, y in p1}
    b = {(x, y): p for p, x, y in p2}
    common = set(a) & set(b)
    assert len(common) > 0.95 * len(b)
    diffs = np.array([abs(a[k] - b[k]) / 1e7 for k in common])
    assert np.median(diffs) < 1e-3
    assert abs(s1["log_likelihood"] - s2["log_likelihood"]) / abs(s2["log_likelihood"]) < 1e-6


# --------------------------------------------------------------------------
# HDP density query
# --------------------------------------------------------------------------

def test_spline_invariants(oracle_lib):
    """impl/hdp_math_utils.c:402-496: the Hermite spline interpolates the knots,
    is C2 (the slopes solve the tridiagonal system), reproduces linear data exactly
    and extrapolates linearly."""
    x = oracle.linspace(30.0, 130.0, 50)
    assert x[0] == 30.0 and x[-1] == 130.0
    y = np.exp(-0.5 * ((x - 80) / 7.0) ** 2)
    k = oracle.spline_knot_slopes(x, y)
    assert np.allclose(models.spline_knot_slopes_rows(x, y)[0], k, rtol=0, atol=1e-15)
    h = x[1] - x[0]
    for i in range(1, len(x) - 1):
        lhs = k[i - 1] / h + 2 * k[i] * (2 / h) + k[i + 1] / h
        rhs = 3 * ((y[i] - y[i - 1]) / h ** 2 + (y[i + 1] - y[i]) / h ** 2)
        assert abs(lhs - rhs) < 1e-9
    assert abs(2 * k[0] + k[1] - 3 * (y[1] - y[0]) / h) < 1e-9
    for i in range(0, len(x), 7):
        assert abs(oracle.grid_spline_interp(x[i], x, y, k) - y[i]) < 1e-12
    lin = 2.0 * x + 1.0
    kl = oracle.spline_knot_slopes(x, lin)
    assert np.allclose(kl, 2.0, atol=1e-9)
    for q in [31.3, 77.7, 129.9, 10.0, 150.0]:
        assert abs(oracle.grid_spline_interp(q, x, lin, kl) - (2 * q + 1)) < 1e-8


def test_hdp_emission_closed_form(oracle_lib):
    """impl/stateMachine.c:468-495: log((1/var) * max(0, spline(descaled mean))) with the
    k-mer's observed-ancestor row; unobserved k-mers read their parent's grid."""
    m = models.synthetic_hdp_model("ACGT", 2, seed=8, grid_length=300)
    rows = m.kmer_row.copy()
    rows[5] = 3  # k-mer 5 unobserved: falls back to DP row 3
    match = np.zeros((16, 5))
    match[:, 0] = m.level_mean
    om = oracle.Model(oracle.HDP, m.alphabet, 2, m.transitions, match, match, None, 1.02, 1.5, 1.2, m.grid,
                      m.density, m.slopes, rows)
    for kmer in ["AC", "CC", "TT"]:
        kid = oracle.kmer_id(kmer, "ACGT")
        mu = m.level_mean[kid]
        for ev in [55.0, 80.0, 101.0, 20.0]:
            u = (ev + 1.2 * mu - 1.02 * mu - 1.5) / 1.2
            r = rows[kid]
            dens = max(0.0, oracle.grid_spline_interp(u, m.grid, m.density[r], m.slopes[r]))
            want = math.log(dens / 1.2) if dens > 0 else -math.inf
            got = oracle.emission(om, kmer, (ev, 1.0), True)
            assert got == want or abs(got - want) < 1e-12
            assert oracle.emission(om, kmer, (ev, 1.0), False) == got
    assert abs(models.expected_value(m.grid, m.density[0]) - oracle.hdp_expected_value(m.grid, m.density[0])) < 1e-12
