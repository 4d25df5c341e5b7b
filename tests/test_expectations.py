"""EM expectations (getExpectationsUsingAnchors with diagonalCalculation_Expectations,
impl/pairwiseAligner.c:1255-1275, 1983-2003, and the ContinuousPairHmm accumulators,
impl/continuousHmm.c:84-170): the oracle pinned by path enumeration, and the GPU
path held to the oracle.
"""
import math

import numpy as np
import pytest

import oracle
from signalalign_amd import models


def _enumerate(m, ref, ev, ragged_left, ragged_right, exact_emission, descale):
    """Every 3-state path from (0,0) to (lX,lY) (match steps emit (k-mer x, event y),
    gapX steps log 0.1, gapY steps (k-mer x-1, event y) with x >= 1), weighted by its
    posterior.  Returns (log total, expected transition counts (3,3) from the start
    state on, expected match count per k-mer, expected sum of descaled event means
    per k-mer)."""
    k = m.kmer_length
    lX, lY = len(ref) - k + 1, len(ev)
    t = m.transitions
    trans = {(0, 0): t[0], (1, 0): t[1], (2, 0): t[2], (0, 1): t[3], (0, 2): t[4], (1, 1): t[5], (2, 2): t[6],
             (2, 1): t[7], (1, 2): -math.inf}
    start = [-math.inf, 0.0, 0.0] if ragged_left else [0.0, -math.inf, -math.inf]
    end = ([(t[3] + t[4]) / 2, t[5], t[6]] if ragged_right else [t[0], t[1], t[2]])
    kid = models.kmer_ids(ref, m.alphabet, k)
    em = {}
    for x in range(lX):
        for y in range(lY):
            em[(x, y, 1)] = exact_emission(ref[x:x + k], ev[y], True)
            em[(x, y, 0)] = exact_emission(ref[x:x + k], ev[y], False)
    paths = []

    def rec(x, y, s, score, steps):
        if x == lX and y == lY:
            paths.append((score + end[s], list(steps)))
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
            steps.append((s, ns, x, y))
            rec(nx, ny, ns, score + e + tp, steps)
            steps.pop()

    for s0 in range(3):
        if start[s0] > -math.inf:
            rec(0, 0, s0, start[s0], [])
    sc = np.array([s for s, _ in paths])
    mx = sc.max()
    total = mx + math.log(np.exp(sc - mx).sum())
    tr = np.zeros((3, 3))
    n_k = {}
    s_k = {}
    for s, steps in paths:
        w = math.exp(s - total)
        for f, to, x, y in steps:
            tr[f, to] += w
            if to == 0:
                kk = int(kid[x])
                n_k[kk] = n_k.get(kk, 0.0) + w
                s_k[kk] = s_k.get(kk, 0.0) + w * descale(kk, ev[y][0])
    return total, tr, n_k, s_k


@pytest.mark.parametrize("ragged", [(False, False), (True, True), (True, False)])
def test_expectations_equal_path_enumeration(ragged, oracle_lib):
    """With the exact log-sum-exp, a one-window read's transition expectations are
    the expected transition counts over all paths, the likelihood is the total times
    the number of posterior diagonals, and each k-mer's posterior sum and
    posterior-weighted descaled event sum are the expected match count and sum over
    all paths (1e-9).  Pseudocounts 0.001 as hmmContinuous_getExpectationsHmm sets."""
    m = models.synthetic_gaussian_model("ACGT", 3, seed=4)
    rng = np.random.default_rng(21)
    ref = "".join(rng.choice(list("ACGT"), 7))
    kids = models.kmer_ids(ref, "ACGT", 3)
    ev = np.zeros((4, 4))
    for j in range(4):
        t = m.match[kids[min(j, len(kids) - 1)]]
        ev[j, 0] = (t[0] + rng.normal(0, t[1])) * 1.1 + 3.0
        ev[j, 1] = rng.wald(t[2], t[4])
    scale, shift, var = 1.1, 3.0, 1.2
    om = oracle.Model(oracle.GAUSS_DESCALED, m.alphabet, 3, m.transitions, m.match, m.gapy, m.gapx, scale, shift,
                      var)
    exact_em = lambda kmer, e, match: oracle.emission(om, kmer, e, match)  # noqa: E731
    descale = lambda kk, x: (x + var * m.match[kk][0] - scale * m.match[kk][0] - shift) / var  # noqa: E731
    total, tr, n_k, s_k = _enumerate(m, ref, ev, ragged[0], ragged[1], exact_em, descale)
    oracle.set_exact_logadd(True)
    try:
        ex = oracle.get_expectations(om, ref, ev, [], oracle.Params(diagonal_expansion=20), *ragged)
    finally:
        oracle.set_exact_logadd(False)
    N = len(ref) - 3 + 1 + len(ev)
    assert abs(ex["likelihood"] - N * total) < 1e-9 * N
    assert np.allclose(ex["transitions"] - 0.001, tr, rtol=0, atol=1e-9)
    for kk in range(len(ex["posteriors"])):
        assert abs(ex["posteriors"][kk] - 0.001 - n_k.get(kk, 0.0)) < 1e-9
        assert abs(ex["event_exp"][kk, 0] - s_k.get(kk, 0.0)) < 1e-9 * max(1.0, abs(s_k.get(kk, 0.0)))
    # a k-mer is observed when a match transition into it was evaluated
    assert set(np.flatnonzero(ex["observed"])) == set(int(k) for k in kids)


def test_expectation_records_follow_the_call_order(oracle_lib):
    """The per-cell match-transition records of a multi-window read: diagonals in
    the windowed traceback's order (each window from its top posterior diagonal
    down), cells in ascending x, and the running variance term equal to replaying
    them (down to the smallest posterior)."""
    m = models.synthetic_gaussian_model("ACGT", 6, seed=1)
    from signalalign_amd import synth
    r = synth.single_read(m, 1500, seed=3)
    om = oracle.Model(oracle.GAUSS_DESCALED, m.alphabet, 6, m.transitions, m.match, m.gapy, m.gapx,
                      r["scale"], r["shift"], 1.0)
    (ex, wins) = oracle.record_windows(oracle.get_expectations, om, r["ref"], r["events"], r["anchors"],
                                       oracle.Params(diagonal_expansion=20), rec_threshold=0.0,
                                       rec_cap=400000)
    rec = ex["records"]
    assert len(wins) > 1
    d = rec[:, 0] + rec[:, 1] + 2  # the cell's anti-diagonal
    order = []
    for tb_to, _, tb_from, _ in wins:
        order += list(range(int(tb_from), int(tb_to), -1))
    # the diagonals appear in the schedule's order, each once, x ascending inside
    seen = [int(v) for i, v in enumerate(d) if i == 0 or d[i - 1] != v]
    assert seen == [e for e in order if e in set(seen)]
    for e in set(seen):
        xs = rec[d == e, 0]
        assert np.all(np.diff(xs) > 0)
    # the first posterior diagonal of every window after the first has no match records
    for tb_to, _, tb_from, _ in wins[1:]:
        assert int(tb_to) + 1 not in set(seen)
    # replaying the records reproduces the event expectations
    k = 6
    kid = models.kmer_ids(r["ref"], m.alphabet, k)
    s1 = np.zeros(len(ex["posteriors"]))
    s0 = np.full(len(ex["posteriors"]), 0.001)
    s2 = np.zeros(len(ex["posteriors"]))
    for x, y, p1, p2, p3 in rec:
        kk = int(kid[int(x)])
        mu = m.match[kk][0]
        xd = (r["events"][int(y)][0] + mu - r["scale"] * mu - r["shift"]) / 1.0
        for p in (p1, p2, p3):
            s1[kk] += p * xd
            s0[kk] += p
            u = s1[kk] / s0[kk]
            s2[kk] += p * (xd - u) * (xd - u)
    assert np.allclose(s0, ex["posteriors"], rtol=1e-12, atol=0)
    assert np.allclose(s1, ex["event_exp"][:, 0], rtol=1e-12, atol=1e-12)
    assert np.allclose(s2, ex["event_exp"][:, 1], rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# the GPU path (sa_expect) against the oracle
# ---------------------------------------------------------------------------
REC_T = 0.0  # every cell whose match transitions add to the sums


def _gpu_vs_oracle(gpu_ctx, m, read, params, kind=None):
    from signalalign_amd import native
    from tests import _parity as P
    om = P.oracle_model(m, read, kind)
    want = oracle.get_expectations(om, read["ref"], read["events"], read["anchors"],
                                   oracle.Params(params.threshold, params.min_diags_between_traceback,
                                                 params.traceback_diagonals, params.diagonal_expansion,
                                                 params.split_matrix_bigger_than_this),
                                   rec_threshold=REC_T)
    dm = gpu_ctx.model(m) if kind is None else gpu_ctx.model(m, kind)
    got = gpu_ctx.expect(dm, P.native_read(read), params, record_threshold=REC_T)
    assert np.allclose(got["transitions"] + 0.001, want["transitions"], rtol=1e-9, atol=1e-12)
    assert abs(got["likelihood"] - want["likelihood"]) <= 1e-12 * abs(want["likelihood"])
    assert np.array_equal(got["observed"], want["observed"])
    assert np.allclose(got["match_posterior"] + 0.001, want["posteriors"], rtol=1e-9, atol=1e-12)
    assert np.allclose(got["match_event_sum"], want["event_exp"][:, 0], rtol=1e-9, atol=1e-9)
    gr, wr = got["records"], want["records"]
    assert gr.shape == wr.shape and len(gr) > 0
    assert np.array_equal(gr[:, :2], wr[:, :2])
    assert np.allclose(gr[:, 2:], wr[:, 2:], rtol=1e-9, atol=1e-15)
    # the running variance term replayed from the records (order-dependent), against
    # the oracle's
    kid = models.kmer_ids(read["ref"], m.alphabet, m.kmer_length)
    nk = len(got["observed"])
    mu = m.level_mean if isinstance(m, models.HdpModel) else m.match[:, 0]
    s0, s1, s2 = np.full(nk, 0.001), np.zeros(nk), np.zeros(nk)
    for x, y, p1, p2, p3 in gr:
        kk = int(kid[int(x)])
        e = read["events"][int(y)][0]
        xd = (e + read["var"] * mu[kk] - read["scale"] * mu[kk] - read["shift"]) / read["var"]
        for p in (p1, p2, p3):
            s1[kk] += p * xd
            s0[kk] += p
            u = s1[kk] / s0[kk]
            s2[kk] += p * (xd - u) * (xd - u)
    assert np.allclose(s2, want["event_exp"][:, 1], rtol=1e-9, atol=1e-12)
    return got, want


@pytest.mark.gpu
@pytest.mark.parametrize("kernels", ["register", "lds"])
@pytest.mark.parametrize("n,expansion,seed", [(300, 20, 1), (2500, 50, 2), (4000, 20, 3)])
def test_gpu_expectations_match_oracle(gpu_ctx, oracle_lib, monkeypatch, n, expansion, seed, kernels):
    """Synthetic reads on a 6-mer Gaussian model (descaled emissions), one to several
    traceback windows: transition and k-mer expectations, likelihood, observed
    k-mers and the ordered match records equal the oracle's (FP64), with the forward
    pass on the register sweeps (the default) and on the LDS sweeps."""
    from signalalign_amd import native, synth
    if kernels == "lds":
        monkeypatch.setenv("SA_KERNELS", "lds")
    m = models.synthetic_gaussian_model("ACGT", 6, seed=7)
    r = synth.single_read(m, n, seed=seed)
    print(f"expect n={n} e={expansion} {kernels}", flush=True)
    got, want = _gpu_vs_oracle(gpu_ctx, m, r, native.Params(diagonal_expansion=expansion))
    print(len(got["records"]), got["likelihood"])


@pytest.mark.gpu
def test_gpu_expectations_split_and_hdp(gpu_ctx, oracle_lib):
    """A read cut into sub-alignments (split matrix 100^2 with sparse anchors) and a
    read on an HDP density grid (level means from the grid)."""
    from signalalign_amd import native, synth
    m = models.synthetic_gaussian_model("ACGT", 6, seed=7)
    r = synth.single_read(m, 1200, seed=5)
    r = dict(r, anchors=r["anchors"][::6])
    print("expect split", flush=True)
    _gpu_vs_oracle(gpu_ctx, m, r, native.Params(diagonal_expansion=20, split_matrix_bigger_than_this=100 * 100))
    h = models.synthetic_hdp_model("ACGT", 5, seed=2)
    r2 = synth.single_read(h, 800, seed=6)
    print("expect hdp", flush=True)
    _gpu_vs_oracle(gpu_ctx, h, r2, native.Params(diagonal_expansion=20))


@pytest.mark.gpu
def test_gpu_expectations_batch(gpu_ctx, oracle_lib):
    """Several reads in one sa_expect_batch equal the oracle read by read; records at
    threshold 0.5 (window slots sized for a few per diagonal, refilled when they
    overflow) equal the oracle's records at that threshold."""
    from signalalign_amd import native, synth
    from tests import _parity as P
    m = models.synthetic_gaussian_model("ACGT", 6, seed=7)
    reads = [synth.single_read(m, n, seed=40 + i) for i, n in enumerate([200, 900, 1700, 3100])]
    params = native.Params(diagonal_expansion=30)
    dm = gpu_ctx.model(m)
    print("expect batch", flush=True)
    for thr in (0.5, 0.0):
        got = gpu_ctx.expect_batch(dm, [P.native_read(r) for r in reads], params, record_threshold=thr)
        for r, g in zip(reads, got):
            om = P.oracle_model(m, r)
            w = oracle.get_expectations(om, r["ref"], r["events"], r["anchors"],
                                        oracle.Params(diagonal_expansion=30), rec_threshold=thr)
            assert np.allclose(g["transitions"] + 0.001, w["transitions"], rtol=1e-9, atol=1e-12)
            assert abs(g["likelihood"] - w["likelihood"]) <= 1e-12 * abs(w["likelihood"])
            assert np.allclose(g["match_posterior"] + 0.001, w["posteriors"], rtol=1e-9, atol=1e-12)
            assert np.array_equal(g["records"][:, :2], w["records"][:, :2])
            assert np.allclose(g["records"][:, 2:], w["records"][:, 2:], rtol=1e-9, atol=1e-15)
        print(thr, [len(g["records"]) for g in got])
