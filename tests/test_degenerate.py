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


# ---------------------------------------------------------------------------
# the GPU path (multi-path sweeps) against the oracle
# ---------------------------------------------------------------------------
def _with_x(ref, every=1):
    """The C of every CpG (every `every`-th one) becomes 'X'."""
    r = list(ref)
    n = 0
    for i in range(len(r) - 1):
        if r[i] == "C" and r[i + 1] == "G":
            if n % every == 0:
                r[i] = "X"
            n += 1
    return "".join(r)


def _compare_paths(got, gpaths, want, threshold, p_tol=1e-4, edge=1e-3):
    g = {(int(x), int(y), int(pa)): p / 1e7 for (p, x, y), pa in zip(got, gpaths)}
    w = {(int(x), int(y), int(pa)): p / 1e7 for p, x, y, pa in want}
    assert len(g) == len(got)
    only_g = [k for k in g if k not in w and abs(g[k] - threshold) > edge]
    only_w = [k for k in w if k not in g and abs(w[k] - threshold) > edge]
    common = [k for k in g if k in w]
    dmax = max((abs(g[k] - w[k]) for k in common), default=0.0)
    out = {"n_gpu": len(g), "n_ref": len(w), "only_gpu": len(only_g), "only_ref": len(only_w), "max_abs_dp": dmax}
    assert not only_g and not only_w, out
    assert dmax <= p_tol, out
    cs = set(common)
    og = [(int(x), int(y), int(pa)) for (p, x, y), pa in zip(got, gpaths) if (int(x), int(y), int(pa)) in cs]
    ow = [(int(x), int(y), int(pa)) for p, x, y, pa in want if (int(x), int(y), int(pa)) in cs]
    assert og == ow, "pair order differs from the reference list order"
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n,options,every,expansion,split", [
    (300, "ECO", 1, 20, None), (2500, "OEC", 1, 50, None), (1800, "EC", 2, 20, None), (1200, "CEO", 1, 20, 100)])
def test_gpu_degenerate_reads_match_oracle(gpu_ctx, oracle_lib, n, options, every, expansion, split):
    """Reads whose CpG cytosines are 'X' (2 or 3 options per site), one to several
    windows, with anchors, and cut into sub-alignments: the pairs, their paths,
    posteriors and order, and the log-likelihood equal the oracle's (FP64)."""
    from signalalign_amd import native, synth
    from tests import _parity as P
    m = models.synthetic_gaussian_model("ACEGOT", 6, seed=3)
    r = synth.single_read(m, n, seed=11 + n)
    r = dict(r, ref=_with_x(r["ref"].decode() if isinstance(r["ref"], bytes) else r["ref"], every))
    if split:
        r["anchors"] = r["anchors"][::6]
    params = native.Params(diagonal_expansion=expansion,
                           split_matrix_bigger_than_this=(split * split if split else 3000 * 3000))
    om = P.oracle_model(m, r)
    want, st = oracle.get_aligned_pairs(om, r["ref"], r["events"], r["anchors"],
                                        oracle.Params(params.threshold, params.min_diags_between_traceback,
                                                      params.traceback_diagonals, params.diagonal_expansion,
                                                      params.split_matrix_bigger_than_this), options=options)
    dm = gpu_ctx.model(m)
    read = native.Read(r["ref"], r["events"], r["anchors"], r["scale"], r["shift"], r["var"],
                       degenerate_options=options)
    print(f"degenerate n={n} {options} every={every} split={split}", flush=True)
    res, _ = gpu_ctx.align_batch(dm, [read], params, native.SA_FP64)
    out = _compare_paths(res[0].pairs, res[0].paths, want, params.threshold)
    assert P.rel(res[0].log_likelihood, st["log_likelihood"]) <= P.LL_REL
    assert int(np.max(want[:, 3])) > 0
    print(out, res[0].n_subalignments)


@pytest.mark.gpu
def test_gpu_degenerate_batch_mixed_and_refusals(gpu_ctx, oracle_lib):
    """A batch mixing a degenerate read and a plain one (the plain one's pairs all on
    path 0, equal to its single-path run); FP32 is refused for degenerate batches;
    an 'X' without options is a bad k-mer."""
    from signalalign_amd import native, synth
    from tests import _parity as P
    m = models.synthetic_gaussian_model("ACEGOT", 6, seed=3)
    r1 = synth.single_read(m, 700, seed=21)
    r2 = synth.single_read(m, 900, seed=22)
    ref1 = _with_x(r1["ref"].decode() if isinstance(r1["ref"], bytes) else r1["ref"])
    params = native.Params(diagonal_expansion=20)
    dm = gpu_ctx.model(m)
    reads = [native.Read(ref1, r1["events"], r1["anchors"], r1["scale"], r1["shift"], r1["var"],
                         degenerate_options="CEO"),
             native.Read(r2["ref"], r2["events"], r2["anchors"], r2["scale"], r2["shift"], r2["var"],
                         degenerate_options="CEO")]
    print("degenerate batch", flush=True)
    res, _ = gpu_ctx.align_batch(dm, reads, params, native.SA_FP64)
    plain, _ = gpu_ctx.align_batch(dm, [P.native_read(r2)], params, native.SA_FP64)
    # the multi-path sweeps and the single-path sweeps order the posterior's additions
    # differently; a posterior is truncated to 1e-7 steps, so one within an ulp of a
    # step (most often p = 1) can land one step apart
    got, want = res[1].pairs.astype(np.int64), plain[0].pairs.astype(np.int64)
    assert np.array_equal(got[:, 1:], want[:, 1:]) and np.all(res[1].paths == 0)
    assert np.max(np.abs(got[:, 0] - want[:, 0])) <= 1
    assert res[1].log_likelihood == plain[0].log_likelihood
    with pytest.raises(native.SignalAlignError):
        gpu_ctx.align_batch(dm, reads, params, native.SA_FP32)
    with pytest.raises(native.SignalAlignError):
        gpu_ctx.align_batch(dm, [native.Read(ref1, r1["events"], r1["anchors"], r1["scale"], r1["shift"],
                                             r1["var"])], params, native.SA_FP64)


# ---------------------------------------------------------------------------
# the reference's known answer on a real read
# ---------------------------------------------------------------------------
DEGENERATE_PAIRS = 7349


def _zymo_descaled(sub):
    """tests/stateMachineTests.c:896-947 (test_DegenerateNucleotides): the Zymo read's
    template events, the descaling machine, threshold 0.01, neither end ragged, and
    ZymoRef with every C replaced by `sub` (degenerate bases "CEO")."""
    from tests.test_real_read import _inputs, _machines
    r, ref, m = _inputs()
    name, mm, kind, read = [t for t in _machines(r, ref, m) if t[0] == "descaled"][0]
    return mm, kind, dict(read, ref=ref.replace("C", sub).encode())


def test_oracle_reproduces_the_reference_degenerate_count(oracle_lib):
    """The reference expects 1076 pairs with C->E and C->O, and 7349 (four-tuples,
    several paths per cell) with C->X.  The reference's run has lastz anchors, absent
    here: without anchors the band is the whole matrix, which gives the same counts."""
    from tests import _parity as P
    for sub, want in [("E", 1076), ("O", 1076), ("X", DEGENERATE_PAIRS)]:
        mm, kind, read = _zymo_descaled(sub)
        pairs, _ = oracle.get_aligned_pairs(P.oracle_model(mm, read, oracle.GAUSS_DESCALED), read["ref"],
                                            read["events"], read["anchors"], oracle.Params(0.01), False, False,
                                            options=("CEO" if sub == "X" else None))
        assert len(pairs) == want, (sub, len(pairs))
        if sub == "X":
            assert pairs.shape[1] == 4 and len({(x, y) for _, x, y, _ in pairs}) < len(pairs)


@pytest.mark.gpu
def test_gpu_reproduces_the_reference_degenerate_count(gpu_ctx, oracle_lib):
    """The same run on the GPU: 7349 pairs, equal to the oracle's (pairs, paths,
    order, posteriors) and its log-likelihood.  The full-matrix band with up to 3^6
    paths per column puts the rings in HBM and runs the strided sweeps."""
    from signalalign_amd import native
    from tests import _parity as P
    mm, kind, read = _zymo_descaled("X")
    params = native.Params(threshold=0.01)
    want, st = oracle.get_aligned_pairs(P.oracle_model(mm, read, oracle.GAUSS_DESCALED), read["ref"],
                                        read["events"], read["anchors"], oracle.Params(0.01), False, False,
                                        options="CEO")
    dm = gpu_ctx.model(mm, kind)
    rd = native.Read(read["ref"], read["events"], read["anchors"], read["scale"], read["shift"], read["var"],
                     False, False, degenerate_options="CEO")
    print("zymo degenerate", flush=True)
    res, tm = gpu_ctx.align_batch(dm, [rd], params, native.SA_FP64)
    out = _compare_paths(res[0].pairs, res[0].paths, want, params.threshold)
    assert len(res[0].pairs) == DEGENERATE_PAIRS, out
    assert P.rel(res[0].log_likelihood, st["log_likelihood"]) <= P.LL_REL
    # 7349 pairs over about 1.5k posterior diagonals overflow the first guess of pair
    # slots (a few per diagonal): the overflowing window ran a second backward pass
    assert tm["backward_launches"] == 2, tm
    print(out, tm)


@pytest.mark.gpu
def test_gpu_degenerate_two_groups_with_hbm_rings(gpu_ctx, oracle_lib):
    """The Zymo degenerate read (whole-matrix band, up to 729 paths per column: rings
    in HBM) batched with short degenerate reads: the short ones, under half the
    longest sweep, run forward and backward on the second stream while the long one's
    forward runs, with their ring slices after the long read's.  Every read equals the
    oracle (pairs, paths, order, log-likelihood)."""
    from signalalign_amd import native, synth
    from tests import _parity as P
    mm, kind, read = _zymo_descaled("X")
    dm = gpu_ctx.model(mm, kind)
    params = native.Params(threshold=0.01)
    reads = [dict(read, ragged=(False, False))]
    for i in range(3):
        r = synth.single_read(mm, 160 + 40 * i, seed=600 + i)
        r["ref"] = _with_x(r["ref"].decode() if isinstance(r["ref"], bytes) else r["ref"]).encode()
        reads.append(dict(r, ragged=(True, True)))
    nreads = [native.Read(r["ref"], r["events"], r["anchors"], r["scale"], r["shift"], r["var"], *r["ragged"],
                          degenerate_options="CEO") for r in reads]
    print("zymo degenerate with short reads", flush=True)
    res, tm = gpu_ctx.align_batch(dm, nreads, params, native.SA_FP64)
    assert tm["forward_launches"] == 2, tm
    for r, g in zip(reads, res):
        want, st = oracle.get_aligned_pairs(P.oracle_model(mm, r, oracle.GAUSS_DESCALED), r["ref"], r["events"],
                                            r["anchors"], oracle.Params(0.01), *r["ragged"], options="CEO")
        _compare_paths(g.pairs, g.paths, want, params.threshold)
        assert P.rel(g.log_likelihood, st["log_likelihood"]) <= P.LL_REL
    assert len(res[0].pairs) == DEGENERATE_PAIRS


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,options", [(0, "CE"), (1, "CEO")])
def test_compat_degenerate_alignment_matches_oracle(tmp_path, oracle_lib, dtype, options):
    """A reference with 'X' through sequence_constructReferenceKmerSequence (the
    DegenerateType's options) and getAlignedPairsUsingAnchors (C): the pairs, list
    order, log-likelihood and the path k-mer of every pair (the 'X' resolved to its
    path's option) equal the oracle's."""
    import subprocess
    from signalalign_amd import synth
    from tests import _parity as P
    from tests.test_compat import COMPAT_BIN, _write_input
    from signalalign_amd import native
    m = synth.gaussian_model("ACEGOT", 6, seed=4)
    mpath = str(tmp_path / "g6.model")
    models.write_model_file(m, mpath)
    r = synth.single_read(m, 1500, seed=91)
    r = dict(r, ref=_with_x(r["ref"].decode() if isinstance(r["ref"], bytes) else r["ref"]))
    params = native.Params(diagonal_expansion=20, split_matrix_bigger_than_this=60 * 60)
    inp, out = str(tmp_path / "in.txt"), str(tmp_path / "out.tsv")
    _write_input(inp, r, params, (True, True))
    print(f"compat degenerate {options}", flush=True)
    res = subprocess.run([COMPAT_BIN, "align", mpath, inp, out, str(dtype)], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    lines = open(out).read().splitlines()
    ll = float(lines[0])
    rows = [ln.split("\t") for ln in lines[1:]]
    want, st = oracle.get_aligned_pairs(P.oracle_model(m, r), r["ref"], r["events"], r["anchors"],
                                        oracle.Params(params.threshold, params.min_diags_between_traceback,
                                                      params.traceback_diagonals, params.diagonal_expansion,
                                                      params.split_matrix_bigger_than_this), options=options)
    got_paths = [oracle.path_kmers(r["ref"][int(b):int(b) + 6], options).index(km) for _, b, _, km in rows]
    _compare_paths([(int(a), int(b), int(c)) for a, b, c, _ in rows], got_paths, want, params.threshold)
    assert P.rel(ll, st["log_likelihood"]) <= P.LL_REL
    assert max(got_paths) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("n,options,expansion", [(900, "CE", 20), (2000, "EC", 20)])
def test_gpu_degenerate_hdp_reads_match_oracle(gpu_ctx, oracle_lib, n, options, expansion):
    """Degenerate references on an HDP density grid (the methylation-calling setup:
    ACEGT 5-mers, every CpG cytosine 'X'): pairs, paths, posteriors, order and
    log-likelihood equal the oracle's (FP64)."""
    from signalalign_amd import native, synth
    from tests import _parity as P
    h = models.synthetic_hdp_model("ACEGT", 5, seed=2)
    r = synth.single_read(h, n, seed=31 + n)
    r = dict(r, ref=_with_x(r["ref"].decode() if isinstance(r["ref"], bytes) else r["ref"]))
    params = native.Params(diagonal_expansion=expansion)
    want, st = oracle.get_aligned_pairs(P.oracle_model(h, r), r["ref"], r["events"], r["anchors"],
                                        oracle.Params(diagonal_expansion=expansion), options=options)
    dm = gpu_ctx.model(h)
    read = native.Read(r["ref"], r["events"], r["anchors"], r["scale"], r["shift"], r["var"],
                       degenerate_options=options)
    print(f"degenerate hdp n={n} {options}", flush=True)
    res, _ = gpu_ctx.align_batch(dm, [read], params, native.SA_FP64)
    out = _compare_paths(res[0].pairs, res[0].paths, want, params.threshold)
    assert P.rel(res[0].log_likelihood, st["log_likelihood"]) <= P.LL_REL
    assert int(np.max(want[:, 3])) > 0
    print(out)


@pytest.mark.gpu
@pytest.mark.parametrize("options,ragged", [("ACGT", (True, True)), ("CEO", (False, False)), ("CE", (False, True))])
def test_gpu_degenerate_options_and_ends(gpu_ctx, oracle_lib, options, ragged):
    """canonicalVariants (four options per site) and non-ragged ends: pairs, paths,
    posteriors, order and log-likelihood equal the oracle's (FP64)."""
    from signalalign_amd import native, synth
    from tests import _parity as P
    m = models.synthetic_gaussian_model("ACEGOT", 6, seed=3)
    r = synth.single_read(m, 1500, seed=77)
    r = dict(r, ref=_with_x(r["ref"].decode() if isinstance(r["ref"], bytes) else r["ref"]))
    params = native.Params(diagonal_expansion=30)
    want, st = oracle.get_aligned_pairs(P.oracle_model(m, r), r["ref"], r["events"], r["anchors"],
                                        oracle.Params(diagonal_expansion=30), *ragged, options=options)
    dm = gpu_ctx.model(m)
    read = native.Read(r["ref"], r["events"], r["anchors"], r["scale"], r["shift"], r["var"], *ragged,
                       degenerate_options=options)
    print(f"degenerate {options} ragged={ragged}", flush=True)
    res, _ = gpu_ctx.align_batch(dm, [read], params, native.SA_FP64)
    out = _compare_paths(res[0].pairs, res[0].paths, want, params.threshold)
    assert P.rel(res[0].log_likelihood, st["log_likelihood"]) <= P.LL_REL
    print(out)
