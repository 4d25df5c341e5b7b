"""HIP path vs the FP64 oracle on identical seeded inputs (GPU tests).

Bar (FP64 drop-in): same aligned-pair set except pairs within 1e-3 of the
threshold, posteriors within 1e-4 absolute, total log-likelihood within 1e-6
relative, and the reference's list order.  FP32 results: log-likelihood within
1e-6 relative, posteriors within P_TOL_FP32 with at most 1 in 500 beyond 1e-4
(tests/_parity.py states why); their errors are printed.
"""
import numpy as np
import pytest

from signalalign_amd import models, native, synth
from tests import _parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["reg1", "reg4", "reg4b", "lds"], autouse=True)
def kernel_path(request, monkeypatch):
    """Every case runs through each sweep: the register-resident kernels with one
    wavefront per forward sweep, with several (synchronised every 8th diagonal through
    halo lanes, and -- reg4b -- every diagonal), and the LDS-ring kernels (the fallback
    for bands wider than a register window)."""
    if request.param == "lds":
        monkeypatch.setenv("SA_KERNELS", "lds")
    else:
        monkeypatch.setenv("SA_KERNELS", "reg")
        monkeypatch.setenv("SA_FWD_WAVES", request.param[3])
        monkeypatch.setenv("SA_HALO", "0" if request.param.endswith("b") else "1")
    return request.param


@pytest.fixture(scope="module")
def g6():
    return synth.gaussian_model("ACGT", 6, seed=1)


def _check(ctx, m, r, params, precision, p_tol=P.P_TOL, ll_rel=P.LL_REL, ragged=(True, True), kind=None):
    dm = ctx.model(m, kind)
    res, _ = ctx.align_batch(dm, [P.native_read(r, *ragged)], params, precision)
    ref, st = P.run_oracle(m, r, params, *ragged, kind=kind)
    out = P.compare(res[0].pairs, ref, params.threshold, p_tol=p_tol)
    out["ll_rel"] = P.rel(res[0].log_likelihood, st["log_likelihood"])
    out["cells"] = res[0].cells
    assert res[0].cells == st["cells"]
    assert res[0].n_windows == st["n_windows"]
    assert out["ll_rel"] <= ll_rel, out
    print(out)
    return out


def test_c1_fp64(gpu_ctx, g6):
    r = synth.single_read(g6, 2000, seed=11)
    _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=20), native.SA_FP64)


def test_c1_fp32(gpu_ctx, g6):
    r = synth.single_read(g6, 2000, seed=11)
    _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=20), native.SA_FP32, p_tol=P.P_TOL_FP32, ll_rel=P.LL_REL_FP32)


def test_reference_5mer_kat_on_gpu(gpu_ctx):
    """tests/stateMachineTests.c:565-692 inputs through the drop-in: 7 pairs at
    threshold 0.2 at (0,0) (1,1) (2,2) (3,3) (5,4) (6,5) (8,6)."""
    import os
    m = models.load_model_file(os.path.join(os.path.dirname(__file__), "golden", "testModelR9_5mer_acgt_template.model"))
    from tests.test_oracle_kats import KAT5_EVENTS, KAT5_REF
    dm = gpu_ctx.model(m, native.SA_EMIT_GAUSS)
    for prec in (native.SA_FP64, native.SA_FP32):
        res, _ = gpu_ctx.align_batch(dm, [native.Read(KAT5_REF, KAT5_EVENTS, [], 1.0, 0.0, 1.0, False, False)],
                                     native.Params(threshold=0.2, diagonal_expansion=2), prec)
        got = sorted((int(x), int(y)) for _, x, y in res[0].pairs)
        assert got == [(0, 0), (1, 1), (2, 2), (3, 3), (5, 4), (6, 5), (8, 6)]


def test_batch_mixed_lengths_fp64(gpu_ctx, g6):
    reads = [synth.single_read(g6, n, seed=100 + i) for i, n in enumerate([60, 300, 900, 1500, 2500, 4000])]
    params = native.Params(diagonal_expansion=20)
    dm = gpu_ctx.model(g6)
    res, _ = gpu_ctx.align_batch(dm, [P.native_read(r) for r in reads], params, native.SA_FP64)
    for r, g in zip(reads, res):
        ref, st = P.run_oracle(g6, r, params)
        P.compare(g.pairs, ref, params.threshold)
        assert P.rel(g.log_likelihood, st["log_likelihood"]) <= P.LL_REL


@pytest.mark.parametrize("ragged", [(False, False), (True, False), (False, True)])
def test_ragged_ends(gpu_ctx, g6, ragged):
    r = synth.single_read(g6, 1200, seed=21)
    _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=20), native.SA_FP64, ragged=ragged)


def test_no_anchors_full_band(gpu_ctx, g6):
    r = synth.single_read(g6, 150, seed=22)
    r["anchors"] = np.zeros((0, 2), dtype=np.int64)
    _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=20), native.SA_FP64)


@pytest.mark.parametrize("n", [700, 1400])
def test_full_band_wider_than_lds(gpu_ctx, g6, n):
    """No anchors on a longer read: the band is the whole matrix, diagonals of up to
    about n cells, wider than any register window; the LDS sweeps keep their rings in
    per-workgroup HBM slices where LDS is too small."""
    r = synth.single_read(g6, n, seed=90 + n)
    r["anchors"] = np.zeros((0, 2), dtype=np.int64)
    _check(gpu_ctx, g6, r, native.Params(), native.SA_FP64)
    _check(gpu_ctx, g6, r, native.Params(), native.SA_FP32, p_tol=P.P_TOL_FP32, ll_rel=P.LL_REL_FP32)


def test_split_blocks(gpu_ctx, g6):
    """A small split_matrix_bigger_than_this makes every anchor gap a split
    (getSplitPoints, impl/pairwiseAligner.c:1706-1757)."""
    r = synth.single_read(g6, 1500, seed=23)
    params = native.Params(diagonal_expansion=20, split_matrix_bigger_than_this=50 * 50)
    out = _check(gpu_ctx, g6, r, params, native.SA_FP64)
    assert out["n_ref"] > 0


@pytest.mark.parametrize("expansion", [2, 20, 50, 100, 200, 300])
def test_window_sizes(gpu_ctx, g6, expansion):
    """Band expansions that need every register window size (64 to 512 positions)."""
    r = synth.single_read(g6, 2500, seed=40 + expansion)
    _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=expansion), native.SA_FP64)


def test_window_sizes_fp32(gpu_ctx, g6):
    for expansion in (20, 100, 300):
        r = synth.single_read(g6, 2500, seed=60 + expansion)
        _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=expansion), native.SA_FP32, p_tol=P.P_TOL_FP32, ll_rel=P.LL_REL_FP32)


def _window_classes(r, m, expansion):
    """Register-window class (cells per lane: 1, 2, 3; 4 = wider) of each aligned block
    of 64 diagonals, forward and backward spans, from the oracle's band."""
    import oracle
    lX, lY = len(r["ref"]) - m.kmer_length + 1, len(r["events"])
    L, R = oracle.band(r["anchors"], lX, lY, expansion)
    d = np.arange(lX + lY + 1)
    lo = np.trunc((d + L) / 2).astype(np.int64)
    hi = np.trunc((d + R) / 2).astype(np.int64)
    lmin = lo.copy()
    lmin[1:] = np.minimum(lmin[1:], lo[:-1])
    lmin[2:] = np.minimum(lmin[2:], lo[:-2])
    hmax = hi.copy()
    hmax[:-1] = np.maximum(hmax[:-1], hi[1:])
    hmax[:-2] = np.maximum(hmax[:-2], hi[2:])
    span = np.maximum(hi - lmin + 1, hmax - lo + 1)
    cls = np.select([span <= 64, span <= 127, span <= 190], [1, 2, 3], 4)
    return [int(cls[b:b + 64].max()) for b in range(0, len(cls), 64)]


@pytest.mark.parametrize("expansion,gaps", [(10, [(500, 600), (1800, 1850)]), (20, [(500, 650), (1800, 1875)]),
                                             (36, [(500, 600), (1800, 1850)])])
def test_window_width_switching(gpu_ctx, g6, expansion, gaps):
    """Anchor-free stretches widen the band there, so one sweep runs its blocks of 64
    diagonals at different window widths (1, 2 or 3 cells per lane), switching both
    ways through LDS; FP64 against the oracle and FP32 to its bar."""
    r = synth.single_read(g6, 2600, seed=70 + expansion)
    keep = [a for a in r["anchors"] if not any(g0 <= a[1] < g1 for g0, g1 in gaps)]
    r["anchors"] = np.array(keep, dtype=np.int64).reshape(-1, 2)
    cls = _window_classes(r, g6, expansion)
    assert max(cls) <= 3 and len(set(cls)) >= 2, cls  # the sweeps switch width (a class 4 would pin it)
    _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=expansion), native.SA_FP64)
    _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=expansion), native.SA_FP32, p_tol=P.P_TOL_FP32, ll_rel=P.LL_REL_FP32)


def test_wide_expansion_and_threshold(gpu_ctx, g6):
    r = synth.single_read(g6, 3000, seed=24)
    _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=100, threshold=0.05), native.SA_FP64)


def test_general_gaussian_tables(gpu_ctx, g6):
    """Extra-event and skip tables that are not the .model loader's (sd x 1.75, one
    skip value) take the general emission path."""
    m = models.GaussianModel(g6.alphabet, 6, g6.transitions, g6.match, g6.gapy.copy(), g6.gapx.copy())
    m.gapy[:, 4] *= 1.3
    m.gapy[:, 0] += 0.7
    m.gapx[::7] = np.log(0.05)
    r = synth.single_read(g6, 1500, seed=28)
    _check(gpu_ctx, m, r, native.Params(diagonal_expansion=20), native.SA_FP64)
    _check(gpu_ctx, m, r, native.Params(diagonal_expansion=20), native.SA_FP32, p_tol=P.P_TOL_FP32, ll_rel=P.LL_REL_FP32)


def test_undescaled_emission(gpu_ctx, g6):
    r = synth.single_read(g6, 800, seed=25)
    r["scale"], r["shift"] = 1.0, 0.0
    import oracle
    _check(gpu_ctx, g6, r, native.Params(diagonal_expansion=20), native.SA_FP64, kind=native.SA_EMIT_GAUSS)


def test_acegt_methylated(gpu_ctx):
    m = synth.gaussian_model("ACEGT", 6, seed=5)
    rng = synth.rng_for(31, 0, 0)
    fwd, _ = synth.random_reference(1500, rng, methyl_cpg=0.3)
    r = synth.simulate_read(m, fwd, synth.rng_for(31, 0, 1))
    _check(gpu_ctx, m, r, native.Params(diagonal_expansion=50), native.SA_FP64)


def test_hdp_fp64_and_fp32(gpu_ctx):
    m = synth.hdp_model("ACEGT", 4, seed=7, grid_length=400)
    rng = synth.rng_for(41, 0, 0)
    fwd, _ = synth.random_reference(1200, rng, methyl_cpg=0.3)
    r = synth.simulate_read(m, fwd, synth.rng_for(41, 0, 1))
    _check(gpu_ctx, m, r, native.Params(diagonal_expansion=50), native.SA_FP64)
    _check(gpu_ctx, m, r, native.Params(diagonal_expansion=50), native.SA_FP32, p_tol=P.P_TOL_FP32, ll_rel=P.LL_REL_FP32)


def test_empty_and_short_inputs(gpu_ctx, g6):
    dm = gpu_ctx.model(g6)
    reads = [native.Read("ACG", np.zeros((0, 4)), []),                    # no k-mers, no events
             native.Read("ACGTACGTAC", np.zeros((0, 4)), []),             # k-mers only
             native.Read("AC", np.array([[80.0, 1.0, 0.01, 0.0]] * 3), [])]  # events only
    res, _ = gpu_ctx.align_batch(dm, reads, native.Params(), native.SA_FP64)
    for g in res:
        assert len(g.pairs) == 0


def test_bad_kmer_is_an_error(gpu_ctx, g6):
    dm = gpu_ctx.model(g6)
    r = synth.single_read(g6, 300, seed=26)
    bad = r["ref"][:100] + b"N" + r["ref"][101:]
    with pytest.raises(native.SignalAlignError) as e:
        gpu_ctx.align(dm, native.Read(bad, r["events"], r["anchors"], r["scale"], r["shift"]))
    assert e.value.code == -4


def test_bad_anchors_is_an_error(gpu_ctx, g6):
    dm = gpu_ctx.model(g6)
    r = synth.single_read(g6, 300, seed=27)
    anc = np.array([[10, 10], [5, 20]])
    with pytest.raises(native.SignalAlignError) as e:
        gpu_ctx.align(dm, native.Read(r["ref"], r["events"], anc, r["scale"], r["shift"]))
    assert e.value.code == -5
