"""sa_align_noalloc (the plain per-read call with no allocation: intermediates in the
library's fixed device scratch) against sa_align on the same reads (GPU tests).

Both run the same kernels on the same shapes, so the pairs, their order and the
log-likelihood must be equal bit for bit; every case also meets the hot path's bar
against the FP64 oracle.  Each case calls the entry point twice in a row, the second
time with a smaller read, so the second call runs over the scratch contents the
first left behind.
"""
import numpy as np
import pytest

from signalalign_amd import native, synth
from tests import _parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def g6():
    return synth.gaussian_model("ACGT", 6, seed=1)


def _both(ctx, dm, m, r, params, precision, label):
    print(f"noalloc {label}: {len(r['ref'])} nt, {len(r['events'])} events, precision {precision}", flush=True)
    rd = P.native_read(r)
    got, ll = ctx.align_noalloc(dm, rd, params, precision)
    want, _ = ctx.align_batch(dm, [rd], params, precision)
    assert np.array_equal(got, want[0].pairs), label
    assert ll == want[0].log_likelihood, (label, ll, want[0].log_likelihood)
    ref, st = P.run_oracle(m, r, params)
    if precision == native.SA_FP64:
        P.compare(got, ref, params.threshold)
        assert P.rel(ll, st["log_likelihood"]) <= P.LL_REL
    else:
        P.compare(got, ref, params.threshold, p_tol=P.P_TOL_FP32)
        assert P.rel(ll, st["log_likelihood"]) <= P.LL_REL_FP32
    return len(got)


@pytest.mark.parametrize("precision", [native.SA_FP64, native.SA_FP32])
def test_noalloc_equals_sa_align_two_calls(ctx, g6, precision):
    dm = ctx.model(g6)
    params = native.Params(diagonal_expansion=50)
    big = synth.single_read(g6, 9000, seed=501)
    small = synth.single_read(g6, 1200, seed=502)
    n1 = _both(ctx, dm, g6, big, params, precision, "9 kb then")
    n2 = _both(ctx, dm, g6, small, params, precision, "1.2 kb")
    assert n1 > n2 > 0


def test_noalloc_split_blocks_and_lds_sweeps(ctx, g6):
    """Split blocks (a small split_matrix_bigger_than_this), then a read without
    anchors whose whole-matrix band is wider than any register window (LDS sweeps,
    their rings in scratch where LDS is too small), then a short read."""
    dm = ctx.model(g6)
    r = synth.single_read(g6, 1500, seed=503)
    _both(ctx, dm, g6, r, native.Params(diagonal_expansion=20, split_matrix_bigger_than_this=50 * 50),
          native.SA_FP64, "split blocks")
    r2 = synth.single_read(g6, 1100, seed=504)
    r2["anchors"] = np.zeros((0, 2), dtype=np.int64)
    _both(ctx, dm, g6, r2, native.Params(), native.SA_FP64, "full band")
    _both(ctx, dm, g6, synth.single_read(g6, 300, seed=505), native.Params(), native.SA_FP64, "short")


def test_noalloc_hdp_read(ctx):
    m = synth.hdp_model("ACEGT", 4, seed=7, grid_length=400)
    rng = synth.rng_for(41, 0, 0)
    fwd, _ = synth.random_reference(2500, rng, methyl_cpg=0.3)
    r = synth.simulate_read(m, fwd, synth.rng_for(41, 0, 1))
    dm = ctx.model(m)
    params = native.Params(diagonal_expansion=50)
    _both(ctx, dm, m, r, params, native.SA_FP32, "HDP")
    _both(ctx, dm, m, r, params, native.SA_FP64, "HDP")


def test_noalloc_long_read_fits(ctx, g6):
    """One 50 kb read at band expansion 100 (about 17M band cells), FP32 and FP64."""
    dm = ctx.model(g6)
    r = synth.single_read(g6, 50000, seed=2074)
    _both(ctx, dm, g6, r, native.Params(diagonal_expansion=100), native.SA_FP32, "50 kb")
    _both(ctx, dm, g6, r, native.Params(diagonal_expansion=100), native.SA_FP64, "50 kb")


def test_noalloc_capacity_errors(ctx, g6):
    """Beyond the fixed scratch, or with a pair buffer too small, the call fails with
    SA_ERR_CAPACITY and a message; the next call still works."""
    dm = ctx.model(g6)
    r = synth.single_read(g6, 7000, seed=506)
    r["anchors"] = np.zeros((0, 2), dtype=np.int64)  # whole-matrix band: about 65M cells, 520 MB in FP64
    print("noalloc capacity: whole-matrix 7 kb read", flush=True)
    with pytest.raises(native.SignalAlignError) as e:
        ctx.align_noalloc(dm, P.native_read(r), native.Params(split_matrix_bigger_than_this=10 ** 12),
                          native.SA_FP64)
    assert e.value.code == native.SA_ERR_CAPACITY and "scratch" in str(e.value)
    small = synth.single_read(g6, 800, seed=507)
    print("noalloc capacity: pair buffer of 5", flush=True)
    with pytest.raises(native.SignalAlignError) as e:
        ctx.align_noalloc(dm, P.native_read(small), native.Params(), native.SA_FP64, pairs_cap=5)
    assert e.value.code == native.SA_ERR_CAPACITY and "pairs" in str(e.value)
    _both(ctx, dm, g6, small, native.Params(), native.SA_FP64, "after the errors")
