"""Batched runs replayed from a recorded HIP graph, and the launch autotuner: both
change how the work is launched, never the results."""
import ctypes

import numpy as np
import pytest

from signalalign_amd import native, synth
from tests import _parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu_ctx):
    m = synth.gaussian_model("ACEGT", 6, seed=5)
    reads = synth.batch_reads(m, 80, 3000, seed=401, methyl_cpg=0.3)
    return m, gpu_ctx.model(m), [P.native_read(r) for r in reads]


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.array_equal(x.pairs, y.pairs)
        assert x.log_likelihood == y.log_likelihood
        assert x.cells == y.cells


@pytest.mark.parametrize("precision", [native.SA_FP32, native.SA_FP64])
def test_graph_replays_match_eager(gpu_ctx, setup, precision):
    """Eager run, then runs with graphs on: the second records, the rest replay.
    Every replay equals the eager result, with other device work (a different
    batch on the same context) and host synchronisation between replays."""
    m, dm, reads = setup
    params = native.Params(diagonal_expansion=50)
    b = native.Batch(gpu_ctx, dm, reads, params)
    other = native.Batch(gpu_ctx, dm, reads[::3], params)
    try:
        b.set_tuning(use_graph=0)
        eager, _ = b.run(precision)
        b.set_tuning(use_graph=1)
        replays = 0
        for i in range(5):
            got, tm = b.run(precision)
            print(i, tm["replayed"], tm["total_ms"], tm["forward_ms"], tm["backward_ms"])
            _same(eager, got)
            replays += tm["replayed"]
            other.run(precision)
        assert replays >= 3
        assert tm["cells"] == sum(r.cells for r in eager)
    finally:
        other.close()
        b.close()


def test_autotune_keeps_results(gpu_ctx, setup):
    """The autotuner tries launch shapes (wavefronts per forward sweep, stream split,
    occupancy cap) and keeps the fastest; results are those of the default shape."""
    m, dm, reads = setup
    params = native.Params(diagonal_expansion=50)
    b = native.Batch(gpu_ctx, dm, reads, params)
    try:
        base, _ = b.run(native.SA_FP32)
        tuning, ms = b.autotune(native.SA_FP32)
        print(tuning, ms)
        assert ms > 0 and tuning["waves_per_simd"] in (0, 2, 3, 4)
        after, _ = b.run(native.SA_FP32)
        _same(base, after)
        after, tm = b.run(native.SA_FP32)
        _same(base, after)
    finally:
        b.close()



def _pairs_of(L, res):
    out = []
    rr = native._ReadResult()
    for i in range(L.sa_result_n_reads(res)):
        assert L.sa_result_read(res, i, ctypes.byref(rr)) == 0
        if rr.n_pairs:
            buf = np.ctypeslib.as_array(ctypes.cast(rr.pairs, ctypes.POINTER(ctypes.c_int32)),
                                        shape=(rr.n_pairs * 3,))
            out.append(buf.reshape(-1, 3).copy())
        else:
            out.append(np.zeros((0, 3), dtype=np.int32))
    return out


def test_large_results_stay_valid():
    """Large pair lists (at least 2^18 pairs) arrive in pinned buffers that results
    hand back to their context when freed: a result stays valid while its batch runs
    again, while a later result's buffer is reused, and after its context is
    destroyed."""
    m = synth.gaussian_model("ACEGT", 6, seed=6)
    reads = [P.native_read(r) for r in synth.batch_reads(m, 64, 6000, seed=402, methyl_cpg=0.3)]
    params = native.Params(diagonal_expansion=50)
    ctx = native.Context(0)
    L = ctx._L
    dm = ctx.model(m)
    b = native.Batch(ctx, dm, reads, params)
    res = [ctypes.c_void_p() for _ in range(3)]
    try:
        ctx._check(L.sa_batch_run(b._h, native.SA_FP32, ctypes.byref(res[0])), "run 0")
        first = _pairs_of(L, res[0])
        assert sum(len(p) for p in first) >= 1 << 18, "the batch must take the pinned path"
        ctx._check(L.sa_batch_run(b._h, native.SA_FP32, ctypes.byref(res[1])), "run 1")
        L.sa_result_free(res[1])  # its buffer goes back to the context
        res[1] = ctypes.c_void_p()
        ctx._check(L.sa_batch_run(b._h, native.SA_FP32, ctypes.byref(res[2])), "run 2")  # reuses it
        for a, c in zip(first, _pairs_of(L, res[2])):
            assert np.array_equal(a, c)
        for a, c in zip(first, _pairs_of(L, res[0])):  # untouched by the later runs
            assert np.array_equal(a, c)
        b.close()
        dm.close()
        ctx.close()
        for a, c in zip(first, _pairs_of(L, res[2])):  # outlives its context
            assert np.array_equal(a, c)
    finally:
        for r in res:
            if r:
                L.sa_result_free(r)
        b.close()
        dm.close()
        ctx.close()
