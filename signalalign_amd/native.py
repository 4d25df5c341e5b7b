"""ctypes binding of libsignalalign_amd.so (include/signalalign_amd.h).

This is the Python face of the drop-in: `Aligner.get_aligned_pairs` takes the
arguments of getAlignedPairsUsingAnchors (impl/pairwiseAligner.c:1871-1899) in
flat form and returns the reference's (floor(p*1e7), x, y) tuples in the
reference's list order.  The library must be built (``__graft_entry__.build()``
or ``make -C signalalign_amd/csrc``); importing this module without it raises.
"""
import ctypes
import os

import numpy as np

from . import models as _models

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsignalalign_amd.so")

SA_FP64, SA_FP32 = 0, 1
# status codes (include/signalalign_amd.h)
SA_OK, SA_ERR_INVALID_ARGUMENT, SA_ERR_HIP, SA_ERR_OUT_OF_MEMORY = 0, -1, -2, -3
SA_ERR_BAD_KMER, SA_ERR_ANCHORS, SA_ERR_UNSUPPORTED, SA_ERR_NO_DEVICE = -4, -5, -6, -7
SA_ERR_CAPACITY = -8
SA_EMIT_GAUSS_DESCALED, SA_EMIT_GAUSS, SA_EMIT_HDP = 0, 1, 2

_ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "out of device memory", -4: "k-mer outside alphabet",
           -5: "invalid anchors", -6: "unsupported", -7: "no device", -8: "fixed capacity exceeded"}

_F64P = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)


class _Params(ctypes.Structure):
    _fields_ = [("threshold", ctypes.c_double), ("min_diags_between_traceback", ctypes.c_int64),
                ("traceback_diagonals", ctypes.c_int64), ("diagonal_expansion", ctypes.c_int64),
                ("split_matrix_bigger_than_this", ctypes.c_int64)]


class _ModelDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("kmer_length", ctypes.c_int32), ("alphabet", ctypes.c_char_p),
                ("transitions", ctypes.c_double * 9), ("match", _F64P), ("gapy", _F64P), ("gapx", _F64P),
                ("grid_length", ctypes.c_int64), ("grid", _F64P), ("n_rows", ctypes.c_int64),
                ("density", _F64P), ("slopes", _F64P), ("kmer_row", _I64P), ("level_mean", _F64P)]


class _Read(ctypes.Structure):
    _fields_ = [("ref", ctypes.c_char_p), ("ref_length", ctypes.c_int64), ("events", _F64P),
                ("n_events", ctypes.c_int64), ("event_stride", ctypes.c_int64), ("anchors", _I64P),
                ("n_anchors", ctypes.c_int64), ("scale", ctypes.c_double), ("shift", ctypes.c_double),
                ("var", ctypes.c_double), ("ragged_left", ctypes.c_int32), ("ragged_right", ctypes.c_int32),
                ("degenerate_options", ctypes.c_char_p)]


class _Pair(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int32), ("x", ctypes.c_int32), ("y", ctypes.c_int32)]


class _MatchRecord(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("p", ctypes.c_double * 3)]


class _Expectations(ctypes.Structure):
    _fields_ = [("transitions", ctypes.c_double * 9), ("likelihood", ctypes.c_double), ("n_kmers", ctypes.c_int64),
                ("match_posterior", ctypes.POINTER(ctypes.c_double)),
                ("match_event_sum", ctypes.POINTER(ctypes.c_double)),
                ("observed", ctypes.POINTER(ctypes.c_int8)), ("records", ctypes.POINTER(_MatchRecord)),
                ("n_records", ctypes.c_int64)]


class _ReadResult(ctypes.Structure):
    _fields_ = [("pairs", ctypes.POINTER(_Pair)), ("n_pairs", ctypes.c_int64), ("log_likelihood", ctypes.c_double),
                ("cells", ctypes.c_int64), ("n_windows", ctypes.c_int64), ("n_subalignments", ctypes.c_int64),
                ("paths", ctypes.POINTER(ctypes.c_int32))]


class _Timing(ctypes.Structure):
    _fields_ = [("prep_ms", ctypes.c_double), ("forward_ms", ctypes.c_double), ("backward_ms", ctypes.c_double),
                ("collect_ms", ctypes.c_double), ("total_ms", ctypes.c_double), ("cells", ctypes.c_int64),
                ("forward_launches", ctypes.c_int64), ("backward_launches", ctypes.c_int64),
                ("fb_ms", ctypes.c_double), ("emission_path", ctypes.c_int32), ("replayed", ctypes.c_int32),
                ("device_start_ms", ctypes.c_double), ("device_end_ms", ctypes.c_double),
                ("emission_ms", ctypes.c_double)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("lanes", ctypes.c_int32), ("max_wave_reads", ctypes.c_int32),
                ("length_ratio", ctypes.c_double), ("batch_delay_ms", ctypes.c_double)]


class _ReqTiming(ctypes.Structure):
    _fields_ = [("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("wave", ctypes.c_int64), ("wave_reads", ctypes.c_int32), ("pad", ctypes.c_int32)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("waves", ctypes.c_int64), ("reads", ctypes.c_int64), ("cells", ctypes.c_int64),
                ("max_wave_reads", ctypes.c_int64), ("device_ms", ctypes.c_double), ("now_s", ctypes.c_double),
                ("pending", ctypes.c_int64), ("inflight_bytes", ctypes.c_int64), ("fill_ms", ctypes.c_double),
                ("run_ms", ctypes.c_double)]


class _Tuning(ctypes.Structure):
    _fields_ = [("use_graph", ctypes.c_int32), ("waves_per_simd", ctypes.c_int32), ("long_factor", ctypes.c_double),
                ("stream_split", ctypes.c_double), ("persist_waves", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class _Launch(ctypes.Structure):
    _fields_ = [("direction", ctypes.c_int32), ("cpt", ctypes.c_int32), ("waves", ctypes.c_int32),
                ("stream", ctypes.c_int32), ("items", ctypes.c_int64), ("cells", ctypes.c_int64),
                ("start_ms", ctypes.c_double), ("ms", ctypes.c_double), ("sync_every", ctypes.c_int32),
                ("persistent", ctypes.c_int32)]



_lib = None


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(make -C signalalign_amd/csrc); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.sa_params_default.argtypes = [ctypes.POINTER(_Params)]
    L.sa_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.sa_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.sa_context_destroy.argtypes = [vp]
    L.sa_model_create.argtypes = [vp, ctypes.POINTER(_ModelDesc), ctypes.POINTER(vp)]
    L.sa_model_destroy.argtypes = [vp]
    L.sa_align.argtypes = [vp, vp, ctypes.POINTER(_Read), ctypes.POINTER(_Params), ctypes.c_int, ctypes.POINTER(vp)]
    L.sa_align_noalloc.argtypes = [vp, vp, ctypes.POINTER(_Read), ctypes.POINTER(_Params), ctypes.c_int,
                                   ctypes.POINTER(_Pair), ctypes.c_int64, _I64P, _F64P]
    L.sa_align_batch.argtypes = [vp, vp, ctypes.POINTER(_Read), ctypes.c_int64, ctypes.POINTER(_Params),
                                 ctypes.c_int, ctypes.POINTER(vp)]
    L.sa_batch_create.argtypes = [vp, vp, ctypes.POINTER(_Read), ctypes.c_int64, ctypes.POINTER(_Params),
                                  ctypes.POINTER(vp)]
    L.sa_batch_run.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    L.sa_batch_set_tuning.argtypes = [vp, ctypes.POINTER(_Tuning)]
    L.sa_batch_get_tuning.argtypes = [vp, ctypes.POINTER(_Tuning)]
    L.sa_batch_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_Tuning), ctypes.POINTER(ctypes.c_double)]
    L.sa_batch_refill.argtypes = [vp, vp, ctypes.POINTER(_Read), ctypes.c_int64, ctypes.POINTER(_Params)]
    L.sa_context_create_ex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.sa_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(_ServerConfig), ctypes.POINTER(vp)]
    L.sa_server_submit.argtypes = [vp, vp, ctypes.POINTER(_Read), ctypes.POINTER(_Params), ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int64)]
    L.sa_server_submit_batch.argtypes = [vp, vp, ctypes.POINTER(_Read), ctypes.c_int64, ctypes.POINTER(_Params),
                                         ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.sa_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(_ReadResult),
                                 ctypes.POINTER(_ReqTiming)]
    L.sa_server_release.argtypes = [vp, ctypes.c_int64]
    L.sa_server_get_stats.argtypes = [vp, ctypes.POINTER(_ServerStats)]
    L.sa_server_destroy.argtypes = [vp]
    L.sa_batch_destroy.argtypes = [vp]
    L.sa_result_n_reads.argtypes = [vp]
    L.sa_result_n_reads.restype = ctypes.c_int64
    L.sa_result_read.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(_ReadResult)]
    L.sa_result_timing.argtypes = [vp, ctypes.POINTER(_Timing)]
    L.sa_result_launches.argtypes = [vp, ctypes.POINTER(_Launch), ctypes.c_int64]
    L.sa_result_launches.restype = ctypes.c_int64
    L.sa_result_free.argtypes = [vp]
    L.sa_expect.argtypes = [vp, vp, ctypes.POINTER(_Read), ctypes.POINTER(_Params), ctypes.c_double,
                            ctypes.POINTER(vp)]
    L.sa_expect_get.argtypes = [vp, ctypes.POINTER(_Expectations)]
    L.sa_expect_batch.argtypes = [vp, vp, ctypes.POINTER(_Read), ctypes.c_int64, ctypes.POINTER(_Params),
                                  ctypes.c_double, ctypes.POINTER(vp)]
    L.sa_expect_read.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(_Expectations)]
    L.sa_expect_timing.argtypes = [vp, ctypes.POINTER(_Timing)]
    L.sa_expect_free.argtypes = [vp]
    L.sa_last_error.argtypes = [vp]
    L.sa_last_error.restype = ctypes.c_char_p
    I32P = ctypes.POINTER(ctypes.c_int32)
    L.sa_band_geometry.argtypes = [_I64P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, I32P, I32P]
    L.sa_window_schedule.argtypes = [I32P, I32P, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(_Params), I32P,
                                     ctypes.c_int64]
    L.sa_window_schedule.restype = ctypes.c_int64
    L.sa_split_points.argtypes = [_I64P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(_Params),
                                  ctypes.c_int32, ctypes.c_int32, _I64P, ctypes.c_int64]
    L.sa_split_points.restype = ctypes.c_int64
    _lib = L
    return L


def band_geometry(anchors, lX, lY, expansion):
    """Per-diagonal x ranges (xlo, xhi) of the band, computed by the library on the host."""
    a = np.ascontiguousarray(np.asarray(anchors, dtype=np.int64).reshape(-1, 2))
    lo = np.zeros(lX + lY + 1, dtype=np.int32)
    hi = np.zeros(lX + lY + 1, dtype=np.int32)
    I32P = ctypes.POINTER(ctypes.c_int32)
    rc = load().sa_band_geometry(a.ctypes.data_as(_I64P), len(a), lX, lY, expansion, lo.ctypes.data_as(I32P),
                                 hi.ctypes.data_as(I32P))
    if rc != 0:
        raise SignalAlignError(rc, "sa_band_geometry")
    return lo, hi


def window_schedule(xlo, xhi, lX, lY, params=None):
    """Traceback windows (tracedBackTo, start diagonal, tracedBackFrom, last) as rows."""
    params = params or Params()
    I32P = ctypes.POINTER(ctypes.c_int32)
    lo = np.ascontiguousarray(xlo, dtype=np.int32)
    hi = np.ascontiguousarray(xhi, dtype=np.int32)
    cap = (lX + lY) // 10 + 16
    out = np.zeros((cap, 4), dtype=np.int32)
    ps = params._struct()
    n = load().sa_window_schedule(lo.ctypes.data_as(I32P), hi.ctypes.data_as(I32P), lX, lY, ctypes.byref(ps),
                                  out.ctypes.data_as(I32P), cap)
    return out[:n].copy()


def split_points(anchors, lX, lY, params=None, ragged_left=True, ragged_right=True):
    params = params or Params()
    a = np.ascontiguousarray(np.asarray(anchors, dtype=np.int64).reshape(-1, 2))
    ps = params._struct()
    n = load().sa_split_points(a.ctypes.data_as(_I64P), len(a), lX, lY, ctypes.byref(ps), int(ragged_left),
                               int(ragged_right), None, 0)
    if n < 0:
        raise SignalAlignError(int(n), "sa_split_points")
    out = np.zeros((max(n, 1), 4), dtype=np.int64)
    load().sa_split_points(a.ctypes.data_as(_I64P), len(a), lX, lY, ctypes.byref(ps), int(ragged_left),
                           int(ragged_right), out.ctypes.data_as(_I64P), n)
    return out[:n]


def device_count():
    n = ctypes.c_int(0)
    load().sa_device_count(ctypes.byref(n))
    return n.value


class SignalAlignError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params:
    """PairwiseAlignmentParameters fields used by the path (defaults of
    pairwiseAlignmentBandingParameters_construct, impl/pairwiseAligner.c:1843-1856)."""

    def __init__(self, threshold=0.01, min_diags_between_traceback=1000, traceback_diagonals=40,
                 diagonal_expansion=20, split_matrix_bigger_than_this=3000 * 3000):
        self.threshold = threshold
        self.min_diags_between_traceback = min_diags_between_traceback
        self.traceback_diagonals = traceback_diagonals
        self.diagonal_expansion = diagonal_expansion
        self.split_matrix_bigger_than_this = split_matrix_bigger_than_this

    def _struct(self):
        return _Params(self.threshold, self.min_diags_between_traceback, self.traceback_diagonals,
                       self.diagonal_expansion, self.split_matrix_bigger_than_this)


class Read:
    """One strand of one read: reference nucleotides, event records (mean, noise, ...),
    0-based anchors (k-mer index, event index), descaling and ragged ends."""

    def __init__(self, ref, events, anchors, scale=1.0, shift=0.0, var=1.0, ragged_left=True, ragged_right=True,
                 degenerate_options=None):
        self.ref = ref.encode() if isinstance(ref, str) else bytes(ref)
        ev = np.asarray(events, dtype=np.float64)
        if ev.ndim != 2 or ev.shape[1] < 2:
            raise ValueError("events must be (n, >=2)")
        self.events = np.ascontiguousarray(ev)
        self.anchors = np.ascontiguousarray(np.asarray(anchors, dtype=np.int64).reshape(-1, 2))
        self.scale, self.shift, self.var = float(scale), float(shift), float(var)
        self.ragged_left, self.ragged_right = bool(ragged_left), bool(ragged_right)
        self.degenerate_options = degenerate_options.encode() if degenerate_options else None

    def _struct(self):
        return _Read(self.ref, len(self.ref), self.events.ctypes.data_as(_F64P), len(self.events),
                     self.events.shape[1], self.anchors.ctypes.data_as(_I64P), len(self.anchors), self.scale,
                     self.shift, self.var, int(self.ragged_left), int(self.ragged_right), self.degenerate_options)


class ReadResult:
    def __init__(self, pairs, log_likelihood, cells, n_windows, n_subalignments, paths=None):
        self.pairs = pairs              # (n, 3) int64 [p_int, x, y]
        self.log_likelihood = log_likelihood
        self.cells = cells
        self.n_windows = n_windows
        self.n_subalignments = n_subalignments
        self.paths = paths              # (n,) int32 path per pair, degenerate references only


class Context:
    """A device context (one HIP stream) with resident models."""

    def __init__(self, device=0):
        L = load()
        h = ctypes.c_void_p()
        rc = L.sa_context_create(device, ctypes.byref(h))
        if rc != 0:
            raise SignalAlignError(rc, "sa_context_create")
        self._h = h
        self._L = L

    def _check(self, rc, what):
        if rc != 0:
            raise SignalAlignError(rc, f"{what}: {self._L.sa_last_error(self._h).decode(errors='replace')}")

    def close(self):
        if self._h:
            self._L.sa_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def model(self, m, emission=None):
        """Upload a models.GaussianModel (emission SA_EMIT_GAUSS_DESCALED by default,
        or SA_EMIT_GAUSS) or a models.HdpModel."""
        return DeviceModel(self, m, emission)

    def align(self, model, read, params=None, precision=SA_FP64):
        return self.align_batch(model, [read], params, precision)[0][0]

    def align_noalloc(self, model, read, params=None, precision=SA_FP64, pairs_cap=None):
        """sa_align_noalloc: the plain per-read call with no allocation in the library
        (intermediates in its fixed device scratch).  Returns (pairs (n, 3) int64,
        log-likelihood).  With pairs_cap too small it raises SA_ERR_CAPACITY."""
        cap = 8 * (len(read.ref) + len(read.events)) + 1024 if pairs_cap is None else int(pairs_cap)
        buf = (_Pair * max(cap, 1))()
        n, ll = ctypes.c_int64(0), ctypes.c_double(0.0)
        rs, ps = read._struct(), (params or Params())._struct()
        self._check(self._L.sa_align_noalloc(self._h, model._h, ctypes.byref(rs), ctypes.byref(ps), precision, buf,
                                             cap, ctypes.byref(n), ctypes.byref(ll)), "sa_align_noalloc")
        pairs = np.frombuffer(buf, dtype=np.int32, count=3 * n.value).reshape(-1, 3).astype(np.int64)
        return pairs, ll.value

    def expect(self, model, read, params=None, record_threshold=1.0):
        """EM expectations of one read (sa_expect, FP64).  Returns a dict:
        transitions (3, 3) [from, to] without pseudocounts, likelihood,
        match_posterior / match_event_sum (n_kmers,), observed (bool), records (n, 5)
        float64: k-mer position, event, p(M->M), p(X->M), p(Y->M) in call order."""
        return self.expect_batch(model, [read], params, record_threshold)[0]

    def expect_batch(self, model, reads, params=None, record_threshold=1.0, timing=None):
        """sa_expect_batch: one dict (as expect) per read; `timing`, a dict, receives
        the run's device timing."""
        params = params or Params()
        arr = (_Read * max(1, len(reads)))()
        for i, r in enumerate(reads):
            arr[i] = r._struct()
        ps = params._struct()
        res = ctypes.c_void_p()
        L = self._L
        self._check(L.sa_expect_batch(self._h, model._h, arr, len(reads), ctypes.byref(ps), float(record_threshold),
                                      ctypes.byref(res)), "sa_expect_batch")
        try:
            if timing is not None:
                t = _Timing()
                L.sa_expect_timing(res, ctypes.byref(t))
                timing.update({k: getattr(t, k) for k, _ in _Timing._fields_})
            out = []
            e = _Expectations()
            for i in range(len(reads)):
                L.sa_expect_read(res, i, ctypes.byref(e))
                nk = e.n_kmers
                d = {"transitions": np.array(list(e.transitions)).reshape(3, 3), "likelihood": e.likelihood,
                     "match_posterior": np.ctypeslib.as_array(e.match_posterior, shape=(nk,)).copy(),
                     "match_event_sum": np.ctypeslib.as_array(e.match_event_sum, shape=(nk,)).copy(),
                     "observed": np.ctypeslib.as_array(e.observed, shape=(nk,)).astype(bool)}
                rec = np.zeros((e.n_records, 5))
                if e.n_records:
                    raw = np.ctypeslib.as_array(ctypes.cast(e.records, ctypes.POINTER(ctypes.c_uint8)),
                                                shape=(e.n_records * ctypes.sizeof(_MatchRecord),))
                    st = np.frombuffer(raw.tobytes(),
                                       dtype=np.dtype([("x", "<i4"), ("y", "<i4"), ("p", "<f8", 3)]))
                    rec[:, 0] = st["x"]
                    rec[:, 1] = st["y"]
                    rec[:, 2:] = st["p"]
                d["records"] = rec
                out.append(d)
            return out
        finally:
            L.sa_expect_free(res)

    def align_batch(self, model, reads, params=None, precision=SA_FP64):
        """sa_align_batch (the context refills one batch call after call).  Returns
        ([ReadResult], timing dict)."""
        reads = list(reads)
        arr = (_Read * max(1, len(reads)))()
        for i, r in enumerate(reads):
            arr[i] = r._struct()
        ps = (params or Params())._struct()
        res = ctypes.c_void_p()
        self._check(self._L.sa_align_batch(self._h, model._h, arr, len(reads), ctypes.byref(ps), precision,
                                           ctypes.byref(res)), "sa_align_batch")
        return _take_result(self._L, res)


class DeviceModel:
    def __init__(self, ctx, m, emission=None):
        self.ctx = ctx
        self.src = m
        d = _ModelDesc()
        keep = []
        if isinstance(m, _models.HdpModel):
            d.kind = SA_EMIT_HDP
            for name in ("grid", "density", "slopes", "level_mean"):
                a = np.ascontiguousarray(getattr(m, name), dtype=np.float64)
                keep.append(a)
                setattr(d, name, a.ctypes.data_as(_F64P))
            kr = np.ascontiguousarray(m.kmer_row, dtype=np.int64)
            keep.append(kr)
            d.kmer_row = kr.ctypes.data_as(_I64P)
            d.grid_length = len(m.grid)
            d.n_rows = m.density.shape[0]
        else:
            d.kind = SA_EMIT_GAUSS_DESCALED if emission is None else emission
            for name in ("match", "gapy", "gapx"):
                a = np.ascontiguousarray(getattr(m, name), dtype=np.float64)
                keep.append(a)
                setattr(d, name, a.ctypes.data_as(_F64P))
        d.kmer_length = m.kmer_length
        d.alphabet = m.alphabet.encode()
        for i in range(9):
            d.transitions[i] = float(m.transitions[i])
        h = ctypes.c_void_p()
        ctx._check(ctx._L.sa_model_create(ctx._h, ctypes.byref(d), ctypes.byref(h)), "sa_model_create")
        self._h = h
        self.kind = d.kind
        self.kmer_length = m.kmer_length

    def close(self):
        if self._h:
            self.ctx._L.sa_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _take_result(L, res, fetch=True):
    """([ReadResult], timing dict) of an sa_result, which is freed."""
    try:
        t = _Timing()
        L.sa_result_timing(res, ctypes.byref(t))
        timing = {k: getattr(t, k) for k, _ in _Timing._fields_}
        n = L.sa_result_launches(res, None, 0)
        arr = (_Launch * max(n, 1))()
        L.sa_result_launches(res, arr, n)
        timing["launches"] = [{k: getattr(arr[i], k) for k, _ in _Launch._fields_} for i in range(n)]
        out = []
        if fetch:
            rr = _ReadResult()
            for i in range(L.sa_result_n_reads(res)):
                L.sa_result_read(res, i, ctypes.byref(rr))
                if rr.n_pairs:
                    buf = np.ctypeslib.as_array(ctypes.cast(rr.pairs, ctypes.POINTER(ctypes.c_int32)),
                                                shape=(rr.n_pairs * 3,))
                    pairs = buf.reshape(-1, 3).astype(np.int64)
                else:
                    pairs = np.zeros((0, 3), dtype=np.int64)
                paths = (np.ctypeslib.as_array(rr.paths, shape=(rr.n_pairs,)).copy()
                         if rr.n_pairs and bool(rr.paths) else None)
                out.append(ReadResult(pairs, rr.log_likelihood, rr.cells, rr.n_windows, rr.n_subalignments,
                                      paths))
        return out, timing
    finally:
        L.sa_result_free(res)


class Batch:
    """Reads uploaded to HBM once, runnable any number of times (sa_batch_*)."""

    def __init__(self, ctx, model, reads, params=None):
        self.ctx = ctx
        self.model = model
        self.reads = list(reads)
        params = params or Params()
        arr = (_Read * max(1, len(self.reads)))()
        for i, r in enumerate(self.reads):
            arr[i] = r._struct()
        ps = params._struct()
        h = ctypes.c_void_p()
        ctx._check(ctx._L.sa_batch_create(ctx._h, model._h, arr, len(self.reads), ctypes.byref(ps),
                                          ctypes.byref(h)), "sa_batch_create")
        self._h = h

    def run(self, precision=SA_FP64, fetch=True):
        L = self.ctx._L
        res = ctypes.c_void_p()
        self.ctx._check(L.sa_batch_run(self._h, precision, ctypes.byref(res)), "sa_batch_run")
        return _take_result(L, res, fetch)

    def tuning(self):
        t = _Tuning()
        self.ctx._L.sa_batch_get_tuning(self._h, ctypes.byref(t))
        return {k: getattr(t, k) for k, _ in _Tuning._fields_}

    def set_tuning(self, **kw):
        t = self.tuning()
        t.update(kw)
        self.ctx._check(self.ctx._L.sa_batch_set_tuning(self._h, ctypes.byref(_Tuning(**t))), "sa_batch_set_tuning")

    def autotune(self, precision=SA_FP32):
        """Times candidate tunings, keeps the fastest; returns (tuning dict, device ms)."""
        t, ms = _Tuning(), ctypes.c_double()
        self.ctx._check(self.ctx._L.sa_batch_autotune(self._h, precision, ctypes.byref(t), ctypes.byref(ms)),
                        "sa_batch_autotune")
        return {k: getattr(t, k) for k, _ in _Tuning._fields_}, ms.value

    def close(self):
        if self._h:
            self.ctx._L.sa_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


SA_PENDING = 1


class Server:
    """Request loop with continuous batching (sa_server_*): submit reads one at a
    time, collect results as waves finish."""

    def __init__(self, device=0, hbm_budget_bytes=0, lanes=0, max_wave_reads=0, length_ratio=0.0,
                 batch_delay_ms=0.0):
        L = load()
        cfg = _ServerConfig(int(hbm_budget_bytes), int(lanes), int(max_wave_reads), float(length_ratio),
                            float(batch_delay_ms))
        h = ctypes.c_void_p()
        rc = L.sa_server_create(device, ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise SignalAlignError(rc, "sa_server_create")
        self._h, self._L = h, L

    def submit_struct(self, model, read_struct, params_struct, precision=SA_FP32):
        t = ctypes.c_int64()
        rc = self._L.sa_server_submit(self._h, model._h, ctypes.byref(read_struct), ctypes.byref(params_struct),
                                      precision, ctypes.byref(t))
        if rc != 0:
            raise SignalAlignError(rc, "sa_server_submit")
        return t.value

    def submit_structs(self, model, read_structs, params_struct, precision=SA_FP32):
        """Queue many reads at once (sa_server_submit_batch): one ticket per read."""
        n = len(read_structs)
        arr = (_Read * n)(*read_structs)
        t = (ctypes.c_int64 * max(n, 1))()
        rc = self._L.sa_server_submit_batch(self._h, model._h, arr, n, ctypes.byref(params_struct), precision, t)
        if rc != 0:
            raise SignalAlignError(rc, "sa_server_submit_batch")
        return list(t[:n])

    def submit(self, model, read, params=None, precision=SA_FP32):
        return self.submit_struct(model, read._struct(), (params or Params())._struct(), precision)

    def wait(self, ticket, timeout=60.0, fetch=True):
        """(ReadResult or None, timing dict), or None while the request is pending."""
        rr, tm = _ReadResult(), _ReqTiming()
        rc = self._L.sa_server_wait(self._h, ticket, float(timeout), ctypes.byref(rr), ctypes.byref(tm))
        if rc == SA_PENDING:
            return None
        timing = {k: getattr(tm, k) for k, _ in _ReqTiming._fields_ if k != "pad"}
        if rc != 0:
            raise SignalAlignError(rc, f"request {ticket}")
        res = None
        if fetch:
            if rr.n_pairs:
                buf = np.ctypeslib.as_array(ctypes.cast(rr.pairs, ctypes.POINTER(ctypes.c_int32)),
                                            shape=(rr.n_pairs * 3,))
                pairs = buf.reshape(-1, 3).astype(np.int64)
            else:
                pairs = np.zeros((0, 3), dtype=np.int64)
            paths = (np.ctypeslib.as_array(rr.paths, shape=(rr.n_pairs,)).copy()
                     if rr.n_pairs and bool(rr.paths) else None)
            res = ReadResult(pairs, rr.log_likelihood, rr.cells, rr.n_windows, rr.n_subalignments, paths)
        else:
            res = ReadResult(None, rr.log_likelihood, rr.cells, rr.n_windows, rr.n_subalignments)
        return res, timing

    def release(self, ticket):
        self._L.sa_server_release(self._h, ticket)

    def stats(self):
        st = _ServerStats()
        self._L.sa_server_get_stats(self._h, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}

    def close(self):
        if self._h:
            self._L.sa_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
