"""ctypes view of the FP64 CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
The oracle restates signalAlign's getAlignedPairsUsingAnchors path
(impl/pairwiseAligner.c:1871-1899 and the functions it calls); see oracle.c
for the file:line each routine follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

GAUSS_DESCALED, GAUSS, HDP = 0, 1, 2

_I64P = ctypes.POINTER(ctypes.c_int64)
_F64P = ctypes.POINTER(ctypes.c_double)


class _Model(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("kmer_length", ctypes.c_int32),
        ("alphabet_size", ctypes.c_int32),
        ("alphabet", ctypes.c_char * 16),
        ("trans", ctypes.c_double * 9),
        ("match", _F64P),
        ("gapy", _F64P),
        ("gapx", _F64P),
        ("scale", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("var", ctypes.c_double),
        ("grid_length", ctypes.c_int64),
        ("grid", _F64P),
        ("density", _F64P),
        ("slopes", _F64P),
        ("kmer_row", _I64P),
    ]


class _Params(ctypes.Structure):
    _fields_ = [
        ("threshold", ctypes.c_double),
        ("min_diags_between_traceback", ctypes.c_int64),
        ("traceback_diagonals", ctypes.c_int64),
        ("diagonal_expansion", ctypes.c_int64),
        ("split_matrix_bigger_than_this", ctypes.c_int64),
    ]


class _Pair(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int64), ("x", ctypes.c_int64), ("y", ctypes.c_int64), ("path", ctypes.c_int64)]


class _Expect(ctypes.Structure):
    _fields_ = [
        ("transitions", _F64P),
        ("likelihood", ctypes.c_double),
        ("event_exp", _F64P),
        ("posteriors", _F64P),
        ("observed", ctypes.POINTER(ctypes.c_uint8)),
        ("rec", _F64P),
        ("rec_cap", ctypes.c_int64),
        ("n_rec", ctypes.c_int64),
        ("rec_threshold", ctypes.c_double),
        ("x1", ctypes.c_int64),
        ("y1", ctypes.c_int64),
    ]


class _Stats(ctypes.Structure):
    _fields_ = [("n_windows", ctypes.c_int64), ("cells", ctypes.c_int64), ("log_likelihood", ctypes.c_double)]


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "all"], cwd=_HERE)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_set_exact_logadd.argtypes = [ctypes.c_int]
        L.or_record_windows.argtypes = [_I64P, ctypes.c_int64]
        L.or_recorded_windows.restype = ctypes.c_int64
        L.or_lookup.restype = ctypes.c_double
        L.or_lookup.argtypes = [ctypes.c_double]
        L.or_logadd.restype = ctypes.c_double
        L.or_logadd.argtypes = [ctypes.c_double, ctypes.c_double]
        L.or_kmer_id.restype = ctypes.c_int64
        L.or_kmer_id.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64]
        L.or_log_gauss_pdf.restype = ctypes.c_double
        L.or_log_gauss_pdf.argtypes = [ctypes.c_double] * 3
        L.or_log_inv_gauss_pdf.restype = ctypes.c_double
        L.or_log_inv_gauss_pdf.argtypes = [ctypes.c_double] * 3
        L.or_emission.restype = ctypes.c_double
        L.or_emission.argtypes = [ctypes.POINTER(_Model), ctypes.c_char_p, _F64P, ctypes.c_int]
        L.or_grid_spline_interp.restype = ctypes.c_double
        L.or_grid_spline_interp.argtypes = [ctypes.c_double, _F64P, _F64P, _F64P, ctypes.c_int64]
        L.or_spline_knot_slopes.argtypes = [_F64P, _F64P, ctypes.c_int64, _F64P]
        L.or_linspace.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int64, _F64P]
        L.or_hdp_expected_value.restype = ctypes.c_double
        L.or_hdp_expected_value.argtypes = [_F64P, _F64P, ctypes.c_int64]
        L.or_band.restype = ctypes.c_int
        L.or_band.argtypes = [_I64P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _I64P, _I64P]
        L.or_split_points.restype = ctypes.c_int64
        L.or_split_points.argtypes = [_I64P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_int, ctypes.c_int, _I64P, ctypes.c_int64]
        L.or_get_aligned_pairs.restype = ctypes.c_int
        L.or_get_aligned_pairs.argtypes = [ctypes.POINTER(_Model), ctypes.c_char_p, ctypes.c_int64, _F64P,
                                           ctypes.c_int64, _I64P, ctypes.c_int64, ctypes.POINTER(_Params),
                                           ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Pair), ctypes.c_int64,
                                           _I64P, ctypes.POINTER(_Stats)]
        L.or_get_aligned_pairs_degenerate.restype = ctypes.c_int
        L.or_get_aligned_pairs_degenerate.argtypes = [ctypes.POINTER(_Model), ctypes.c_char_p, ctypes.c_int64,
                                                      _F64P, ctypes.c_int64, _I64P, ctypes.c_int64,
                                                      ctypes.POINTER(_Params), ctypes.c_int, ctypes.c_int,
                                                      ctypes.c_char_p, ctypes.POINTER(_Pair), ctypes.c_int64, _I64P,
                                                      ctypes.POINTER(_Stats)]
        L.or_path_kmers.restype = ctypes.c_int64
        L.or_path_kmers.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64]
        L.or_path_legal.restype = ctypes.c_int
        L.or_path_legal.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.or_full_forward_backward.restype = ctypes.c_int
        L.or_full_forward_backward.argtypes = [ctypes.POINTER(_Model), ctypes.c_char_p, ctypes.c_int64, _F64P,
                                               ctypes.c_int64, _I64P, ctypes.c_int64, ctypes.c_int64,
                                               ctypes.c_double, ctypes.POINTER(_Pair), ctypes.c_int64, _I64P,
                                               _F64P, _F64P, _F64P]
        L.or_get_expectations.restype = ctypes.c_int
        L.or_get_expectations.argtypes = [ctypes.POINTER(_Model), ctypes.c_char_p, ctypes.c_int64, _F64P,
                                          ctypes.c_int64, _I64P, ctypes.c_int64, ctypes.POINTER(_Params),
                                          ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Expect)]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _p64(a):
    return a.ctypes.data_as(_F64P) if a is not None else None


def _pi64(a):
    return a.ctypes.data_as(_I64P) if a is not None else None


class Model:
    """Emission/transition tables in the reference's units.

    match, gapy: (n_kmers, 5) level_mean, level_sd, noise_mean, noise_sd, noise_lambda.
    gapx: (n_kmers,) log-space skip emission.
    trans: 9 log transitions in StateMachine3 field order.
    HDP: grid (G,), density/slopes (rows, G), kmer_row (n_kmers,)."""

    def __init__(self, kind, alphabet, kmer_length, trans, match, gapy=None, gapx=None,
                 scale=1.0, shift=0.0, var=1.0, grid=None, density=None, slopes=None, kmer_row=None):
        self.kind = kind
        self.alphabet = "".join(sorted(alphabet))
        self.kmer_length = int(kmer_length)
        n = len(self.alphabet) ** self.kmer_length
        self.trans = _f64(trans)
        self.match = _f64(match).reshape(n, 5)
        self.gapy = _f64(gapy if gapy is not None else self.match).reshape(n, 5)
        self.gapx = _f64(gapx if gapx is not None else np.full(n, np.log(0.1)))
        self.scale, self.shift, self.var = float(scale), float(shift), float(var)
        self.grid = _f64(grid) if grid is not None else None
        self.density = _f64(density) if density is not None else None
        self.slopes = _f64(slopes) if slopes is not None else None
        self.kmer_row = _i64(kmer_row) if kmer_row is not None else None

    def with_scaling(self, scale, shift, var):
        m = Model.__new__(Model)
        m.__dict__.update(self.__dict__)
        m.scale, m.shift, m.var = float(scale), float(shift), float(var)
        return m

    def _struct(self):
        s = _Model()
        s.kind = self.kind
        s.kmer_length = self.kmer_length
        s.alphabet_size = len(self.alphabet)
        s.alphabet = self.alphabet.encode()
        for i in range(9):
            s.trans[i] = float(self.trans[i])
        s.match = _p64(self.match)
        s.gapy = _p64(self.gapy)
        s.gapx = _p64(self.gapx)
        s.scale, s.shift, s.var = self.scale, self.shift, self.var
        if self.grid is not None:
            s.grid_length = len(self.grid)
            s.grid = _p64(self.grid)
            s.density = _p64(self.density)
            s.slopes = _p64(self.slopes)
            s.kmer_row = _pi64(self.kmer_row)
        return s


class Params:
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


def record_windows(fn, *args, cap=4096, **kw):
    """Run fn(*args, **kw) and return (its result, the traceback windows the oracle ran
    as rows (tracedBackTo, traceback diagonal, tracedBackFrom, atEnd))."""
    buf = np.zeros((cap, 4), dtype=np.int64)
    lib().or_record_windows(_pi64(buf), cap)
    try:
        out = fn(*args, **kw)
        n = lib().or_recorded_windows()
    finally:
        lib().or_record_windows(None, 0)
    return out, buf[:min(n, cap)].copy()


def set_exact_logadd(on):
    lib().or_set_exact_logadd(1 if on else 0)


def lookup(x):
    return lib().or_lookup(float(x))


def logadd(x, y):
    return lib().or_logadd(float(x), float(y))


def kmer_id(kmer, alphabet, kmer_length=None):
    alphabet = "".join(sorted(alphabet))
    k = len(kmer) if kmer_length is None else kmer_length
    return lib().or_kmer_id(kmer.encode(), alphabet.encode(), len(alphabet), k)


def log_gauss_pdf(x, mu, sigma):
    return lib().or_log_gauss_pdf(x, mu, sigma)


def log_inv_gauss_pdf(x, mu, lam):
    return lib().or_log_inv_gauss_pdf(x, mu, lam)


def emission(model, kmer, event, match=True):
    s = model._struct()
    ev = _f64(list(event) + [0.0] * (4 - len(event)))
    return lib().or_emission(ctypes.byref(s), kmer.encode() if kmer is not None else None, _p64(ev),
                             1 if match else 0)


def grid_spline_interp(q, x, y, slope):
    x, y, slope = _f64(x), _f64(y), _f64(slope)
    return lib().or_grid_spline_interp(float(q), _p64(x), _p64(y), _p64(slope), len(x))


def spline_knot_slopes(x, y):
    x, y = _f64(x), _f64(y)
    k = np.empty_like(x)
    lib().or_spline_knot_slopes(_p64(x), _p64(y), len(x), _p64(k))
    return k


def linspace(start, stop, n):
    out = np.empty(n, dtype=np.float64)
    lib().or_linspace(float(start), float(stop), n, _p64(out))
    return out


def hdp_expected_value(grid, distr):
    grid, distr = _f64(grid), _f64(distr)
    return lib().or_hdp_expected_value(_p64(grid), _p64(distr), len(grid))


def band(anchors, lX, lY, expansion):
    """Per-diagonal (xmyL, xmyR) of band_construct, or raises on invalid anchors."""
    a = _i64(np.asarray(anchors, dtype=np.int64).reshape(-1, 2))
    L = np.empty(lX + lY + 1, dtype=np.int64)
    R = np.empty(lX + lY + 1, dtype=np.int64)
    rc = lib().or_band(_pi64(a), len(a), lX, lY, expansion, _pi64(L), _pi64(R))
    if rc != 0:
        raise ValueError(f"or_band failed: {rc}")
    return L, R


def split_points(anchors, lX, lY, split_matrix, ragged_left, ragged_right):
    a = _i64(np.asarray(anchors, dtype=np.int64).reshape(-1, 2))
    n = lib().or_split_points(_pi64(a), len(a), lX, lY, split_matrix, int(ragged_left), int(ragged_right), None, 0)
    if n < 0:
        raise ValueError("invalid anchors")
    out = np.zeros((max(n, 1), 4), dtype=np.int64)
    lib().or_split_points(_pi64(a), len(a), lX, lY, split_matrix, int(ragged_left), int(ragged_right),
                          _pi64(out), n)
    return out[:n]


def path_kmers(kmer, options):
    """The path k-mers of one k-mer with degenerate 'X' bases, in path order."""
    k = len(kmer)
    n = lib().or_path_kmers(kmer.encode(), k, options.encode(), None, 0)
    if n < 0:
        raise ValueError("an X without options")
    buf = ctypes.create_string_buffer(int(n) * k)
    lib().or_path_kmers(kmer.encode(), k, options.encode(), buf, n)
    raw = buf.raw.decode()
    return [raw[i * k:(i + 1) * k] for i in range(int(n))]


def path_legal(from_kmer, to_kmer, k=6):
    return bool(lib().or_path_legal(from_kmer.encode() if from_kmer else None,
                                    to_kmer.encode() if to_kmer else None, k))


def get_aligned_pairs(model, ref, events, anchors, params=None, ragged_left=True, ragged_right=True,
                      cap=None, options=None):
    """Oracle getAlignedPairsUsingAnchors.  ref: str/bytes of nucleotides (k-mer i is
    ref[i:i+k]); events: (lY, >=2) float array of (mean, noise[, duration, start]).
    Returns (pairs (n,3) int64 [p_int, x, y] in the reference's list order, stats dict).
    With `options` (the degenerate base options of 'X', e.g. "CEO") the pairs are
    (n,4) [p_int, x, y, path]."""
    params = params or Params()
    refb = ref.encode() if isinstance(ref, str) else bytes(ref)
    lX = len(refb) - (model.kmer_length - 1) if len(refb) > 0 else 0
    ev = np.zeros((len(events), 4), dtype=np.float64)
    e = np.asarray(events, dtype=np.float64)
    if len(e):
        ev[:, : min(4, e.shape[1])] = e[:, :4]
    a = _i64(np.asarray(anchors, dtype=np.int64).reshape(-1, 2))
    if cap is None:
        cap = 8 * (lX + len(ev)) + 1024
    out = (_Pair * cap)()
    n = ctypes.c_int64(0)
    st = _Stats()
    ms = model._struct()
    ps = params._struct()
    rc = lib().or_get_aligned_pairs_degenerate(ctypes.byref(ms), refb, lX, _p64(ev), len(ev), _pi64(a), len(a),
                                               ctypes.byref(ps), int(ragged_left), int(ragged_right),
                                               options.encode() if options else None, out, cap,
                                               ctypes.byref(n), ctypes.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle getAlignedPairs failed: {rc}")
    if n.value > cap:
        return get_aligned_pairs(model, ref, events, anchors, params, ragged_left, ragged_right, n.value, options)
    arr = np.frombuffer(out, dtype=np.int64, count=4 * n.value).reshape(-1, 4).copy()
    if not options:
        arr = arr[:, :3].copy()
    return arr, {"n_windows": st.n_windows, "cells": st.cells, "log_likelihood": st.log_likelihood}


def get_expectations(model, ref, events, anchors, params=None, ragged_left=True, ragged_right=True,
                     transition_pseudocount=0.001, emission_pseudocount=0.001, rec_threshold=None, rec_cap=None):
    """Oracle getExpectationsUsingAnchors with the ContinuousPairHmm accumulators
    (pseudocounts as hmmContinuous_getExpectationsHmm sets them).  Returns a dict:
    transitions (3, 3) [from, to], likelihood, event_exp (n_kmers, 2), posteriors,
    observed (bool), and with rec_threshold the match-transition records (n, 5):
    k-mer position, event, p(M->M), p(X->M), p(Y->M) in call order."""
    params = params or Params()
    refb = ref.encode() if isinstance(ref, str) else bytes(ref)
    lX = len(refb) - (model.kmer_length - 1) if len(refb) > 0 else 0
    ev = np.zeros((len(events), 4), dtype=np.float64)
    e = np.asarray(events, dtype=np.float64)
    if len(e):
        ev[:, : min(4, e.shape[1])] = e[:, :4]
    a = _i64(np.asarray(anchors, dtype=np.int64).reshape(-1, 2))
    nk = len(model.alphabet) ** model.kmer_length
    trans = np.full(9, transition_pseudocount)
    eexp = np.zeros(2 * nk)
    post = np.full(nk, emission_pseudocount)
    obs = np.zeros(nk, dtype=np.uint8)
    if rec_cap is None:
        rec_cap = 0 if rec_threshold is None else 8 * (lX + len(ev)) + 1024
    rec = np.zeros(5 * max(rec_cap, 1))
    ex = _Expect(_p64(trans), 0.0, _p64(eexp), _p64(post), obs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                 _p64(rec), rec_cap, 0, 0.0 if rec_threshold is None else rec_threshold, 0, 0)
    ms = model._struct()
    ps = params._struct()
    rc = lib().or_get_expectations(ctypes.byref(ms), refb, lX, _p64(ev), len(ev), _pi64(a), len(a),
                                   ctypes.byref(ps), int(ragged_left), int(ragged_right), ctypes.byref(ex))
    if rc != 0:
        raise RuntimeError(f"oracle getExpectations failed: {rc}")
    if rec_threshold is not None and ex.n_rec > rec_cap:
        return get_expectations(model, ref, events, anchors, params, ragged_left, ragged_right,
                                transition_pseudocount, emission_pseudocount, rec_threshold, ex.n_rec)
    out = {"transitions": trans.reshape(3, 3), "likelihood": ex.likelihood, "event_exp": eexp.reshape(-1, 2),
           "posteriors": post, "observed": obs.astype(bool)}
    if rec_threshold is not None:
        out["records"] = rec[:5 * ex.n_rec].reshape(-1, 5).copy()
    return out


def full_forward_backward(model, ref, events, anchors, expansion, threshold):
    refb = ref.encode() if isinstance(ref, str) else bytes(ref)
    lX = len(refb) - (model.kmer_length - 1)
    ev = np.zeros((len(events), 4), dtype=np.float64)
    e = np.asarray(events, dtype=np.float64)
    ev[:, : min(4, e.shape[1])] = e[:, :4]
    a = _i64(np.asarray(anchors, dtype=np.int64).reshape(-1, 2))
    cap = 4 * (lX + 1) * (len(ev) + 1)
    out = (_Pair * cap)()
    n = ctypes.c_int64(0)
    tf = ctypes.c_double(0)
    tb = ctypes.c_double(0)
    dt = np.empty(lX + len(ev) + 1, dtype=np.float64)
    ms = model._struct()
    rc = lib().or_full_forward_backward(ctypes.byref(ms), refb, lX, _p64(ev), len(ev), _pi64(a), len(a),
                                        expansion, threshold, out, cap, ctypes.byref(n), ctypes.byref(tf),
                                        ctypes.byref(tb), _p64(dt))
    if rc != 0:
        raise RuntimeError("oracle full forward-backward failed")
    arr = np.frombuffer(out, dtype=np.int64, count=4 * n.value).reshape(-1, 4)[:, :3].copy()
    return arr, tf.value, tb.value, dt
