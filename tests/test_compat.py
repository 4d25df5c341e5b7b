"""The reference's pairwiseAligner C API (include/signalalign_compat.h) over the
C-ABI: exports, the reference-layout state machine built by the C model loader,
and alignments driven from C the way the reference's tests drive them
(tests/stateMachineTests.c:565-716)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from signalalign_amd import models, native, synth
from tests import _parity as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMPAT_LIB = os.path.join(ROOT, "signalalign_amd", "libsignalalign_compat.so")
COMPAT_BIN = os.path.join(ROOT, "tests", "compat", "compat_align")
HEADER = os.path.join(ROOT, "include", "signalalign_compat.h")
MODEL5 = os.path.join(ROOT, "tests", "golden", "testModelR9_5mer_acgt_template.model")


def _declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M)
    return sorted({n for n in names if n not in ("double", "void", "int64_t")})


def test_compat_exports_every_declared_function():
    """Every function signalalign_compat.h and signalalign_machine.h declare."""
    lib = ctypes.CDLL(COMPAT_LIB)
    names = _declared_functions()
    assert len(names) >= 30, names
    machine = _declared_functions(os.path.join(ROOT, "include", "signalalign_machine.h"))
    assert len(machine) >= 40, machine
    missing = [n for n in names + machine if not hasattr(lib, n)]
    assert not missing, missing


class _SM(ctypes.Structure):
    """struct _stateMachine + the StateMachine3 transitions (inc/stateMachine.h)."""
    _fields_ = [("type", ctypes.c_int), ("stateNumber", ctypes.c_int64), ("matchState", ctypes.c_int64),
                ("parameterSetSize", ctypes.c_int64), ("scale", ctypes.c_double), ("shift", ctypes.c_double),
                ("var", ctypes.c_double), ("alphabet", ctypes.c_char_p), ("alphabetSize", ctypes.c_int64),
                ("kmerLength", ctypes.c_int64), ("match", ctypes.POINTER(ctypes.c_double)),
                ("gapx", ctypes.POINTER(ctypes.c_double)), ("gapy", ctypes.POINTER(ctypes.c_double)),
                ("fp", ctypes.c_void_p * 6), ("trans", ctypes.c_double * 9)]


class _NPP(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("scale", "shift", "var", "shift_sd", "scale_sd", "var_sd", "drift")]


def test_c_model_loader_matches_the_pinned_loader():
    """getStateMachine3 / getStateMachine3_descaled (C) against models.load_model_file,
    which the test_poreModel KAT pins: same tables, transitions, alphabet, priors."""
    lib = ctypes.CDLL(COMPAT_LIB)
    lib.getStateMachine3.restype = ctypes.POINTER(_SM)
    lib.getStateMachine3.argtypes = [ctypes.c_char_p]
    lib.getStateMachine3_descaled.restype = ctypes.POINTER(_SM)
    lib.getStateMachine3_descaled.argtypes = [ctypes.c_char_p, _NPP, ctypes.c_bool]
    lib.stateMachine_destruct.argtypes = [ctypes.POINTER(_SM)]
    ref = models.load_model_file(MODEL5)
    sm = lib.getStateMachine3(MODEL5.encode())
    s = sm.contents
    n = ref.n_kmers
    assert (s.type, s.stateNumber, s.parameterSetSize, s.kmerLength) == (2, 3, n, ref.kmer_length)
    assert s.alphabet.decode() == ref.alphabet
    assert np.array_equal(np.ctypeslib.as_array(s.match, (n * 5,)), ref.match.reshape(-1))
    assert np.array_equal(np.ctypeslib.as_array(s.gapy, (n * 5,)), ref.gapy.reshape(-1))
    assert np.array_equal(np.ctypeslib.as_array(s.gapx, (n,)), ref.gapx)
    assert np.array_equal(np.array(s.trans[:]), ref.transitions)
    # start / end priors (impl/stateMachine.c:1022-1061) through the struct's pointers
    F = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(_SM), ctypes.c_int64)
    start, end, rend, rstart = (F(s.fp[i]) for i in range(4))
    t = ref.transitions
    assert [start(sm, i) for i in range(3)] == [0.0, -np.inf, -np.inf]
    assert [rstart(sm, i) for i in range(3)] == [-np.inf, 0.0, 0.0]
    assert [end(sm, i) for i in range(3)] == [t[models.T_MATCH_CONTINUE], t[models.T_MATCH_FROM_GAP_X],
                                              t[models.T_MATCH_FROM_GAP_Y]]
    assert [rend(sm, i) for i in range(3)] == [(t[models.T_GAP_OPEN_X] + t[models.T_GAP_OPEN_Y]) / 2.0,
                                               t[models.T_GAP_EXTEND_X], t[models.T_GAP_EXTEND_Y]]
    lib.stateMachine_destruct(sm)
    # descaled machine with noise scaling (emissions_signal_scaleNoise, :609-629)
    npp = _NPP(1.02, -3.0, 1.1, 0.0, 1.3, 0.8, 0.0)
    sm = lib.getStateMachine3_descaled(MODEL5.encode(), npp, True)
    s = sm.contents
    want = models.scale_noise(ref, 1.3, 0.8)
    assert (s.scale, s.shift, s.var) == (1.02, -3.0, 1.1)
    assert np.allclose(np.ctypeslib.as_array(s.match, (n * 5,)), want.match.reshape(-1), rtol=1e-15, atol=0)
    assert np.allclose(np.ctypeslib.as_array(s.gapy, (n * 5,)), want.gapy.reshape(-1), rtol=1e-15, atol=0)
    lib.stateMachine_destruct(sm)


def test_compat_without_a_device_aborts_like_the_reference(tmp_path):
    """No GPU: the call fails loudly (abort with a message), as the reference's
    st_errAbort does; nothing runs on the host instead."""
    if native.device_count() > 0:
        pytest.skip("a device is visible")
    r = subprocess.run([COMPAT_BIN, "kat5", MODEL5], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "no usable GPU" in r.stderr


@pytest.mark.gpu
def test_compat_5mer_kat():
    """tests/stateMachineTests.c:565-692 through getAlignedPairsUsingAnchors."""
    r = subprocess.run([COMPAT_BIN, "kat5", MODEL5], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "ok" in r.stdout


def _write_input(path, r, params, ragged):
    with open(path, "w") as f:
        ref = r["ref"].decode() if isinstance(r["ref"], bytes) else r["ref"]
        f.write(f"{len(ref)}\n{ref}\n")
        f.write(f"{r['scale']!r} {r['shift']!r} {r['var']!r} {int(ragged[0])} {int(ragged[1])}\n")
        f.write(f"{params.diagonal_expansion} {params.threshold!r} {params.min_diags_between_traceback} "
                f"{params.traceback_diagonals} {params.split_matrix_bigger_than_this}\n")
        ev = np.asarray(r["events"], dtype=np.float64)
        f.write(f"{len(ev)}\n")
        for row in ev:
            f.write(" ".join(repr(float(v)) for v in row) + "\n")
        anc = np.asarray(r["anchors"], dtype=np.int64).reshape(-1, 2)
        f.write(f"{len(anc)}\n")
        for x, y in anc:
            f.write(f"{x} {y}\n")


@pytest.mark.gpu
@pytest.mark.parametrize("ragged,split", [((True, True), 3000 * 3000), ((False, True), 3000 * 3000),
                                          ((True, True), 50 * 50)])
def test_compat_alignment_matches_oracle(tmp_path, ragged, split):
    """A synthetic read through getStateMachine3_descaled + getAlignedPairsUsingAnchors
    (C), against the FP64 oracle: pair set, posteriors, list order, log-likelihood,
    and the path k-mer of every pair."""
    m = synth.gaussian_model("ACGT", 6, seed=1)
    mpath = str(tmp_path / "g6.model")
    models.write_model_file(m, mpath)
    r = synth.single_read(m, 2500, seed=77)
    params = native.Params(diagonal_expansion=20, split_matrix_bigger_than_this=split)
    inp, out = str(tmp_path / "in.txt"), str(tmp_path / "out.tsv")
    _write_input(inp, r, params, ragged)
    res = subprocess.run([COMPAT_BIN, "align", mpath, inp, out], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = open(out).read().splitlines()
    ll = float(lines[0])
    rows = [ln.split("\t") for ln in lines[1:]]
    pairs = np.array([[int(a), int(b), int(c)] for a, b, c, _ in rows], dtype=np.int64).reshape(-1, 3)
    ref = r["ref"].decode() if isinstance(r["ref"], bytes) else r["ref"]
    assert all(km == ref[int(x):int(x) + 6] for _, x, _, km in rows)
    want, st = P.run_oracle(m, r, params, *ragged)
    out_ = P.compare(pairs, want, params.threshold)
    assert P.rel(ll, st["log_likelihood"]) <= P.LL_REL
    print(out_)
