"""The C-ABI library without a GPU: it loads, exports every symbol its header
declares, rejects calls when no device is present, and its host-side copies of
the kernels' integer geometry (band, traceback windows, split points) agree
with the oracle (CPU tests)."""
import os
import re

import numpy as np
import pytest

import oracle
from signalalign_amd import native, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sa_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = native.load()
    names = _declared("signalalign_amd.h")
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/signalalign_amd.h but not exported"


@pytest.mark.parametrize("cfg", [dict(length_ratio=0.5), dict(lanes=5)])
def test_server_config_checked_before_any_device_call(cfg):
    """A length class narrower than one length, or more than four lanes, is refused
    before the server touches a device."""
    with pytest.raises(native.SignalAlignError) as e:
        native.Server(0, **cfg)
    assert e.value.code == -1  # SA_ERR_INVALID_ARGUMENT


def test_no_device_is_an_error_not_a_fallback():
    if native.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(native.SignalAlignError) as e:
        native.Context(0)
    assert e.value.code == -7


def test_band_geometry_matches_oracle_kat(oracle_lib):
    lo, hi = native.band_geometry([(1, 0), (2, 1), (3, 3)], 6, 5, 2)
    L, R = oracle.band([(1, 0), (2, 1), (3, 3)], 6, 5, 2)
    d = np.arange(12)
    assert np.array_equal(lo, (d + L) // 2) and np.array_equal(hi, (d + R) // 2)


@pytest.mark.parametrize("seed", range(6))
def test_band_geometry_matches_oracle_random(oracle_lib, seed):
    m = synth.gaussian_model("ACGT", 6, seed=1)
    r = synth.single_read(m, 400 + 300 * seed, seed=200 + seed)
    lX, lY = len(r["ref"]) - 5, len(r["events"])
    for e in (0, 2, 20, 50, 100):
        lo, hi = native.band_geometry(r["anchors"], lX, lY, e)
        L, R = oracle.band(r["anchors"], lX, lY, e)
        d = np.arange(lX + lY + 1)
        assert np.array_equal(lo, (d + L) // 2)
        assert np.array_equal(hi, (d + R) // 2)
        assert np.all(np.diff(lo) >= 0) and np.all(np.diff(hi) >= 0)  # the register sweeps rely on this
    with pytest.raises(native.SignalAlignError):
        native.band_geometry([(5, 5), (4, 9)], lX, lY, 20)


@pytest.mark.parametrize("params", [native.Params(diagonal_expansion=20),
                                    native.Params(diagonal_expansion=50, min_diags_between_traceback=300,
                                                  traceback_diagonals=17),
                                    native.Params(diagonal_expansion=4, min_diags_between_traceback=50,
                                                  traceback_diagonals=3)])
def test_window_schedule_matches_oracle(oracle_lib, params):
    m = synth.gaussian_model("ACGT", 6, seed=1)
    r = synth.single_read(m, 3000, seed=301)
    lX, lY = len(r["ref"]) - 5, len(r["events"])
    lo, hi = native.band_geometry(r["anchors"], lX, lY, params.diagonal_expansion)
    ours = native.window_schedule(lo, hi, lX, lY, params)
    op = oracle.Params(params.threshold, params.min_diags_between_traceback, params.traceback_diagonals,
                       params.diagonal_expansion, params.split_matrix_bigger_than_this)
    om = oracle.Model(oracle.GAUSS_DESCALED, m.alphabet, 6, m.transitions, m.match, m.gapy, m.gapx,
                      r["scale"], r["shift"], 1.0)
    _, theirs = oracle.record_windows(oracle.get_aligned_pairs, om, r["ref"], r["events"], r["anchors"], op)
    assert len(ours) > 1
    assert np.array_equal(ours.astype(np.int64), theirs)


def test_split_points_match_oracle(oracle_lib):
    anchors = [(2000, 2000), (4002, 4001), (5000, 5000), (8000, 6000), (9000, 9000), (10000, 14000),
               (15000, 15000), (16000, 16000)]
    p = native.Params(split_matrix_bigger_than_this=2000 * 2000)
    for rl in (0, 1):
        for rr in (0, 1):
            for a in ([], anchors):
                ours = native.split_points(a, 20000, 25000, p, rl, rr)
                theirs = oracle.split_points(a, 20000, 25000, 2000 * 2000, rl, rr)
                assert np.array_equal(ours, theirs)
