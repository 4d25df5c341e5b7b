"""The reference's real-read known answers (tests/stateMachineTests.c:823-851).

ZymoC_ch_1_file1.npRead (an R7.3 read, its template strand) against ZymoRef.txt with
testModelR73_acegot_template.model, threshold 0.01, both ends ragged: the reference
expects 1076 aligned pairs from the model scaled to the read (emissions_signal_scaleModel)
and from the descaling machine (getStateMachine3_descaled with noise scaling), with
lastz anchors, and 1076 from the unbanded full matrix.  lastz is not available here,
so these runs give no anchors: the band is then the whole matrix (as unbanded) and
the traceback windows are the banded path's.  The three files are copies of the
reference's test data.
"""
import os

import numpy as np
import pytest

import oracle
from signalalign_amd import formats, models, native
from tests import _parity as P

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
EXPECTED_PAIRS = 1076


def _inputs():
    r = formats.load_npread(os.path.join(GOLDEN, "ZymoC_ch_1_file1.npRead"))
    ref = open(os.path.join(GOLDEN, "ZymoRef.txt")).read().split()[0]
    m = models.load_model_file(os.path.join(GOLDEN, "testModelR73_acegot_template.model"))
    return r, ref, m


def _machines(r, ref, m):
    """(name, model, emission kind, read dict) for the scaled and the descaling machine."""
    t = r["template"]
    p = t["params"]
    base = {"ref": ref.encode(), "events": t["events"], "anchors": np.zeros((0, 2), dtype=np.int64)}
    scaled = models.scale_model(m, p["scale"], p["shift"], p["var"], p["scale_sd"], p["var_sd"])
    descaling = models.scale_noise(m, p["scale_sd"], p["var_sd"])
    return [("scaled", scaled, native.SA_EMIT_GAUSS, dict(base, scale=1.0, shift=0.0, var=1.0)),
            ("descaled", descaling, native.SA_EMIT_GAUSS_DESCALED,
             dict(base, scale=p["scale"], shift=p["shift"], var=p["var"]))]


def test_npread_loader_reads_the_reference_file():
    r = formats.load_npread(os.path.join(GOLDEN, "ZymoC_ch_1_file1.npRead"))
    t, c = r["template"], r["complement"]
    assert r["read_length"] == 950 and r["two_d"]
    assert t["events"].shape == (799, 4) and c["events"].shape == (670, 4)
    assert len(t["read"]) == 879 and len(t["strand_event_map"]) == 879
    assert len(c["read"]) == 766 and len(c["strand_event_map"]) == 766
    assert len(t["event_map"]) == 950 and len(c["event_map"]) == 950
    assert t["params"]["scale"] == 1.02618833016 and t["params"]["var_sd"] == 1.21382992558
    assert c["params"]["drift"] == 0.0110887440008
    assert len(t["model_state"]) == 799 and len(t["p_model"]) == 799
    # event maps index events of their strand
    assert t["strand_event_map"].min() >= 0 and t["strand_event_map"].max() < 799


def test_npread_loader_rejects_bad_files(tmp_path):
    good = open(os.path.join(GOLDEN, "ZymoC_ch_1_file1.npRead")).read().split("\n")
    bad = list(good)
    bad[0] = " ".join(bad[0].split()[:17])
    f = tmp_path / "bad1.npRead"
    f.write_text("\n".join(bad))
    with pytest.raises(ValueError, match="line 1"):
        formats.load_npread(str(f))
    bad = list(good)
    bad[7] = " ".join(bad[7].split()[:-4])  # one template event short
    f = tmp_path / "bad2.npRead"
    f.write_text("\n".join(bad))
    with pytest.raises(ValueError, match="template events"):
        formats.load_npread(str(f))


def test_scale_model_formulas():
    """emissions_signal_scaleModel (impl/stateMachine.c:631-667) on one row, by hand."""
    _, _, m = _inputs()
    s = models.scale_model(m, 1.5, 10.0, 1.25, 1.1, 0.9)
    i = 123
    mu, sd, nmu, nsd, lam = m.match[i]
    assert np.allclose(s.match[i], [mu * 1.5 + 10.0, sd * 1.25, nmu * 1.1, np.sqrt((nmu * 1.1) ** 3 / (lam * 0.9)),
                                    lam * 0.9], rtol=0, atol=0)
    gmu, gsd, gnmu, _, glam = m.gapy[i]
    assert s.gapy[i][0] == gmu * 1.5 + 10.0 and s.gapy[i][1] == gsd * 1.25 and s.gapy[i][4] == glam * 0.9
    assert s.gapy[i][3] == np.sqrt((gnmu * 1.1) ** 3 / (lam * 0.9))  # the match row's lambda, as the reference


def test_oracle_reproduces_the_reference_count(oracle_lib):
    r, ref, m = _inputs()
    for name, mm, kind, read in _machines(r, ref, m):
        pairs, st = P.run_oracle(mm, read, native.Params(threshold=0.01),
                                 kind=oracle.GAUSS if kind == native.SA_EMIT_GAUSS else oracle.GAUSS_DESCALED)
        assert len(pairs) == EXPECTED_PAIRS, (name, len(pairs))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gpu_reproduces_the_reference_count(gpu_ctx, precision):
    """The full-matrix band (≈800 cells per diagonal) runs the LDS sweeps, their rings
    in HBM where LDS is too small (FP64)."""
    r, ref, m = _inputs()
    prec = native.SA_FP64 if precision == "fp64" else native.SA_FP32
    for name, mm, kind, read in _machines(r, ref, m):
        dm = gpu_ctx.model(mm, kind)
        params = native.Params(threshold=0.01)
        res, _ = gpu_ctx.align_batch(dm, [P.native_read(read)], params, prec)
        want, st = P.run_oracle(mm, read, params,
                                kind=oracle.GAUSS if kind == native.SA_EMIT_GAUSS else oracle.GAUSS_DESCALED)
        if precision == "fp64":
            out = P.compare(res[0].pairs, want, params.threshold)
            assert P.rel(res[0].log_likelihood, st["log_likelihood"]) <= P.LL_REL
            assert len(res[0].pairs) == EXPECTED_PAIRS, (name, out)
        else:
            out = P.compare(res[0].pairs, want, params.threshold, p_tol=P.P_TOL_FP32)
            assert P.rel(res[0].log_likelihood, st["log_likelihood"]) <= P.LL_REL_FP32
        print(name, precision, out)


def _hdp_model(m):
    h = formats.load_nhdp(os.path.join(GOLDEN, "templateSingleLevelFixed.nhdp"))
    return h, models.HdpModel(h["alphabet"], h["kmer_length"], m.transitions, h["grid"], h["density"], h["slopes"],
                              h["kmer_row"], m.match[:, 0])


def test_nhdp_loader_reads_the_reference_file():
    """Structure of the serialised HDP, and its spline slopes: written by the
    reference's finalize_distributions, they equal spline_knot_slopes of the stored
    posterior predictive densities exactly (pins the slope restatement on reference
    output)."""
    h = formats.load_nhdp(os.path.join(GOLDEN, "templateSingleLevelFixed.nhdp"))
    assert h["alphabet"] == "ACEGOT" and h["kmer_length"] == 6 and h["n_dps"] == 6 ** 6 + 1
    assert np.array_equal(h["grid"], np.array(models.linspace(0.0, 100.0, 100)))
    assert h["density"].shape == h["slopes"].shape == (352, 100)
    assert (h["density"] >= 0).all()
    assert np.array_equal(models.spline_knot_slopes_rows(h["grid"], h["density"]), h["slopes"])
    rows, counts = np.unique(h["kmer_row"], return_counts=True)
    assert len(rows) == 352 and counts.max() == 46305  # unobserved k-mers fall back to the root
    assert (counts[counts != 46305] == 1).all()


@pytest.mark.gpu
def test_gpu_real_read_on_the_reference_hdp(gpu_ctx):
    """The Zymo read's events against the reference's HDP density grid (descaled by
    the read's template parameters, threshold 0.1, full band): FP64 to the oracle's
    bar, FP32 to its stated tolerance."""
    r, ref, m = _inputs()
    h, hm = _hdp_model(m)
    t = r["template"]
    p = t["params"]
    read = {"ref": ref.encode(), "events": t["events"], "anchors": np.zeros((0, 2), dtype=np.int64),
            "scale": p["scale"], "shift": p["shift"], "var": p["var"]}
    dm = gpu_ctx.model(hm)
    params = native.Params(threshold=0.1)
    want, st = P.run_oracle(hm, read, params)
    for prec, p_tol, ll in ((native.SA_FP64, P.P_TOL, P.LL_REL), (native.SA_FP32, P.P_TOL_FP32, P.LL_REL_FP32)):
        res, _ = gpu_ctx.align_batch(dm, [P.native_read(read)], params, prec)
        out = P.compare(res[0].pairs, want, params.threshold, p_tol=p_tol)
        assert P.rel(res[0].log_likelihood, st["log_likelihood"]) <= ll
        print(prec, out)
