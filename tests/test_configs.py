"""The benchmark configurations at their own shapes, against the FP64 oracle (GPU tests).

  * one long read as the plain drop-in: 50 kb ACGT reference, 6-mer Gaussian table,
    band expansion 100, FP64 and FP32;
  * HDP emissions: ACEGT 5-mer template and complement HDP models (3125 k-mers,
    mixture densities on a 2000-point grid over 30-130 pA), reads aligned on both
    strands -- the complement read against the reverse-complement target with the
    complement model -- band expansion 50, FP32 and FP64;
  * batched: ACEGT 6-mer Gaussian table (15625 k-mers), reads of uneven length (one of
    24 kb) in one batch, band expansion 50, FP32 and FP64.

FP64: the hot path's bar (same pair set except pairs within 1e-3 of the threshold,
posteriors within 1e-4, log-likelihood within 1e-6 relative, the reference's list
order).  FP32: log-likelihood within 1e-6 relative, posteriors within P_TOL_FP32 and
at most 1 in 500 pairs beyond 1e-4 (tests/_parity.py states why).
"""
import numpy as np
import pytest

from signalalign_amd import native, synth
from tests import _parity as P

pytestmark = pytest.mark.gpu


def _fp32_ok(out, ll_rel):
    assert ll_rel <= P.LL_REL_FP32, (ll_rel, out)


def _check_batch(ctx, m, reads, params, label, ragged=(True, True)):
    dm = ctx.model(m)
    b = native.Batch(ctx, dm, [P.native_read(r, *ragged) for r in reads], params)
    try:
        refs = [P.run_oracle(m, r, params, *ragged) for r in reads]
        for prec, name in ((native.SA_FP64, "fp64"), (native.SA_FP32, "fp32")):
            print(f"{label}: {len(reads)} reads {name}", flush=True)
            res, tm = b.run(prec)
            for i, (g, (ref, st)) in enumerate(zip(res, refs)):
                assert g.cells == st["cells"] and g.n_windows == st["n_windows"]
                ll = P.rel(g.log_likelihood, st["log_likelihood"])
                if prec == native.SA_FP64:
                    out = P.compare(g.pairs, ref, params.threshold)
                    assert ll <= P.LL_REL, (i, ll)
                else:
                    out = P.compare(g.pairs, ref, params.threshold, p_tol=P.P_TOL_FP32)
                    _fp32_ok(out, ll)
                out["ll_rel"] = ll
                print(f"  read {i} ({len(reads[i]['ref'])} nt, {len(reads[i]['events'])} events, {g.cells} cells): "
                      f"{out}", flush=True)
    finally:
        b.close()
        dm.close()


def test_long_read_50kb_e100():
    """The plain drop-in on one 50 kb read (about 67k events, e=100, about 17M cells,
    over 100 traceback windows), with one and with several wavefronts per forward
    sweep."""
    ctx = native.Context(0)
    try:
        m = synth.gaussian_model("ACGT", 6, seed=1)
        r = synth.single_read(m, 50000, seed=2074)
        params = native.Params(diagonal_expansion=100)
        _check_batch(ctx, m, [r], params, "50 kb read")
    finally:
        ctx.close()


def test_long_read_50kb_one_wavefront(monkeypatch):
    monkeypatch.setenv("SA_FWD_WAVES", "1")
    ctx = native.Context(0)
    try:
        m = synth.gaussian_model("ACGT", 6, seed=1)
        r = synth.single_read(m, 50000, seed=2075)
        _check_batch(ctx, m, [r], native.Params(diagonal_expansion=100), "50 kb read, one wavefront")
    finally:
        ctx.close()


def test_hdp_5mer_2000_point_grid_both_strands():
    """HDP configuration shapes: template reads with the template model, complement
    reads (reverse-complement target) with the complement model."""
    ctx = native.Context(0)
    try:
        mt = synth.hdp_model("ACEGT", 5, seed=3)
        mc = synth.hdp_model("ACEGT", 5, seed=4)
        assert mt.density.shape[1] == 2000 and mt.n_kmers == 3125
        reads = synth.batch_reads(mt, 4, 5000, seed=77, methyl_cpg=0.3, strands=2, complement_model=mc)
        params = native.Params(diagonal_expansion=50)
        for strand, m in ((0, mt), (1, mc)):
            rs = [r for r in reads if r["strand"] == strand]
            if strand == 1:
                # the complement read's target is the reverse complement of its template's
                t = [r for r in reads if r["strand"] == 0]
                comp = bytes.maketrans(b"ACGTE", b"TGCAE")
                for a, c in zip(t, rs):
                    assert a["ref"].replace(b"E", b"C")[::-1].translate(comp) == c["ref"].replace(b"E", b"C")
            _check_batch(ctx, m, rs, params, f"HDP strand {strand}")
    finally:
        ctx.close()


def test_hdp_strands_in_one_batch_under_a_merged_table():
    """Template and complement reads in one batch under models.merge_strand_models
    (complement references lower-cased): every read's pairs and log-likelihood equal
    its run with its own strand's model, bit for bit, FP32 and FP64."""
    from signalalign_amd import models
    ctx = native.Context(0)
    try:
        mt = synth.hdp_model("ACEGT", 5, seed=3)
        mc = synth.hdp_model("ACEGT", 5, seed=4)
        reads = synth.batch_reads(mt, 3, 3000, seed=78, methyl_cpg=0.3, strands=2, complement_model=mc)
        params = native.Params(diagonal_expansion=50)
        dmm = ctx.model(models.merge_strand_models(mt, mc))
        dmt, dmc = ctx.model(mt), ctx.model(mc)
        merged = [native.Read(r["ref"] if r["strand"] == 0 else r["ref"].lower(), r["events"], r["anchors"],
                              r["scale"], r["shift"], r["var"]) for r in reads]
        for prec in (native.SA_FP32, native.SA_FP64):
            print(f"merged strands, precision {prec}", flush=True)
            got, _ = ctx.align_batch(dmm, merged, params, prec)
            for r, g in zip(reads, got):
                want, _ = ctx.align_batch(dmt if r["strand"] == 0 else dmc, [P.native_read(r, True, True)],
                                          params, prec)
                assert np.array_equal(g.pairs, want[0].pairs) and g.log_likelihood == want[0].log_likelihood
        for dm in (dmm, dmt, dmc):
            dm.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("waves", ["1", "4"])
def test_hdp_emission_plane_matches_gathers(monkeypatch, waves):
    """HDP batches read each cell's density emission from the emission plane filled
    before the sweeps (emission path 3); SA_HDP_PLANE=0 keeps the sweeps' own density
    gathers (path 2).  The HDP emission is computed without multiply-add fusion, so
    both give the same pairs and log-likelihoods bit for bit, FP64 and FP32, with one
    and with several wavefronts per forward sweep."""
    monkeypatch.setenv("SA_FWD_WAVES", waves)
    ctx = native.Context(0)
    try:
        m = synth.hdp_model("ACEGT", 5, seed=3)
        reads = synth.batch_reads(m, 6, 4000, seed=79, methyl_cpg=0.3)
        params = native.Params(diagonal_expansion=50)
        dm = ctx.model(m)
        nr = [P.native_read(r) for r in reads]
        for prec in (native.SA_FP64, native.SA_FP32):
            got = {}
            for plane in ("1", "0"):
                monkeypatch.setenv("SA_HDP_PLANE", plane)
                res, tm = ctx.align_batch(dm, nr, params, prec)
                assert tm["emission_path"] == (3 if plane == "1" else 2)
                got[plane] = res
            for i, (a, b) in enumerate(zip(got["1"], got["0"])):
                print(f"precision {prec} read {i}: {len(a.pairs)} pairs, logL {a.log_likelihood}", flush=True)
                assert a.cells == b.cells
                assert np.array_equal(a.pairs, b.pairs), (prec, i)
                assert a.log_likelihood == b.log_likelihood, (prec, i)
        dm.close()
    finally:
        ctx.close()


def test_batched_acegt_6mer_uneven_lengths():
    """Batched configuration shapes: the 15625-k-mer ACEGT table, 5mC at 30% of CpG
    sites, lengths 1-24 kb in one batch (the long one on the several-wavefront sweep,
    the rest on one-wavefront sweeps)."""
    ctx = native.Context(0)
    try:
        m = synth.gaussian_model("ACEGT", 6, seed=1)
        assert m.n_kmers == 15625
        reads = []
        for i, n in enumerate([24000, 1000, 3000, 8000, 5500, 12000, 2200, 9000]):
            rng = synth.rng_for(313, i, 0)
            fwd, _ = synth.random_reference(n, rng, methyl_cpg=0.3)
            reads.append(synth.simulate_read(m, fwd, synth.rng_for(313, i, 1)))
        _check_batch(ctx, m, reads, native.Params(diagonal_expansion=50), "ACEGT batch")
    finally:
        ctx.close()


@pytest.mark.parametrize("expansion", [20, 50, 100])
def test_halo_sweep_bit_equal_to_barrier_sweep(monkeypatch, expansion):
    """The several-wavefront forward sweep that synchronises every 8th diagonal through
    halo lanes gives the same pairs and log-likelihood, bit for bit, as the one that
    synchronises every diagonal (SA_HALO=0): bands of about 45, 93 and 190 cells (two,
    two or three and four wavefronts), FP32 and FP64, the model file's Gaussian tables,
    general Gaussian tables, HDP gathers and the HDP emission plane."""
    from signalalign_amd import models
    monkeypatch.setenv("SA_FWD_WAVES", "4")
    ctx = native.Context(0)
    try:
        g = synth.gaussian_model("ACGT", 6, seed=1)
        h = synth.hdp_model("ACGT", 4, seed=3)
        reads = [synth.single_read(g, n, seed=900 + n) for n in (6000, 2500)]
        hreads = [synth.single_read(h, 4000, seed=901)]
        params = native.Params(diagonal_expansion=expansion)
        gen = models.GaussianModel(g.alphabet, 6, g.transitions, g.match, g.gapy.copy(), g.gapx.copy())
        gen.gapy[:, 4] *= 1.3
        gen.gapx[::7] = np.log(0.05)
        cases = [("model file", g, reads), ("general tables", gen, reads[:1]), ("hdp plane", h, hreads),
                 ("hdp gathers", h, hreads)]
        for label, m, rs in cases:
            dm = ctx.model(m)
            nr = [P.native_read(r) for r in rs]
            for prec in (native.SA_FP32, native.SA_FP64):
                if label == "hdp gathers":
                    monkeypatch.setenv("SA_HDP_PLANE", "0")
                monkeypatch.setenv("SA_HALO", "1")
                got, tm = ctx.align_batch(dm, nr, params, prec)
                fwd = [(L["waves"], L["sync_every"]) for L in tm["launches"] if L["direction"] == 0]
                monkeypatch.setenv("SA_HALO", "0")
                want, tm0 = ctx.align_batch(dm, nr, params, prec)
                fwd0 = [(L["waves"], L["sync_every"]) for L in tm0["launches"] if L["direction"] == 0]
                monkeypatch.delenv("SA_HDP_PLANE", raising=False)
                print(f"e={expansion} {label} precision {prec}: forward (waves, sync) halo {fwd}, barrier {fwd0}",
                      flush=True)
                assert fwd and all(w >= 2 and se == 8 for w, se in fwd), fwd
                assert fwd0 and all(w >= 2 and se == 1 for w, se in fwd0), fwd0
                for a, b in zip(got, want):
                    assert np.array_equal(a.pairs, b.pairs), label
                    assert a.log_likelihood == b.log_likelihood, label
            dm.close()
    finally:
        ctx.close()
