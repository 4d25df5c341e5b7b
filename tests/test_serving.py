"""Serving loop (sa_server_*): continuous batching over mixed models, HBM-budget
admission, per-request results identical alone and in a mix."""
import numpy as np
import pytest

from signalalign_amd import native, synth
from tests import _parity as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def models_(gpu_ctx):
    g6 = synth.gaussian_model("ACEGT", 6, seed=5)
    hd = synth.hdp_model("ACEGT", 4, seed=7, grid_length=400)
    return g6, hd, gpu_ctx.model(g6), gpu_ctx.model(hd)


def _reads(model, lens, seed, methyl=0.3):
    out = []
    for i, n in enumerate(lens):
        rng = synth.rng_for(seed, i, 0)
        fwd, _ = synth.random_reference(int(n), rng, methyl_cpg=methyl)
        out.append(synth.simulate_read(model, fwd, synth.rng_for(seed, i, 1)))
    return out


def _collect(srv, tickets):
    out = []
    for t in tickets:
        r = srv.wait(t, timeout=120.0)
        assert r is not None, f"request {t} did not finish"
        out.append(r)
        srv.release(t)
    return out


def test_alone_equals_mixed(models_):
    """Each probe read submitted alone (idle server, a wave of one), then the same
    reads inside a burst of 80 reads of two models arriving together (waves of many
    reads, several waves in flight): identical pairs, posteriors and
    log-likelihood (FP32; the bar is exact equality)."""
    g6, hd, dg, dh = models_
    params = native.Params(diagonal_expansion=50)
    greads = _reads(g6, [300, 900, 2500, 4000, 7000, 12000] + [1500 + 97 * i for i in range(44)], seed=301)
    hreads = _reads(hd, [400, 1200, 3000] + [1000 + 131 * i for i in range(27)], seed=302)
    probes = [(dg, greads[i]) for i in (0, 2, 5)] + [(dh, hreads[i]) for i in (0, 2)]
    srv = native.Server(0, lanes=3, max_wave_reads=16)
    try:
        alone = []
        for dm, r in probes:
            t = srv.submit(dm, P.native_read(r), params, native.SA_FP32)
            (res, tm), = _collect(srv, [t])
            assert tm["wave_reads"] == 1
            alone.append(res)
        mix = [(dg, r) for r in greads] + [(dh, r) for r in hreads]
        order = np.random.default_rng(9).permutation(len(mix))
        tickets = {}
        for i in order:
            dm, r = mix[i]
            tickets[i] = srv.submit(dm, P.native_read(r), params, native.SA_FP32)
        got = dict(zip(order, _collect(srv, [tickets[i] for i in order])))
        waves = {tm["wave"] for _, tm in got.values()}
        assert len(waves) >= 3 and max(tm["wave_reads"] for _, tm in got.values()) > 1
        idx = [0, 2, 5, len(greads) + 0, len(greads) + 2]
        for a, i in zip(alone, idx):
            b, tm = got[i]
            assert np.array_equal(a.pairs, b.pairs), f"read {i}: pairs differ alone vs in wave of {tm['wave_reads']}"
            assert a.log_likelihood == b.log_likelihood
            assert a.cells == b.cells
        st = srv.stats()
        print(st)
        assert st["reads"] == len(probes) + len(mix)
    finally:
        srv.close()


def test_batch_submission_equals_single(models_):
    """Reads queued by one sa_server_submit_batch call give the results they give
    submitted one by one (FP32, exact equality); a refused read in the batch fails the
    call and queues nothing."""
    g6, hd, dg, dh = models_
    params = native.Params(diagonal_expansion=50)
    rs = _reads(g6, [600 + 173 * i for i in range(150)], seed=306)
    srv = native.Server(0, lanes=2, max_wave_reads=64)
    try:
        one = [r for r, _ in _collect(srv, [srv.submit(dg, P.native_read(r), params, native.SA_FP32) for r in rs])]
        objs = [P.native_read(r) for r in rs]
        tickets = srv.submit_structs(dg, [o._struct() for o in objs], params._struct(), native.SA_FP32)
        assert len(tickets) == len(rs) and len(set(tickets)) == len(rs)
        many = _collect(srv, tickets)
        assert max(tm["wave_reads"] for _, tm in many) > 1
        for a, (b, _) in zip(one, many):
            assert np.array_equal(a.pairs, b.pairs) and a.log_likelihood == b.log_likelihood
        st0 = srv.stats()
        bad = native.Read(rs[0]["ref"], rs[0]["events"], rs[0]["anchors"], rs[0]["scale"], rs[0]["shift"],
                          rs[0]["var"])
        bs = bad._struct()
        bs.event_stride = 1  # refused: fewer than two values per event
        with pytest.raises(native.SignalAlignError) as e:
            srv.submit_structs(dg, [objs[1]._struct(), bs], params._struct(), native.SA_FP32)
        assert e.value.code == -1
        assert srv.stats()["pending"] == 0 and srv.stats()["reads"] == st0["reads"]
    finally:
        srv.close()


def test_served_results_match_oracle(models_):
    """FP64 requests through the server against the FP64 oracle (the drop-in bar)."""
    g6, hd, dg, dh = models_
    params = native.Params(diagonal_expansion=20)
    rs = _reads(g6, [800, 2000], seed=303)
    srv = native.Server(0, lanes=2)
    try:
        tickets = [srv.submit(dg, P.native_read(r), params, native.SA_FP64) for r in rs]
        for r, (res, _) in zip(rs, _collect(srv, tickets)):
            want, st = P.run_oracle(g6, r, params)
            P.compare(res.pairs, want, params.threshold)
            assert P.rel(res.log_likelihood, st["log_likelihood"]) <= P.LL_REL
    finally:
        srv.close()


def test_admission_respects_hbm_budget(models_):
    """A budget that holds about three reads' band matrices: waves stay small, all
    requests still finish, and the bytes in flight never exceed the budget by more
    than one read (a read larger than the budget runs alone)."""
    g6, hd, dg, dh = models_
    params = native.Params(diagonal_expansion=50)
    rs = _reads(g6, [3000] * 24, seed=304)
    srv = native.Server(0, hbm_budget_bytes=3 * 3000 * 110 * 6 + (1 << 20), lanes=3)
    try:
        tickets = [srv.submit(dg, P.native_read(r), params, native.SA_FP32) for r in rs]
        got = _collect(srv, tickets)
        assert all(res.cells > 0 for res, _ in got)
        st = srv.stats()
        print(st)
        assert st["max_wave_reads"] <= 4
        assert st["waves"] >= 6
    finally:
        srv.close()


def test_bad_read_fails_alone(models_):
    """A read with a k-mer outside the alphabet fails with its own status; the reads
    batched with it still succeed."""
    g6, hd, dg, dh = models_
    params = native.Params(diagonal_expansion=20)
    rs = _reads(g6, [600, 700, 800], seed=305)
    bad = dict(rs[1])
    bad["ref"] = rs[1]["ref"][:50] + b"N" + rs[1]["ref"][51:]
    srv = native.Server(0, lanes=1)
    try:
        srv.submit(dg, P.native_read(rs[0]), params)          # occupies the lane
        tickets = [srv.submit(dg, P.native_read(r), params) for r in (rs[0], bad, rs[2])]
        assert srv.wait(tickets[0], 120.0) is not None
        with pytest.raises(native.SignalAlignError) as e:
            srv.wait(tickets[1], 120.0)
        assert e.value.code == -4
        assert srv.wait(tickets[2], 120.0)[0].cells > 0
    finally:
        srv.close()


@pytest.mark.gpu
def test_served_degenerate_reads(gpu_ctx, models_):
    """Degenerate references ('X' at CpG cytosines over "CE", FP64) are served in
    waves of their own beside plain requests: each result (pairs, paths,
    log-likelihood) equals the read's own batched run bit for bit; an FP32 request
    with 'X' is refused at submission."""
    import re
    g6, hd, dg, dh = models_
    params = native.Params(diagonal_expansion=20)
    rs = _reads(g6, [900, 1400, 700, 1100], seed=411)
    xr = [dict(r, ref=re.sub(b"[CE](?=G)", b"X", r["ref"])) for r in rs[:3]]

    def dread(r):
        return native.Read(r["ref"], r["events"], r["anchors"], r["scale"], r["shift"], r["var"],
                           degenerate_options="CE")

    srv = native.Server(0, lanes=2)
    try:
        print("serve degenerate", flush=True)
        tickets = [srv.submit(dg, dread(r), params, native.SA_FP64) for r in xr]
        plain = [srv.submit(dg, P.native_read(r), params, native.SA_FP64) for r in rs[2:]]
        got = _collect(srv, tickets)
        got_plain = _collect(srv, plain)
        with pytest.raises(native.SignalAlignError) as e:
            srv.submit(dg, dread(xr[0]), params, native.SA_FP32)
        assert e.value.code == native.SA_ERR_UNSUPPORTED
    finally:
        srv.close()
    for r, (res, _) in zip(xr, got):
        want, _ = gpu_ctx.align_batch(dg, [dread(r)], params, native.SA_FP64)
        assert np.array_equal(res.pairs, want[0].pairs) and np.array_equal(res.paths, want[0].paths)
        assert res.log_likelihood == want[0].log_likelihood
        assert res.paths is not None and res.paths.max() > 0
    for r, (res, _) in zip(rs[2:], got_plain):
        want, _ = gpu_ctx.align_batch(dg, [P.native_read(r)], params, native.SA_FP64)
        assert np.array_equal(res.pairs, want[0].pairs) and res.paths is None
