"""signalMachine's driver around the path (signalalign_machine.h) and the
signalMachine executable: read scaling, guide alignment -> anchors, TSV writers.

The reference's own output file for the Zymo read (its "full" TSV,
tests/test_alignments/zymo_C_test_alignments_sm3/.../ch1_file1_strand.sm.forward.tsv,
copied to tests/golden) pins the read-scaling estimate, the drift correction, the
noise scaling and the TSV writer together: rewriting its aligned pairs through
writePosteriorProbsFull must give the file back byte for byte.  The guide
alignment that produced it (bwa) is not in the reference, so the end-to-end GPU
run uses a guide alignment built from the file's confident pairs and is held to
the oracle on the same anchors.
"""
import os
import random
import subprocess

import numpy as np
import pytest

import oracle
from oracle import driver as D
from signalalign_amd import formats, models, native
from tests import _parity as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CHECK = os.path.join(ROOT, "tests", "compat", "machine_check")
MACHINE = os.path.join(ROOT, "signalalign_amd", "signalMachine")
NPREAD = os.path.join(GOLDEN, "ZymoC_ch_1_file1.npRead")
TMODEL = os.path.join(GOLDEN, "testModelR73_acegot_template.model")
CMODEL = os.path.join(GOLDEN, "testModelR73_acegot_complement.model")
REF = os.path.join(GOLDEN, "ZymoRef.txt")
TSV = os.path.join(GOLDEN, "ZymoC_ch_1_file1.sm.forward.tsv")


def _run(args, stdin=None, check=True):
    r = subprocess.run(args, input=stdin, capture_output=True, text=True, timeout=120)
    if check and r.returncode != 0:
        raise AssertionError(f"{args[:2]} exited {r.returncode}: {r.stderr[-2000:]}")
    return r


def _params():
    out = {}
    for line in _run([CHECK, "params", NPREAD, TMODEL, CMODEL]).stdout.split("\n"):
        t = line.split()
        if t:
            out[t[0]] = dict(zip(("n", "shift", "scale", "var", "drift", "shift_sd", "scale_sd", "var_sd"),
                                 [int(t[1])] + [float(v) for v in t[2:]]))
    return out


def test_read_scaling_within_the_reference_tolerances():
    """tests/stateMachineTests.c:408-418 (test_nanoporeScaleParamsFromStrandRead):
    the strand-read one-D assignments re-estimate the template's scale and shift
    within 5% and var within 50% of the values the .npRead carries."""
    r = formats.load_npread(NPREAD)
    got = _params()
    for s, name in (("t", "template"), ("c", "complement")):
        want = r[name]["params"]
        for key, pct in (("scale", 5.0), ("shift", 5.0), ("var", 50.0)):
            assert abs(got[s][key] - want[key]) / abs(want[key]) * 100 < pct, (s, key, got[s][key], want[key])
    # one assignment per distinct event in strand-read order (impl/nanopore.c:503-533)
    m = r["template"]["strand_event_map"][: len(r["template"]["read"]) - 5]
    assert got["t"]["n"] == int(np.sum(np.diff(np.concatenate([[-1], np.maximum.accumulate(m)])) > 0))


def test_writer_and_scaling_reproduce_the_reference_tsv(tmp_path):
    """All 2008 rows (template and complement) of the reference's full TSV for this
    read, byte for byte: reference k-mers and coordinates, drift-corrected event
    means, model means scaled by the re-estimated scale/shift, noise means scaled
    twice by scale_sd, posteriors, descaled means and path k-mers."""
    out = tmp_path / "rewrite.tsv"
    _run([CHECK, "rewrite", NPREAD, TMODEL, CMODEL, REF, TSV, str(out)])
    assert out.read_bytes() == open(TSV, "rb").read()


def test_reference_tsv_rows_follow_the_reference_kmers():
    """scripts/testSignalAlign.py:103-114: every row's k-mer is the reference's at
    its position, on both strands."""
    ref = open(REF).read().split()[0]
    for line in open(TSV):
        t = line.split("\t")
        assert t[2] == ref[int(t[1]):int(t[1]) + 6]


CIGARS = [
    # (line, trim, rebased (start1, end1, start2, end2), anchors by hand)
    ("cigar: r 0 10 + ref 5 15 + 1 M 10", 2, (0, 10, 0, 10), [(i, i) for i in range(2, 8)]),
    ("cigar: r 3 23 + ref 100 118 + 1 M 8 I 3 D 1 M 9", 1, (0, 18, 3, 23),
     [(i, i + 3) for i in range(1, 7)] + [(i, i + 5) for i in range(10, 17)]),
    ("cigar: r 0 10 + ref 20 10 - 1 M 10", 0, (0, 10, 0, 10), [(i, i) for i in range(10)]),
    ("cigar: r 0 6 + ref 0 6 + 1 M 3 D 2 I 2 M 1", 0, (0, 6, 0, 6), [(0, 0), (1, 1), (2, 2), (5, 5)]),
    ("cigar: r 0 4 + ref 0 4 + 1 M 4", 2, (0, 4, 0, 4), []),
]


@pytest.mark.parametrize("line,trim,span,want", CIGARS)
def test_guide_alignment_to_anchors(line, trim, span, want):
    out = _run([CHECK, "anchors", str(trim)], stdin=line + "\n").stdout.split("\n")
    pa = [int(v) for v in out[0].split()[1:]]
    assert (pa[0], pa[1], pa[3], pa[4]) == span and pa[2] == 1
    got = [tuple(int(v) for v in l.split()[1:]) for l in out[1:] if l.startswith("a ")]
    assert got == want
    assert D.guide_alignment_to_anchors(D.parse_cigar(line), trim)[1] == want


def test_guide_alignment_rejects_inconsistent_spans():
    r = _run([CHECK, "anchors", "0"], stdin="cigar: r 0 10 + ref 0 12 + 1 M 10\n", check=False)
    assert r.returncode != 0 and "reference span 12" in r.stderr


def _dominance_filter(pairs):
    """The reference's definition written out over all pairs: kept when strictly
    below-left of every later pair (through an equal pair that is) and strictly
    above-right of every earlier pair."""
    kept = []
    n = len(pairs)
    for i, (x, y) in enumerate(pairs):
        below = any(pairs[j] == (x, y) and all(pairs[k][0] > x and pairs[k][1] > y for k in range(j + 1, n))
                    for j in range(n))
        above = all(pairs[k][0] < x and pairs[k][1] < y for k in range(i))
        if below and above:
            kept.append((x, y))
    return kept


def test_filter_to_remove_overlap_random():
    rng = random.Random(7)
    for _ in range(60):
        n = rng.randint(0, 14)
        pairs = sorted((rng.randint(0, 9), rng.randint(0, 9)) for _ in range(n))
        args = [str(v) for p in pairs for v in p]
        got = [tuple(int(v) for v in l.split()) for l in _run([CHECK, "filter"] + args).stdout.split("\n") if l]
        assert got == D.filter_to_remove_overlap(pairs) == _dominance_filter(pairs), pairs


def test_remapped_anchors_on_the_read():
    """Anchors from a guide alignment of the Zymo read, remapped to each strand's
    events (the 2D event maps): the library's list equals the restatement's and is
    strictly increasing in both coordinates."""
    line = _guide_cigar()
    pa = D.parse_cigar(line)
    rd = formats.load_npread(NPREAD)
    for strand, name in (("t", "template"), ("c", "complement")):
        out = _run([CHECK, "anchors", "14", NPREAD, "1", strand], stdin=line + "\n").stdout.split("\n")
        got = [tuple(int(v) for v in l.split()[1:]) for l in out if l.startswith("r ")]
        _, anchors = D.guide_alignment_to_anchors(pa, 14)
        want = D.remap(anchors, rd[name]["event_map"], pa["start2"])
        assert got == want and len(got) > 100
        assert all(b[0] > a[0] and b[1] > a[1] for a, b in zip(got, got[1:]))


def test_lineq_solve():
    # a zero leading pivot takes the next usable row
    x = [float(v) for v in _run([CHECK, "solve", "2", "0", "1", "1", "0", "2", "3"]).stdout.split()]
    assert x == [3.0, 2.0]
    rng = np.random.default_rng(3)
    a = rng.normal(size=(3, 3))
    a = a @ a.T + 3 * np.eye(3)
    b = rng.normal(size=3)
    out = _run([CHECK, "solve", "3"] + [float(v).hex() for v in a.ravel()] + [float(v).hex() for v in b]).stdout.split()
    assert np.allclose([float(v) for v in out], np.linalg.solve(a, b), rtol=1e-13, atol=1e-13)


def test_sort_by_x_plus_y_is_stable():
    out = _run([CHECK, "sort", "1", "5", "5", "2", "3", "3", "3", "4", "6", "4", "6", "4", "5", "2", "2"]).stdout
    assert [int(l.split()[0]) for l in out.split("\n") if l] == [5, 2, 1, 3, 4]


def test_npread_c_reader_rejects_bad_files(tmp_path):
    lines = open(NPREAD).read().split("\n")
    bad = list(lines)
    bad[7] = " ".join(bad[7].split()[:-4])  # one template event short
    f = tmp_path / "bad.npRead"
    f.write_text("\n".join(bad))
    r = _run([CHECK, "params", str(f), TMODEL, CMODEL], check=False)
    assert r.returncode != 0 and "template events" in r.stderr


@pytest.mark.parametrize("extra,msg", [
    (["--sm3Hdp"], "needs --templateHdp"),
    (["--sm3Hdp", "--twoD", "-C", TMODEL, "-v", "t.nhdp"], "Need to have template and complement HDPs"),
    (["-p", "somewhere"], "error correction"),
])
def test_signal_machine_refuses_unserved_runs(extra, msg):
    r = _run([MACHINE, "-T", TMODEL, "-f", REF, "-b", REF, "-q", NPREAD] + extra, stdin="", check=False)
    assert r.returncode == 1 and msg in r.stderr


NHDP = os.path.join(GOLDEN, "templateSingleLevelFixed.nhdp")


def _hdp_expectations(h, dp):
    """dir_proc_expected_val / dir_proc_variance (impl/hdp.c:2772-2801), sequential sums."""
    g, d = h["grid"], h["density"][h["row_of"][dp]]
    e = models.expected_value(g, d)
    v = 0.0
    for i in range(1, len(g)):
        v += (g[i] - e) * (g[i] - e) * d[i] * (g[i] - g[i - 1])
    return e, np.sqrt(v)


def _nhdp_python():
    h = formats.load_nhdp(NHDP)
    # observed leaves are the k-mers whose row is their own (their first observed ancestor is themselves)
    rows = h["kmer_row"]
    own = [k for k in range(len(rows)) if np.sum(rows == rows[k]) == 1]
    h["row_of"] = {k: int(rows[k]) for k in own}
    return h, own


def test_nhdp_c_reader_matches_the_python_reader():
    """The C .nhdp reader (grid, observed DPs, densities) and
    stateMachine3_setModelToHdpExpectedValues against the Python reader: observed
    k-mers take their DP's expected value and standard deviation exactly, the
    others keep the .model file's."""
    h, own = _nhdp_python()
    m = models.load_model_file(TMODEL)
    pick = own[:6] + [0, 1, 46655]
    out = _run([CHECK, "nhdp", NHDP, TMODEL] + [str(k) for k in pick]).stdout.split("\n")
    head = out[0].split()
    assert head[1:4] == [str(h["n_dps"]), str(len(h["density"])), str(len(h["grid"]))]
    assert float(head[4]) == h["grid"][0] and float(head[5]) == h["grid"][-1]
    ks = [l.split() for l in out if l.startswith("k ")]
    ms = [l.split() for l in out if l.startswith("m ")]
    for k, kl, ml in zip(pick, ks, ms):
        if k in h["row_of"]:
            e, sd = _hdp_expectations(h, k)
            assert kl[2] == "1" and int(kl[3]) == h["row_of"][k]
            assert float(kl[4]) == e and float(kl[5]) == sd
            assert float(ml[2]) == e and float(ml[3]) == sd
        else:
            assert kl[2] == "0"
            assert float(ml[2]) == m.match[k][0] and float(ml[3]) == m.match[k][1]
    assert len(own) > 100


def test_signal_machine_needs_models():
    r = _run([MACHINE, "--twoD", "-T", TMODEL, "-q", NPREAD], stdin="", check=False)
    assert r.returncode == 1 and "Missing model files" in r.stderr


# ---------------------------------------------------------------------------
# end to end on the GPU
# ---------------------------------------------------------------------------
def _guide_cigar(event_map="event_map"):
    """A guide alignment of the 2D read to the reference from the reference TSV's
    confident template pairs (posterior >= 0.9): reference position against the
    first read base (2D read, or template read with event_map="strand_event_map")
    mapped to that event, thinned to increasing points at least 40
    reference bases apart, and the last confident pair."""
    rd = formats.load_npread(NPREAD)
    first = {}
    for b, e in enumerate(rd["template"][event_map]):
        first.setdefault(int(e), b)
    pts = sorted({(int(t[1]), first[int(t[5])]) for t in (l.split("\t") for l in open(TSV))
                  if t[4] == "t" and float(t[12]) >= 0.9 and int(t[5]) in first})
    chain = []
    for a, b in pts:  # one point per 40 reference bases or more, as a seed-and-extend aligner's blocks
        if not chain or (a >= chain[-1][0] + 40 and b > chain[-1][1]):
            chain.append((a, b))
    last = max(p for p in pts if p[0] > chain[-1][0] and p[1] > chain[-1][1])
    chain.append(last)
    ops = []

    def emit(op, n):
        if n <= 0:
            return
        if ops and ops[-1][0] == op:
            ops[-1][1] += n
        else:
            ops.append([op, n])

    # bases between two confident pairs are paired as matches (mismatches included,
    # as an aligner's M), the surplus on one side is an insertion or deletion
    emit("M", 1)
    for (pa, pb), (a, b) in zip(chain, chain[1:]):
        m = min(a - pa - 1, b - pb - 1)
        emit("M", m)
        emit("D", a - pa - 1 - m)
        emit("I", b - pb - 1 - m)
        emit("M", 1)
    ops = [f"{op} {n}" for op, n in ops]
    (a0, b0), (a1, b1) = chain[0], chain[-1]
    return f"cigar: read {b0} {b1 + 1} + ZYMO {a0} {a1 + 1} + 1 " + " ".join(ops)


def _complement(s):
    return s.translate(str.maketrans("ACGT", "TGCA"))


def _oracle_inputs(strand, pa, prm, rd, ref, event_map="event_map", hdp=None):
    """The model, read (target, drift-corrected event slice, remapped anchors, read
    parameters), first event and target of one strand of a signalMachine run."""
    name = "template" if strand == "t" else "complement"
    ev = rd[name]["events"].copy()
    ev[:, 0] = ev[:, 0] - ev[:, 3] * prm["drift"]
    emap = rd[name][event_map]
    e0, e1 = int(emap[pa["start2"]]), int(emap[pa["end2"]])
    span = ref[pa["start1"]:pa["end1"]]
    target = span if strand == "t" else _complement(span)[::-1]
    _, anchors = D.guide_alignment_to_anchors(pa, 14)
    m = models.load_model_file(TMODEL if strand == "t" else CMODEL)
    m = models.scale_noise(m, prm["scale_sd"], prm["var_sd"])
    if hdp is not None:
        # descaling level means: the .model file's, observed k-mers' from the HDP
        level = m.match[:, 0].copy()
        for kk in hdp["row_of"]:
            level[kk] = _hdp_expectations(hdp, kk)[0]
        m = models.HdpModel(hdp["alphabet"], hdp["kmer_length"], m.transitions, hdp["grid"], hdp["density"],
                            hdp["slopes"], hdp["kmer_row"], level)
    read = {"ref": target.encode(), "events": ev[e0:e1], "anchors": np.array(D.remap(anchors, emap, pa["start2"]),
                                                                             dtype=np.int64).reshape(-1, 2),
            "scale": prm["scale"], "shift": prm["shift"], "var": prm["var"]}
    return m, read, e0, target


def _expected(strand, pa, prm, rd, ref, k=6, event_map="event_map", hdp=None):
    """Oracle pairs in reference coordinates for one strand of the signalMachine run."""
    m, read, e0, target = _oracle_inputs(strand, pa, prm, rd, ref, event_map, hdp)
    pairs, _ = P.run_oracle(m, read, native.Params(threshold=0.01, diagonal_expansion=50),
                            kind=None if hdp is not None else oracle.GAUSS_DESCALED)
    L = len(target)
    out = []
    for p, x, y in pairs:
        xa = x + pa["start1"] if strand == "t" else (L - k) - (x + (L - pa["end1"]))
        out.append((int(p), int(xa), int(y) + e0))
    return out


@pytest.mark.gpu
def test_signal_machine_end_to_end(tmp_path, oracle_lib):
    """The signalMachine executable on the Zymo read (2D, both strands), through the
    library on the GPU (FP64): pairs and posteriors to the oracle's bar on the same
    anchors and re-estimated parameters, rows in the stable x+y order of the
    reference list, and the variantCaller / assignments outputs consistent with it."""
    line = _guide_cigar()
    ref = open(REF).read().split()[0]
    bwd = tmp_path / "bwd.txt"
    bwd.write_text(_complement(ref) + "\n")
    rd = formats.load_npread(NPREAD)
    prm = _params()
    pa = D.parse_cigar(line)
    label = "zymo_read"
    outs = {}
    for fmt in (0, 1, 2):
        out = tmp_path / f"out{fmt}.tsv"
        print(f"signalMachine -s {fmt}", flush=True)
        r = _run([MACHINE, "--twoD", "-s", str(fmt), "-T", TMODEL, "-C", CMODEL, "-f", REF, "-b", str(bwd),
                  "-q", NPREAD, "-u", str(out), "-L", label], stdin=line + "\n")
        outs[fmt] = out.read_text().split("\n")[:-1] if out.exists() else []
        summary = r.stdout.split()
    rows = [l.split("\t") for l in outs[0]]
    assert summary[0] == label
    for strand in ("t", "c"):
        got = [(int(round(float(t[12]) * 1e7)), int(t[1]), int(t[5])) for t in rows if t[4] == strand]
        want = _expected(strand, pa, prm[strand], rd, ref)
        cmp = P.compare(got, want, 0.01, p_tol=1e-4 + 1e-6, check_order=False)
        print(strand, cmp)
        # the rows are the reference list sorted stably by x + y (in target coordinates)
        common = {(x, y) for _, x, y in got} & {(x, y) for _, x, y in want}
        assert len(common) == len(got) == len(want)
        if strand == "t":
            keyf = lambda q: q[1] + q[2]  # noqa: E731
        else:
            keyf = lambda q: -q[1] + q[2]  # x runs backwards on the complement's target  # noqa: E731
        assert [(x, y) for _, x, y in got] == [(x, y) for _, x, y in sorted(want, key=keyf)]
        for t in rows:
            if t[4] == strand:
                assert t[2] == ref[int(t[1]):int(t[1]) + 6]
    # canonical reference: the variant-caller output has no rows
    assert outs[1] == []
    # assignments: path k-mer, strand, descaled mean, posterior of every full row
    assert [l.split("\t") for l in outs[2]] == [[t[15], t[4], t[13], t[12]] for t in rows]


@pytest.mark.gpu
def test_signal_machine_hdp_end_to_end(tmp_path, oracle_lib):
    """signalMachine --sm3Hdp on the Zymo read's template strand (1D: the guide
    alignment in template-read coordinates) with the reference's HDP: read scaling
    from the .model table, level means of observed k-mers set to the HDP's expected
    values, emissions from the density grid; pairs to the oracle's bar (FP64)."""
    line = _guide_cigar("strand_event_map")
    ref = open(REF).read().split()[0]
    bwd = tmp_path / "bwd.txt"
    bwd.write_text(_complement(ref) + "\n")
    rd = formats.load_npread(NPREAD)
    prm = _params()
    pa = D.parse_cigar(line)
    h, _ = _nhdp_python()
    out = tmp_path / "hdp.tsv"
    print("signalMachine --sm3Hdp", flush=True)
    _run([MACHINE, "--sm3Hdp", "-v", NHDP, "-s", "0", "-T", TMODEL, "-f", REF, "-b", str(bwd), "-q", NPREAD,
          "-u", str(out), "-L", "zymo_read"], stdin=line + "\n")
    rows = [l.split("\t") for l in out.read_text().split("\n")[:-1]]
    got = [(int(round(float(t[12]) * 1e7)), int(t[1]), int(t[5])) for t in rows]
    want = _expected("t", pa, prm["t"], rd, ref, event_map="strand_event_map", hdp=h)
    cmp = P.compare(got, want, 0.01, p_tol=1e-4 + 1e-6, check_order=False)
    print("hdp", cmp)
    assert cmp["n_gpu"] > 500 and all(t[4] == "t" for t in rows)
    for t in rows:
        assert t[2] == ref[int(t[1]):int(t[1]) + 6]


def _read_expectations(path):
    return [l.split("\t") for l in open(path).read().split("\n")]


@pytest.mark.gpu
def test_signal_machine_expectations(tmp_path, oracle_lib):
    """signalMachine -t/-c (the expectations routine) on the Zymo read (2D): each
    strand's expectations file in the reference's layout, its values equal to the
    oracle's getExpectationsUsingAnchors on the same inputs (1e-9 relative; the file
    prints six decimals), the event model the noise-scaled table."""
    line = _guide_cigar()
    ref = open(REF).read().split()[0]
    bwd = tmp_path / "bwd.txt"
    bwd.write_text(_complement(ref) + "\n")
    rd = formats.load_npread(NPREAD)
    prm = _params()
    pa = D.parse_cigar(line)
    tf, cf = tmp_path / "t.expectations", tmp_path / "c.expectations"
    print("signalMachine expectations", flush=True)
    _run([MACHINE, "--twoD", "-T", TMODEL, "-C", CMODEL, "-f", REF, "-b", str(bwd), "-q", NPREAD, "-L", "zymo",
          "-t", str(tf), "-c", str(cf)], stdin=line + "\n")
    for strand, f in (("t", tf), ("c", cf)):
        rows = _read_expectations(f)
        assert rows[0][:4] == ["3", "4", "ACGT", "6"] or rows[0][:4] == ["3", "6", "ACEGOT", "6"], rows[0][:4]
        m, read, e0, target = _oracle_inputs(strand, pa, prm[strand], rd, ref)
        om = P.oracle_model(m, read, oracle.GAUSS_DESCALED)
        want = oracle.get_expectations(om, read["ref"], read["events"], read["anchors"],
                                       oracle.Params(threshold=0.01, diagonal_expansion=50))
        vals = [float(v) for v in rows[1] if v]
        assert len(vals) == 10
        assert np.allclose(vals[:9], want["transitions"].ravel(), rtol=1e-9, atol=1e-6)
        assert abs(vals[9] - want["likelihood"]) <= 1e-9 * abs(want["likelihood"]) + 1e-6
        model = np.array([float(v) for v in rows[2] if v])
        assert np.allclose(model, m.match.ravel(), rtol=0, atol=1e-6)
        eexp = np.array([float(v) for v in rows[3] if v]).reshape(-1, 2)
        assert np.allclose(eexp, want["event_exp"], rtol=1e-9, atol=1e-6)
        post = np.array([float(v) for v in rows[4] if v])
        assert np.allclose(post, want["posteriors"], rtol=1e-9, atol=1e-6)
        obs = np.array([int(v) for v in rows[5] if v]).astype(bool)
        assert np.array_equal(obs, want["observed"])
        print(strand, "likelihood", vals[9], "observed", int(obs.sum()))


@pytest.mark.gpu
def test_signal_machine_hdp_expectations(tmp_path, oracle_lib):
    """signalMachine --sm3Hdp -t on the template strand (1D): the HDP expectations
    file's transitions and its assignments (event mean, k-mer of every match
    transition with posterior >= the threshold, in the reference's call order) equal
    the oracle's."""
    line = _guide_cigar("strand_event_map")
    ref = open(REF).read().split()[0]
    bwd = tmp_path / "bwd.txt"
    bwd.write_text(_complement(ref) + "\n")
    rd = formats.load_npread(NPREAD)
    prm = _params()
    pa = D.parse_cigar(line)
    h, _ = _nhdp_python()
    tf = tmp_path / "t.expectations"
    print("signalMachine --sm3Hdp expectations", flush=True)
    _run([MACHINE, "--sm3Hdp", "-v", NHDP, "-T", TMODEL, "-f", REF, "-b", str(bwd), "-q", NPREAD, "-L", "zymo",
          "-t", str(tf)], stdin=line + "\n")
    rows = _read_expectations(tf)
    # the expectations routine does not move the level means to the HDP's (the
    # alignment routine does): the oracle model keeps the .model file's
    m, read, e0, target = _oracle_inputs("t", pa, prm["t"], rd, ref, "strand_event_map")
    hm = models.HdpModel(h["alphabet"], h["kmer_length"], m.transitions, h["grid"], h["density"], h["slopes"],
                         h["kmer_row"], m.match[:, 0].copy())
    om = P.oracle_model(hm, read)
    want = oracle.get_expectations(om, read["ref"], read["events"], read["anchors"],
                                   oracle.Params(threshold=0.01, diagonal_expansion=50), rec_threshold=0.01)
    vals = [float(v) for v in rows[1] if v]
    assert np.allclose(vals[:9], want["transitions"].ravel(), rtol=1e-9, atol=1e-6)
    ev = [float(v) for v in rows[3] if v]
    km = [v for v in rows[4] if v]
    wev, wkm = [], []
    for x, y, p1, p2, p3 in want["records"]:
        for pj in (p1, p2, p3):
            if pj >= 0.01:
                wev.append(read["events"][int(y)][0])
                wkm.append(target[int(x):int(x) + 6])
    assert km == wkm and len(km) > 100
    assert np.allclose(ev, wev, rtol=0, atol=1e-6)
    print("assignments", len(km))


@pytest.mark.gpu
def test_signal_machine_degenerate_variant_calls(tmp_path, oracle_lib):
    """signalMachine on the Zymo read (2D) against ZymoRef with the cytosine of every
    CpG made 'X' (forward and backward reference files), --degenerate 1 (C, E, O):
    the full output's pairs, posteriors and path k-mers equal the oracle's degenerate
    run on the same anchors and parameters, in the stable x+y order of its list; the
    variantCaller output is the reference writer's rows for the 'X' positions of those
    pairs (writePosteriorProbsVC, signalMachine.c:135-199: the called base is the path
    k-mer's option at the 'X', at the reference position of the 'X'); the assignments
    output carries each pair's path k-mer."""
    import re
    line = _guide_cigar()
    ref = re.sub("C(?=G)", "X", open(REF).read().split()[0])
    fwd = tmp_path / "fwd.txt"
    fwd.write_text(ref + "\n")
    bwd = tmp_path / "bwd.txt"
    bwd.write_text(_complement(ref) + "\n")
    rd = formats.load_npread(NPREAD)
    prm = _params()
    pa = D.parse_cigar(line)
    outs = {}
    for fmt in (0, 1, 2):
        out = tmp_path / f"out{fmt}.tsv"
        print(f"signalMachine degenerate -s {fmt}", flush=True)
        _run([MACHINE, "--twoD", "-s", str(fmt), "-o", "1", "-T", TMODEL, "-C", CMODEL, "-f", str(fwd), "-b",
              str(bwd), "-q", NPREAD, "-u", str(out), "-L", "zymo_read"], stdin=line + "\n")
        outs[fmt] = [l.split("\t") for l in out.read_text().split("\n")[:-1]]
    rows = outs[0]
    k = 6
    n_x = 0
    for strand in ("t", "c"):
        m, read, e0, target = _oracle_inputs(strand, pa, prm[strand], rd, ref)
        want, _ = oracle.get_aligned_pairs(P.oracle_model(m, read, oracle.GAUSS_DESCALED), read["ref"],
                                           read["events"], read["anchors"],
                                           oracle.Params(threshold=0.01, diagonal_expansion=50), options="CEO")
        L = len(target)
        wl = []
        for p, x, y, path in want:
            xa = x + pa["start1"] if strand == "t" else (L - k) - (x + (L - pa["end1"]))
            wl.append((int(xa), int(y) + e0, oracle.path_kmers(target[int(x):int(x) + k], "CEO")[int(path)],
                       int(p) / 1e7))
        gl = [(int(t[1]), int(t[5]), t[15], float(t[12])) for t in rows if t[4] == strand]
        g = {q[:3]: q[3] for q in gl}
        w = {q[:3]: q[3] for q in wl}
        assert len(g) == len(gl)
        only_g = [q for q in g if q not in w and abs(g[q] - 0.01) > 1e-3]
        only_w = [q for q in w if q not in g and abs(w[q] - 0.01) > 1e-3]
        assert not only_g and not only_w, (strand, only_g[:5], only_w[:5])
        common = [q for q in g if q in w]
        dmax = max(abs(g[q] - w[q]) for q in common)
        assert dmax <= 1e-4 + 1e-6, dmax  # the TSV prints six decimals
        keyf = (lambda q: q[0] + q[1]) if strand == "t" else (lambda q: -q[0] + q[1])
        cs = set(common)
        assert [q[:3] for q in gl if q[:3] in cs] == [q[:3] for q in sorted(wl, key=keyf) if q[:3] in cs]
        n_x += sum("X" in t[2] for t in rows if t[4] == strand)
        print(strand, len(gl), len(wl), dmax)
    assert n_x > 0
    # variantCaller rows from the full rows, in order
    vc = []
    for t in rows:
        refk, pathk, strand = t[2], t[15], t[4]
        for q in range(k):
            if refk[q] == "X":
                base = pathk[q] if strand == "t" else pathk[k - 1 - q]
                vc.append([t[5], str(int(t[1]) + q), base, t[12], strand, "forward", "zymo_read"])
    assert outs[1] == vc
    assert {v[2] for v in vc} <= set("CEO") and len({v[2] for v in vc}) > 1
    # assignments: the path k-mer (with the called option), strand, descaled mean and
    # posterior of every full row
    assert outs[2] == [[t[15], t[4], t[13], t[12]] for t in rows]
    assert any("E" in t[15] or "O" in t[15] for t in rows)
