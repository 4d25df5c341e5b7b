"""The CPU-side C code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
tests): the oracle (oracle/asan_check.c drives its entry points: Gaussian, split
blocks and several windows, HDP, degenerate references, EM expectations, band and
split geometry) and signalMachine's host pieces in sa_machine.c and sa_compat.c
(tests/compat/machine_check.c: read scaling, the TSV writer, CIGAR anchors, the
.nhdp reader, the solver, the overlap filter and the sort).  Built by
`make -C oracle asan` with gcc; any memory error or undefined behaviour aborts the
run and fails the test."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_DIR = os.path.join(ROOT, "oracle", "_asan")
NPREAD = os.path.join(GOLDEN, "ZymoC_ch_1_file1.npRead")
TMODEL = os.path.join(GOLDEN, "testModelR73_acegot_template.model")
CMODEL = os.path.join(GOLDEN, "testModelR73_acegot_complement.model")
REF = os.path.join(GOLDEN, "ZymoRef.txt")
TSV = os.path.join(GOLDEN, "ZymoC_ch_1_file1.sm.forward.tsv")
NHDP = os.path.join(GOLDEN, "templateSingleLevelFixed.nhdp")


@pytest.fixture(scope="module")
def built():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"])


def _run(args, stdin=None, leaks=True):
    env = dict(os.environ)
    # the machine check links the HIP runtime, whose own allocations live until exit
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=1" if leaks else "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run(args, input=stdin, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, f"{args[:2]} exited {r.returncode}: {r.stderr[-3000:]}"
    assert "runtime error" not in r.stderr, r.stderr[-3000:]
    return r.stdout


def test_oracle_under_sanitizers(built):
    out = _run([os.path.join(REF_DIR, "asan_check")])
    lines = dict((l.split()[0], l) for l in out.splitlines() if l)
    for k in ("gauss", "split", "hdp", "degenerate", "expectations", "band"):
        assert k in lines and " rc 0 " in lines[k] + " ", out


def test_machine_host_code_under_sanitizers(built, tmp_path):
    chk = os.path.join(REF_DIR, "machine_check_asan")
    assert _run([chk, "params", NPREAD, TMODEL, CMODEL], leaks=False).count("\n") == 2
    _run([chk, "rewrite", NPREAD, TMODEL, CMODEL, REF, TSV, str(tmp_path / "out.tsv")], leaks=False)
    assert open(tmp_path / "out.tsv").read() == open(TSV).read()
    line = "cigar: r 0 700 + ref 0 720 + 1 M 300 D 20 M 400\n"
    assert "a " in _run([chk, "anchors", "14", NPREAD, "1", "t"], stdin=line, leaks=False)
    assert _run([chk, "nhdp", NHDP, TMODEL, "0", "5", "4095"], leaks=False).startswith("h ")
    _run([chk, "solve", "2", "0", "1", "1", "0", "2", "3"], leaks=False)
    _run([chk, "filter", "1", "1", "2", "2", "2", "3", "4", "3"], leaks=False)
    _run([chk, "sort", "1", "5", "5", "2", "3", "3", "3", "4", "6"], leaks=False)
