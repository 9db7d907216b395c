"""The profile builder (dummest-amd-build; dummest-build, dummest-build.cc).

Pinned by the reference's own sample: samples/MET.msa built with default
options gives samples/MET-test.hmm (two profiles, 959 positions, every value
printed to 5 decimals) apart from the program name in the header and the BM
line. A closed-form case (gapless, --countonly --pnone) pins the sequence
weights, the emission estimate, the transition and background rows, EFFN and
the number format on their own."""
import gzip
import math
import os
import subprocess

import numpy as np
import pytest

import dummest_amd as D

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "dummest_amd", "dummest-amd-build")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(*args, stdin=None, timeout=120):
    return subprocess.run([BUILD, *args], input=stdin, capture_output=True, text=True, timeout=timeout)


def body(text):
    return [ln for ln in text.splitlines() if not ln.startswith(("HMMER3/f", "BM "))]


def test_met_msa_gives_reference_profiles():
    r = run(os.path.join(GOLD, "MET.msa"), timeout=300)
    assert r.returncode == 0, r.stderr
    want = gzip.open(os.path.join(GOLD, "MET-test.hmm.gz"), "rt").read()
    assert body(r.stdout) == body(want)
    assert "MSA #1: metC" in r.stderr and "MSA #2: metG" in r.stderr
    assert r.stderr.count("Baum-Welch converged") == 2
    # the engine reads what the builder writes
    for k, (name, M) in enumerate([("metC", 467), ("metG", 492)]):
        m, values, params, bg, count = D.parse_hmm_text(r.stdout, index=k)
        assert m == M and count == 2


ALN = """# STOCKHOLM 1.0
#=GF ID tiny

s1 ACDEF
s2 ACDEF
s3 ACDEG
//
"""


def test_gapless_countonly_pnone_closed_form():
    r = run("--countonly", "--pnone", "-", stdin=ALN)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    # Henikoff weights: columns 1-4 give each sequence 3/(1*3) = 1; column 5
    # gives F rows 3/(2*2) and the G row 3/(2*1); mean over 5, scaled to max 1
    w = np.array([4.75 / 5, 4.75 / 5, 5.5 / 5])
    w /= w.max()
    assert f"EFFN  {w.sum():.6f}" in lines
    assert "NSEQ  3" in lines and "LENG  5" in lines
    aa = "ACDEFGHIKLMNPQRSTVWY"
    probs = np.zeros((5, 20))
    for col, letters in enumerate(["AAA", "CCC", "DDD", "EEE", "FFG"]):
        for s, c in enumerate(letters):
            probs[col, aa.index(c)] += w[s]
        probs[col] /= probs[col].sum()
    bg = np.exp(np.log(np.maximum(probs, 1e-6)).mean(0))
    bg /= bg.sum()
    fmt = lambda p: "        *" if p <= 0 else "  %.5f" % abs(math.log(p))
    i = lines.index("           match   insStart delStart insEnd   insExt   delEnd   delExt")
    rows = lines[i + 1:]
    for pos in range(6):
        assert rows[3 * pos] == "        " + "".join(fmt(x) for x in bg)
        # no gaps, no priors: match 1, insert and delete starts 0, the
        # 0/0 end/extend ratios 1/2; the first row's delete pair is (1, 0)
        trans = [1, 0, 0, 0.5, 0.5, 1 if pos == 0 else 0.5, 0 if pos == 0 else 0.5]
        if pos == 5:
            trans[0], trans[2] = 1, 0
        assert rows[3 * pos + 1] == "        " + "".join(fmt(x) for x in trans), pos
        if pos < 5:
            cons = aa[int(np.argmax(probs[pos]))]
            tail = "%7d %s - - -" % (pos + 1, cons)
            assert rows[3 * pos + 2] == "%7d " % (pos + 1) + "".join(fmt(x) for x in probs[pos]) + tail
    assert rows[17] == "//"


def test_counts_output_and_baum_welch_progress():
    r = run("--counts", "--countonly", "-", stdin=ALN)
    assert r.returncode == 0
    assert "notIns    delStart  insEnd    insExt    delEnd    delExt" in r.stdout
    assert "Iteration" not in r.stderr
    r2 = run("-", stdin=ALN)
    assert "Baum-Welch converged" in r2.stderr


def test_hand_needs_rf():
    r = run("--hand", "-", stdin=ALN)
    assert r.returncode == 1 and "missing RF annotation line: needed for --hand" in r.stderr
    with_rf = ALN.replace("//", "#=GC RF xx.xx\n//")
    r = run("--hand", "--countonly", "-", stdin=with_rf)
    assert r.returncode == 0 and "LENG  4" in r.stdout and "RF    yes" in r.stdout


def test_symfrac_and_bm_line():
    gappy = "# STOCKHOLM 1.0\n\ns1 AC-EF\ns2 AC-EF\ns3 ACDEF\n//\n"
    assert "LENG  4" in run("--countonly", "-", stdin=gappy).stdout
    r = run("--countonly", "--symfrac", "0.2", "-", stdin=gappy)
    assert "LENG  5" in r.stdout
    assert ("BM    %s --countonly --symfrac 0.2" % BUILD) in r.stdout


def test_malformed_alignment_is_fatal():
    bad = "# STOCKHOLM 1.0\n\ns1 ACDEF\ns2 ACDEF\n\ns1 AC\n//\n"
    r = run("-", stdin=bad)
    assert r.returncode != 0 and "FATAL: alignment blocks should have same number of sequences" in r.stderr
    bad2 = "# STOCKHOLM 1.0\n\ns1 ACDEF\ns2 ACD\n//\n"
    r = run("-", stdin=bad2)
    assert r.returncode != 0 and "FATAL: aligned sequences should have the same length" in r.stderr


@pytest.mark.parametrize("args", [["--symfrac", "1.5"], ["--ere", "0"], ["--esigma", "-1"], ["--maxiter", "0"],
                                  ["--maxdiff", "-1"]])
def test_bad_option_values(args):
    r = run(*args, "-", stdin=ALN)
    assert r.returncode == 1 and "bad option value" in r.stderr


def test_usage():
    assert run().returncode == 1
    r = run("-h")
    assert r.returncode == 0 and "--countonly" in r.stdout
