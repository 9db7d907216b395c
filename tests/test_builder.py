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


def has_gpu():
    try:
        D.Engine(0).close()
        return True
    except Exception:
        return False


def test_baum_welch_needs_the_gpu():
    # Baum-Welch runs only on the device: without one the builder says so
    # (and --countonly, which needs none, still works)
    if has_gpu():
        pytest.skip("a HIP device is present")
    r = run(os.path.join(GOLD, "MET.msa"))
    assert r.returncode == 1 and "no usable HIP device for Baum-Welch" in r.stderr
    assert run("--countonly", os.path.join(GOLD, "MET.msa")).returncode == 0


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


def test_counts_output():
    r = run("--counts", "--countonly", "-", stdin=ALN)
    assert r.returncode == 0
    assert "notIns    delStart  insEnd    insExt    delEnd    delExt" in r.stdout
    assert "Iteration" not in r.stderr


def _bw_case(rng, A, L):
    p = np.zeros((L + 1, 7 + A))
    for r in range(L + 1):
        ins, ext, de, dext = rng.uniform(0.05, 0.3), rng.uniform(0.1, 0.6), rng.uniform(0.02, 0.2), rng.uniform(0.1, 0.6)
        p[r, :7] = [(1 - ins) * (1 - de), ins, de, 1 - ext, ext, 1 - dext, dext]
        p[r, 7:] = rng.dirichlet(np.ones(A))
    p[0, 5], p[0, 6] = 1, 0
    p[L, 0], p[L, 2] = 1 - p[L, 1], 0
    p[L, 7:] = rng.dirichlet(np.ones(A) * 5)
    return p


@pytest.mark.parametrize("A", [20, 4])
def test_oracle_estep_counts_are_gradients(A):
    # Expected counts are derivatives of the forward total: the emission count
    # of (row r, letter x) is d ln v / d ln S[r][x] (S proportional to the
    # letter probability), and for the insert extension probability e of row r
    # d ln v / d ln e = extensions - starts * e / (1 - e). Central differences
    # of the oracle's forward total pin its backward and count passes.
    import oracle as O
    rng = np.random.default_rng(40 + A)
    L, n = 5, 12
    p = _bw_case(rng, A, L)
    seq = rng.integers(0, A, n)
    c, t = O.bw_estep(L, A, p, seq, [n], [1.0])
    eps = 1e-4

    def dlog(r, col):
        pp, pm = p.copy(), p.copy()
        pp[r, col] *= 1 + eps
        pm[r, col] *= 1 - eps
        return (np.log(O.bw_estep(L, A, pp, seq, [n], [1.0])[1][0]) -
                np.log(O.bw_estep(L, A, pm, seq, [n], [1.0])[1][0])) / (2 * eps)

    scale = np.abs(c).max()
    for r in range(L):
        for x in sorted(set(seq.tolist())):
            assert abs(dlog(r, 7 + x) - c[r, 7 + x]) < 1e-8 * scale, (r, x)
    for r in range(L + 1):
        e = p[r, 4]
        assert abs(dlog(r, 4) - (c[r, 4] - c[r, 3] * e / (1 - e))) < 1e-8 * scale, r
    # weights scale the counts linearly; sequences add
    c2, _ = O.bw_estep(L, A, p, np.concatenate([seq, seq]), [n, n], [0.25, 0.5])
    assert np.allclose(c2, 0.75 * c, rtol=1e-13, atol=0)


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
