"""The profile builder's Baum-Welch on the GPU (bw.hip, dm_bw_estep) against
the oracle's E-step (dummest-build.cc:561-736), and the builder end to end:
samples/MET.msa must give samples/MET-test.hmm."""
import gzip
import os
import subprocess

import numpy as np
import pytest

import dummest_amd as D
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "dummest_amd", "dummest-amd-build")
GOLD = os.path.join(ROOT, "tests", "golden")


def bw_case(rng, A, L):
    p = np.zeros((L + 1, 7 + A))
    for r in range(L + 1):
        ins, ext, de, dext = rng.uniform(0.01, 0.3), rng.uniform(0.1, 0.7), rng.uniform(0.01, 0.2), rng.uniform(0.1, 0.7)
        p[r, :7] = [(1 - ins) * (1 - de), ins, de, 1 - ext, ext, 1 - dext, dext]
        p[r, 7:] = rng.dirichlet(np.ones(A))
    p[0, 5], p[0, 6] = 1, 0
    p[L, 0], p[L, 2] = 1 - p[L, 1], 0
    p[L, 7:] = rng.dirichlet(np.ones(A) * 5)
    return p


# profile lengths around the 64-row strips; empty and one-letter sequences;
# letter code A (outside the alphabet)
CASES = [(20, 1, [0, 1, 7]), (20, 63, [50, 0, 120]), (20, 64, [64, 65, 1]), (20, 65, [200, 3]),
         (20, 130, [150, 151, 40, 0, 260]), (4, 70, [90, 33]), (20, 300, [310, 280])]


@pytest.mark.parametrize("A,L,lens", CASES)
def test_estep_matches_oracle(engine, A, L, lens):
    rng = np.random.default_rng(L * 7 + A)
    p = bw_case(rng, A, L)
    seqs = [rng.integers(0, A + 1, n).astype(np.uint8) for n in lens]
    w = rng.uniform(0.2, 1.0, len(lens))
    print("bw case", A, L, lens, flush=True)
    engine.bw_prepare(A, L, seqs, w)
    got, tot = engine.bw_estep(p)
    want, wtot = O.bw_estep(L, A, p, np.concatenate(seqs) if seqs else [], lens, w)
    # the forward total is summed row by row on the device (cell by cell in
    # the reference): the per-sequence factor differs by rounding only
    assert np.allclose(tot, wtot, rtol=1e-13, atol=0)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-12 * scale
    assert (got[:, 2] == 0).all()


def test_estep_batches_give_the_same_counts(engine):
    rng = np.random.default_rng(9)
    A, L = 20, 90
    p = bw_case(rng, A, L)
    lens = [120, 40, 200, 0, 77, 150]
    seqs = [rng.integers(0, A, n).astype(np.uint8) for n in lens]
    w = rng.uniform(0.2, 1.0, len(lens))
    engine.bw_prepare(A, L, seqs, w)
    one, t1 = engine.bw_estep(p)
    engine.set_option("bw_budget", 6 * 8 * (L + 2) * 130)  # forces several batches
    try:
        engine.bw_prepare(A, L, seqs, w)
        many, t2 = engine.bw_estep(p)
    finally:
        engine.set_option("bw_budget", 8 << 30)
    assert np.array_equal(one, many) and np.array_equal(t1, t2)


def body(text):
    return [ln for ln in text.splitlines() if not ln.startswith(("HMMER3/f", "BM "))]


def test_met_msa_gives_reference_profiles():
    print("dummest-amd-build MET.msa", flush=True)
    r = subprocess.run([BUILD, os.path.join(GOLD, "MET.msa")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    want = gzip.open(os.path.join(GOLD, "MET-test.hmm.gz"), "rt").read()
    assert body(r.stdout) == body(want)
    assert "MSA #1: metC" in r.stderr and "MSA #2: metG" in r.stderr
    assert r.stderr.count("Baum-Welch converged") == 2
    for k, (name, M) in enumerate([("metC", 467), ("metG", 492)]):
        m, values, params, bg, count = D.parse_hmm_text(r.stdout, index=k)
        assert m == M and count == 2
