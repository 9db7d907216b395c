"""dm_pairs_hits -- the post-gate path of many pairs at once on the device
(dump-mode DP, per-column pass, flank totals, MEA tables and tracebacks in
batched kernels) -- against dm_pair_hits pair by pair (the host restatement
post.cpp, checked against the oracle in test_gpu_parity.py): identical hit
lists, bit for bit (scores, anchors, segments), over profiles of one lane
group and of row strips, sequences of 60 to 2,500 letters with zero, one or
two embedded homologs, thresholds from 0 (every anchor a candidate, many
tracebacks rejected for overlap) to above every score, and negative
thresholds (the best anchor only). Also against the oracle directly, the
reference's herpes x medium fixture, and the DM_ERANGE retry."""
import gzip
import os
import tempfile
import time

import numpy as np
import pytest

import dummest_amd as D
import oracle as O
from dummest_amd import synth
from tests.common import GOLDEN, read_fasta, synthetic_case
from tests.test_gpu_parity import check_hits, oracle_cells, reported

pytestmark = pytest.mark.gpu


def add_case(engine, seed, L, lengths, homolog_every):
    text, prof, dnas = synthetic_case(seed, L, lengths, homolog_every)
    M, v, p, b, _ = D.parse_hmm_text(text)
    pid = engine.add_profile(M, v, p, b)
    codes = [D.decode(s) for s in dnas]
    return pid, prof, codes, engine.add_sequences(codes)


def two_copies(engine, seed, L, n):
    """A profile and a sequence carrying two copies of one of its homologs."""
    rng = np.random.default_rng(seed)
    text, prof, _ = synthetic_case(seed, L, [], 0)
    raw = O.parse_hmm(text)[0]
    ins = synth.reverse_translate(rng, np.array(raw.consensus, np.int64) % 20)
    k = (n - 2 * len(ins)) // 3
    dna = synth.random_dna(rng, k) + ins + synth.random_dna(rng, k) + ins + synth.random_dna(rng, k)
    M, v, p, b, _ = D.parse_hmm_text(text)
    pid = engine.add_profile(M, v, p, b)
    c = D.decode(dna)
    return pid, prof, c, engine.add_sequences([c])[0]


def assert_same_hits(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g == w, (g, w)


def test_batch_equals_per_pair(engine):
    P, S, R = [], [], []
    cases = [add_case(engine, 71, 40, [60, 300, 900], 1), add_case(engine, 72, 150, [1200, 2500, 500, 151], 2),
             add_case(engine, 73, 600, [1800, 700], 1)]
    dup = two_copies(engine, 74, 90, 1500)
    ratios = [float(np.exp(8.0)), float(np.exp(20.0)), 1.0, -1.0, float(np.exp(400.0))]
    q = 0
    for pid, _, _, sids in cases:
        for s in sids:
            for r in ratios:
                P.append(pid)
                S.append(s)
                R.append(r)
                q += 1
    P += [dup[0]] * 3
    S += [dup[3]] * 3
    R += [float(np.exp(8.0)), float(np.exp(20.0)), 0.0]
    # a short pair with threshold 0: every column's anchor is a candidate
    P.append(cases[0][0])
    S.append(cases[0][3][0])
    R.append(0.0)
    print("batched hits:", len(P), "pairs", flush=True)
    t = time.perf_counter()
    got = engine.find_similarities_batch(P, S, R)
    tb = time.perf_counter() - t
    t = time.perf_counter()
    want = [engine.find_similarities(p, s, r) for p, s, r in zip(P, S, R)]
    tp = time.perf_counter() - t
    print(f"batched {tb:.3f} s, pair by pair {tp:.3f} s; device {engine.last_hits_ms:.1f} ms", flush=True)
    nh = [len(w) for w in want]
    assert max(nh) >= 2 and min(nh) == 0  # multi-hit pairs and pairs below the gate
    for g, w in zip(got, want):
        assert_same_hits(g, w)


def test_batch_matches_oracle(engine):
    pid, prof, codes, sids = add_case(engine, 300, 150, [1500, 2500, 900], 1)
    thr = float(np.exp(8.0))
    got = engine.find_similarities_batch([pid] * len(sids), sids, thr)
    for c, g in zip(codes, got):
        ref = O.search(prof, c, thr)
        assert len(reported(ref.hits)) >= 1
        check_hits(g, ref.hits, oracle_cells(prof, c))


def test_herpes_medium_batch(engine):
    """The reference's overflow fixture (M = 1359, a 2-strip configuration):
    the batch equals the per-pair path, and the top hit's spans are the ones
    tests/test_overflow.sh:68 fingerprints."""
    with tempfile.NamedTemporaryFile("wb", suffix=".hmm", delete=False) as f:
        f.write(gzip.open(os.path.join(GOLDEN, "herpes.hmm.gz")).read())
        path = f.name
    try:
        pid = engine.read_hmm(path)[0]
    finally:
        os.unlink(path)
    c = D.decode(read_fasta(os.path.join(GOLDEN, "medium.fa"))[0][1])
    sid = engine.add_sequences([c])[0]
    thr = float(np.exp(20.0))
    print("herpes x medium, batched", flush=True)
    got = engine.find_similarities_batch([pid, pid], [sid, sid], [thr, -1.0])
    assert_same_hits(got[0], engine.find_similarities(pid, sid, thr))
    assert_same_hits(got[1], engine.find_similarities(pid, sid, -1.0))
    top = max((h for h in got[0] if h[5]), key=lambda h: h[0])
    segs = top[5]
    assert (segs[0][0], segs[-1][0] + segs[-1][2] - segs[0][0], segs[0][1],
            segs[-1][1] + segs[-1][2] - segs[0][1]) == (428, 504, 1994, 1510)


def test_buffers_too_small(engine):
    pid, prof, c, sid = two_copies(engine, 75, 80, 2000)
    thr = float(np.exp(10.0))
    want = engine.find_similarities(pid, sid, thr)
    assert len(want) >= 2
    got = engine.find_similarities_batch([pid, pid], [sid, sid], thr, hit_cap=1, seg_cap=1)
    assert_same_hits(got[0], want)
    assert_same_hits(got[1], want)


def test_empty_and_invalid(engine):
    assert engine.find_similarities_batch([], [], 1.0) == []
    pid, _, _, sids = add_case(engine, 76, 30, [2], 0)
    with pytest.raises(D.DummestError):
        engine.find_similarities_batch([pid], [sids[0]], 1.0)  # fewer than 3 letters
