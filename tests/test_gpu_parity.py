"""GPU parity: the HIP kernels against the CPU oracle on the same seeded inputs.

Tolerances: the DP runs in FP32 in a different summation order than the
reference (anti-diagonal tiles, fused multiply-adds), so log-odds scores agree
within 1e-4 nats; the null model and profile transform are bit-exact.
"""
import os
import tempfile

import numpy as np
import pytest

import oracle as O
import dummest_amd as D
from tests.common import synthetic_case, ln_es

pytestmark = pytest.mark.gpu

TOL_NATS = 1e-4


def load_case(engine, seed, L, lengths, homolog_every=2):
    text, prof, dnas = synthetic_case(seed, L, lengths, homolog_every)
    with tempfile.NamedTemporaryFile("w", suffix=".hmm", delete=False) as f:
        f.write(text)
        path = f.name
    try:
        pid = engine.read_hmm(path)[0]
    finally:
        os.unlink(path)
    codes = [D.decode(s) for s in dnas]
    sids = engine.add_sequences(codes)
    return pid, prof, codes, sids


@pytest.mark.parametrize("L", [30, 120, 250])
def test_null_model_bit_exact(engine, L):
    print(f"null model L={L}", flush=True)
    pid, prof, codes, sids = load_case(engine, 11 + L, L, [17, 300, 1001], homolog_every=0)
    for c, s in zip(codes, sids):
        ps, ss, sc = engine.pair_null(pid, s, len(c))
        ref = O.null_model(prof.bg, c)
        np.testing.assert_array_equal(ps, ref["prefix_seed"])
        np.testing.assert_array_equal(ss, ref["suffix_seed"])
        np.testing.assert_array_equal(sc, ref["scales"])


@pytest.mark.parametrize("L,lengths", [(30, [3, 4, 50, 400]), (120, [200, 777, 1500]), (400, [300, 1200])])
def test_filter_scores(engine, L, lengths):
    print(f"filter L={L} lengths={lengths}", flush=True)
    pid, prof, codes, sids = load_case(engine, 100 + L, L, lengths)
    got = engine.find_similarities_backward_only(pid, sids)
    for c, g in zip(codes, got):
        ref = O.filter_score(prof, c)
        if not np.isfinite(ref):
            assert not np.isfinite(g)
            continue
        assert g > 0 and ref > 0
        assert abs(np.log(g) - np.log(ref)) < TOL_NATS, (g, ref)


@pytest.mark.parametrize("L,lengths", [(30, [3, 50, 400]), (120, [200, 777, 1500]), (250, [300, 2000]),
                                       (700, [900])])
def test_fwdbwd_scores(engine, L, lengths):
    print(f"fwdbwd L={L} lengths={lengths}", flush=True)
    pid, prof, codes, sids = load_case(engine, 200 + L, L, lengths)
    mm, me, bm, be = engine.fwdbwd_scores(pid, sids, want_bwd=True)
    for k, c in enumerate(codes):
        r = O.search(prof, c, -2.0, need_align=False, want_bwd=True)
        assert abs(ln_es(mm[k], me[k]) - ln_es(*r.global_best)) < TOL_NATS
        assert abs(ln_es(bm[k], be[k]) - ln_es(*r.bwd_best)) < TOL_NATS


@pytest.mark.parametrize("L,n", [(40, 300), (150, 1500), (300, 700)])
def test_pair_matrices(engine, L, n):
    """Dumped w0/W1 against the oracle's, cell by cell, in log space."""
    print(f"matrices L={L} n={n}", flush=True)
    pid, prof, codes, sids = load_case(engine, 400 + L, L, [n], homolog_every=1)
    c, s = codes[0], sids[0]
    (gw0m, gw0e), (gW1m, gW1e) = engine.pair_matrices(pid, s, len(c))
    (ow0m, ow0e), (oW1m, oW1e), _, _ = O.matrices(prof, c)
    for gm, ge, om, oe in ((gw0m, gw0e, ow0m, ow0e), (gW1m, gW1e, oW1m, oW1e)):
        gl = D.es_log(gm, ge)
        ol = D.es_log(om, oe)
        top = np.max(ol)
        live = ol > top - 60.0  # cells within e^-60 of the largest
        assert live.sum() > 0
        err = np.abs(gl[live] - ol[live])
        assert err.max() < TOL_NATS, (err.max(), np.unravel_index(np.argmax(np.where(live, np.abs(gl - ol), 0)), gl.shape))


def test_dump_matrices_equal_oracle(engine):
    """The matrices of the post-gate step evaluate every cell in the
    reference's order of operations: each live cell of w0 and W1 (within e^-80
    of the largest) equals the oracle's FP32 value exactly, under single-strip
    and row-strip kernels, for a strong hit that rescales many times (the
    medium fixture against the herpes profile, M = 1359) and for synthetic
    profiles of 100-700 positions with embedded homologs."""
    import gzip
    from tests.common import GOLDEN, read_fasta
    table = engine.config_table()
    cases = []
    with tempfile.NamedTemporaryFile("wb", suffix=".hmm", delete=False) as f:
        f.write(gzip.open(os.path.join(GOLDEN, "herpes.hmm.gz")).read())
        path = f.name
    try:
        hp = engine.read_hmm(path)[0]
    finally:
        os.unlink(path)
    M, v, p, b = engine.get_profile(hp)
    hc = D.decode(read_fasta(os.path.join(GOLDEN, "medium.fa"))[0][1][1500:2700])
    cases.append((hp, O.Profile("herpes", M, v, p, b, []), hc, engine.add_sequences([hc])[0]))
    for L, n in ((100, 700), (300, 1100), (700, 800)):
        pid, prof, codes, sids = load_case(engine, 1200 + L, L, [n], homolog_every=1)
        cases.append((pid, prof, codes[0], sids[0]))
    try:
        for pid, prof, c, s in cases:
            (ow0m, ow0e), (oW1m, oW1e), _, _ = O.matrices(prof, c)
            top = max(D.es_log(ow0m, ow0e).max(), D.es_log(oW1m, oW1e).max())
            for idx, (G, R, strips, _) in enumerate(table):
                if (G, R, strips) not in ((16, 8, False), (32, 6, False), (64, 16, False), (32, 6, True),
                                          (64, 4, True), (64, 16, True)):
                    continue
                if not strips and G * R < prof.M + 1:
                    continue
                print(f"dump exact M={prof.M} n={len(c)} config {G},{R},{strips}", flush=True)
                engine.set_option("config", idx)
                (gw0m, gw0e), (gW1m, gW1e) = engine.pair_matrices(pid, s, len(c))
                for gm, ge, om, oe in ((gw0m, gw0e, ow0m, ow0e), (gW1m, gW1e, oW1m, oW1e)):
                    live = D.es_log(om, oe) > top - 80.0
                    assert live.sum() > 100
                    bad = live & ((gm != om) | (ge != oe))
                    assert not bad.any(), (G, R, strips, int(bad.sum()), np.argwhere(bad)[:3])
    finally:
        engine.set_option("config", -1)


def reported(hits):
    """Hits that become records: the reference skips empty alignments at emit
    (dummest-search.hh:36-41)."""
    return sorted((h for h in hits if h[5]), key=lambda h: (h[5][0][1], h[5][0][0]))


def check_hits(got, ref, oracle_cell_ln):
    """Reported hits: identical alignments (segments), scores within the DP
    tolerance. The anchor is the argmax of w0*W1 over a ridge of near-equal
    cells along the alignment, so FP32 summation order can move it along the
    ridge (the reference's own FP32 test leaves the anchor unchecked,
    tests/test_overflow.sh:66-71); the check is that the oracle's value at the
    GPU's anchor cell equals the hit score within the tolerance."""
    g, r = reported(got), reported(ref)
    assert [h[5] for h in g] == [h[5] for h in r]
    for a, b in zip(g, r):
        assert abs(a[0] - b[0]) < TOL_NATS
        assert abs(oracle_cell_ln(a[3], a[4]) - b[0]) < TOL_NATS


def oracle_cells(prof, codes):
    (w0m, w0e), (W1m, W1e), _, _ = O.matrices(prof, codes)
    lw = D.es_log(w0m, w0e) + D.es_log(W1m, W1e)
    return lambda i, j: float(lw[i, j])


@pytest.mark.parametrize("min_ln", [8.0, 20.0])
def test_hits_match_oracle(engine, min_ln):
    """Hits above a realistic threshold (E-value gate) against the oracle."""
    print(f"hits min_ln={min_ln}", flush=True)
    pid, prof, codes, sids = load_case(engine, 300, 150, [1500, 2500, 900], homolog_every=1)
    thr = float(np.exp(min_ln))
    for c, s in zip(codes, sids):
        ref = O.search(prof, c, thr)
        got = engine.find_similarities(pid, s, thr)
        assert len(reported(ref.hits)) >= 1
        check_hits(got, ref.hits, oracle_cells(prof, c))


@pytest.mark.parametrize("L,lengths", [(1100, [400, 1300]), (1500, [700])])
def test_strip_scores(engine, L, lengths):
    """Profiles longer than one lane group (row strips with boundary buffers)."""
    print(f"strips L={L} lengths={lengths}", flush=True)
    pid, prof, codes, sids = load_case(engine, 500 + L, L, lengths)
    mm, me, bm, be = engine.fwdbwd_scores(pid, sids, want_bwd=True)
    best = engine.find_similarities_backward_only(pid, sids)
    for k, c in enumerate(codes):
        r = O.search(prof, c, -2.0, need_align=False, want_bwd=True)
        assert abs(ln_es(mm[k], me[k]) - ln_es(*r.global_best)) < TOL_NATS
        assert abs(ln_es(bm[k], be[k]) - ln_es(*r.bwd_best)) < TOL_NATS
        ref = O.filter_score(prof, c)
        if np.isfinite(ref):
            assert abs(np.log(best[k]) - np.log(ref)) < TOL_NATS
        else:
            assert not np.isfinite(best[k])


def test_herpes_medium_fixture(engine):
    """The reference's overflow fixture (M=1359, strong hit, many rescales):
    scores and the hit list match the oracle; the top hit's spans are the ones
    tests/test_overflow.sh:68 fingerprints."""
    import gzip
    from tests.common import GOLDEN, read_fasta
    print("herpes x medium", flush=True)
    with tempfile.NamedTemporaryFile("wb", suffix=".hmm", delete=False) as f:
        f.write(gzip.open(os.path.join(GOLDEN, "herpes.hmm.gz")).read())
        path = f.name
    try:
        pid = engine.read_hmm(path)[0]
    finally:
        os.unlink(path)
    M, v, p, b = engine.get_profile(pid)
    prof = O.Profile("herpes", M, v, p, b, [])
    dna = read_fasta(os.path.join(GOLDEN, "medium.fa"))[0][1]
    c = D.decode(dna)
    sid = engine.add_sequences([c])[0]
    mm, me, bm, be = engine.fwdbwd_scores([pid], [sid], want_bwd=True)
    r = O.search(prof, c, float(np.exp(20.0)))
    assert abs(ln_es(mm[0], me[0]) - ln_es(*r.global_best)) < TOL_NATS
    assert abs(ln_es(bm[0], be[0]) - ln_es(*r.bwd_best)) < TOL_NATS
    got = engine.find_similarities(pid, sid, float(np.exp(20.0)))
    check_hits(got, r.hits, oracle_cells(prof, c))
    top = max((h for h in got if h[5]), key=lambda h: h[0])
    segs = top[5]
    assert (segs[0][0], segs[-1][0] + segs[-1][2] - segs[0][0], segs[0][1],
            segs[-1][1] + segs[-1][2] - segs[0][1]) == (428, 504, 1994, 1510)


def test_calibration_matches_oracle(engine):
    """dm_calibrate: the trial sequences scored in one batch against the
    oracle on the same sequences; scores within the DP tolerance, fits
    (smooth functions of the scores) within 1e-3 relative."""
    print("calibration", flush=True)
    pa, prof_a, _, _ = load_case(engine, 501, 60, [], 0)
    pb, prof_b, _, _ = load_case(engine, 502, 150, [], 0)
    seq_len, border, trials = 300, 30, 48
    mid, flt, fits = engine.calibrate([pa, pb], seq_len, border, trials)
    for k, p in enumerate([prof_a, prof_b]):
        seqs = [O.decode(O.calib_sequence(p.bg, seq_len, border, t)) for t in range(trials)]
        rm, rb = O.fwdbwd_batch(p, seqs)
        ref_mid = np.array([ln_es(m, e) for m, e in rm])
        ref_flt = np.array([ln_es(m, e) for m, e in rb])
        np.testing.assert_allclose(mid[k], ref_mid, atol=TOL_NATS, rtol=0)
        np.testing.assert_allclose(flt[k], ref_flt, atol=TOL_NATS, rtol=0)
        for key, ref in O.gumbel(ref_mid, seq_len).items():
            assert fits[k]["mid"][key] == pytest.approx(ref, rel=1e-3), key
        for key, ref in O.gumbel(ref_flt, seq_len).items():
            assert fits[k]["filter"][key] == pytest.approx(ref, rel=1e-3), key


def test_calibration_trial_anchors(engine):
    """dm_calibrate_anchors: each trial's best anchor (profile and sequence
    position) and score, the table estimateK prints without sequences
    (dummest-calibration.hh:511-519), against the oracle's third similarity
    of the same trial (dummest-dp-forward.hh:536-543, the end- and
    start-anchored entries repeating its anchor with -inf): anchors identical,
    scores equal (the post-gate matrices are the oracle's bit for bit)."""
    print("calibration anchors", flush=True)
    pa, prof_a, _, _ = load_case(engine, 503, 70, [], 0)
    pb, prof_b, _, _ = load_case(engine, 504, 130, [], 0)
    seq_len, border, trials = 260, 25, 20
    anchors, mid = engine.calibrate_anchors([pa, pb], seq_len, border, trials)
    for k, p in enumerate([prof_a, prof_b]):
        for t in range(trials):
            codes = O.decode(O.calib_sequence(p.bg, seq_len, border, t))
            ref = O.search(p, codes, -2.0, need_align=False)
            assert len(ref.hits) == 3
            assert ref.hits[0][0] == -np.inf and ref.hits[1][0] == -np.inf
            assert ref.hits[0][3:5] == ref.hits[2][3:5]
            ln, m, e, a1, a2, _ = ref.hits[2]
            assert (int(anchors[k, t, 0]), int(anchors[k, t, 1])) == (a1, a2), (k, t)
            assert mid[k, t] == pytest.approx(ln, abs=1e-12), (k, t)


def test_every_configuration(engine):
    """Each compiled configuration forced in turn: single-strip kernels with one
    or two waves per workgroup (two waves share the emission table of two units
    of one profile; an odd unit count leaves a padded item) and the row-strip
    kernels, on pairs whose profile fits, in filter and score modes."""
    table = engine.config_table()
    print(f"every configuration: {len(table)}", flush=True)
    cases = {}
    try:
        for idx, (G, R, strips, _) in enumerate(table):
            rows = G * R
            M = int(rows * 2.3) if strips else rows - 3
            npairs = 3 * (64 // G)  # three units of one profile
            if M not in cases:
                lengths = [150 + (37 * k) % 260 for k in range(npairs)]
                pid, prof, codes, sids = load_case(engine, 900 + M, M, lengths, homolog_every=2)
                rm, rb = O.fwdbwd_batch(prof, codes)
                flt = [O.filter_score(prof, c) for c in codes]
                cases[M] = (pid, sids, rm, rb, flt)
            pid, sids, rm, rb, flt = cases[M]
            sids = sids[:npairs]
            engine.set_option("config", idx)
            mm, me, bm, be = engine.fwdbwd_scores(pid, sids, want_bwd=True)
            best = engine.find_similarities_backward_only(pid, sids)
            for k in range(len(sids)):
                assert abs(ln_es(mm[k], me[k]) - ln_es(*rm[k])) < TOL_NATS, (G, R, strips, k)
                assert abs(ln_es(bm[k], be[k]) - ln_es(*rb[k])) < TOL_NATS, (G, R, strips, k)
                if np.isfinite(flt[k]):
                    assert abs(np.log(best[k]) - np.log(flt[k])) < TOL_NATS, (G, R, strips, k)
                else:
                    assert not np.isfinite(best[k])
    finally:
        engine.set_option("config", -1)
