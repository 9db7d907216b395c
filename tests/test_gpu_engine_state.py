"""Engine state across calls: plans after the engine's profiles and sequences
change, gradient sets after the store is cleared, hit buffers that are too
small, and a HIP error (an allocation the device cannot satisfy) after which
the engine still scores correctly."""
import ctypes

import numpy as np
import pytest

import dummest_amd as D
import oracle as O
from tests.common import synthetic_case

pytestmark = pytest.mark.gpu

TOL_NATS = 1e-4


def ln(m, e):
    return O.es_log(float(m), int(e))


def case(e, seed, L, lengths):
    text, prof, dnas = synthetic_case(seed, L, lengths, homolog_every=2)
    M, v, p, b, _ = D.parse_hmm_text(text)
    pid = e.add_profile(M, v, p, b)
    codes = [D.decode(s) for s in dnas]
    sids = e.add_sequences(codes)
    return pid, prof, codes, sids, (M, v, p, b)


def assert_same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_plan_follows_engine_changes():
    print("plans across engine changes", flush=True)
    e = D.Engine(0)
    try:
        pid, prof, codes, sids, (M, v, p, b) = case(e, 31, 90, [400, 700, 250, 900])
        pids = [pid] * len(sids)
        plan = e.plan(pids, sids, "score", want_bwd=True)
        cap = e.plan(pids, sids, "score", want_bwd=True)
        cap.capture()
        first = plan.execute()
        assert_same(cap.execute(), first)
        # profiles added after the plan (the tables are uploaded again and may move)
        for k in range(40):
            case(e, 100 + k, 60 + 7 * k, [300])
        assert_same(plan.execute(), first)
        assert_same(cap.execute(), first)
        # the plan's profile rewritten at the same length: the plans read the new tables
        M2, v2f, p2f, b2f, _ = D.parse_hmm_text(synthetic_case(32, 90, [], 0)[0])
        assert M2 == M
        e.set_profile(pid, M2, v2f, p2f, b2f)
        fresh = e.plan(pids, sids, "score", want_bwd=True)
        want = fresh.execute()
        fresh.close()
        assert not np.array_equal(want[0], first[0])
        assert_same(plan.execute(), want)
        assert_same(cap.execute(), want)
        op = O.Profile("x", M2, *e.get_profile(pid)[1:], [])
        rm, rb = O.fwdbwd_batch(op, codes)
        for k in range(len(codes)):
            assert abs(ln(want[0][k], want[1][k]) - ln(*rm[k])) < TOL_NATS
        # another length: the plans no longer describe their pairs
        Mx, vx, px, bx, _ = D.parse_hmm_text(synthetic_case(77, 130, [], 0)[0])
        e.set_profile(pid, Mx, vx, px, bx)
        for pl in (plan, cap):
            with pytest.raises(D.DummestError) as ex:
                pl.execute()
            assert ex.value.code == D.DM_EINVAL
        e.set_profile(pid, M2, v2f, p2f, b2f)
        assert_same(plan.execute(), want)
        # sequences cleared: the sequence ids of the plans are gone
        e.clear_sequences()
        e.add_sequences(codes)
        for pl in (plan, cap):
            with pytest.raises(D.DummestError) as ex:
                pl.execute()
            assert ex.value.code == D.DM_EINVAL
        plan.close()
        cap.close()
    finally:
        e.close()


def test_grad_set_invalid_after_clear():
    print("gradient set after clear_sequences", flush=True)
    e = D.Engine(0)
    try:
        text, prof, dnas = synthetic_case(41, 40, [300, 500], homolog_every=1)
        M, v, p, b, _ = D.parse_hmm_text(text)
        pid = e.add_profile(M, v, p, b)
        sid = e.grad_set(dnas)
        lnz, _ = e.grad_set_counts(pid, sid)
        assert np.all(np.isfinite(lnz))
        e.clear_sequences()
        with pytest.raises(D.DummestError) as ex:
            e.grad_set_counts(pid, sid)
        assert ex.value.code == D.DM_EINVAL
        # a set registered afterwards works
        sid2 = e.grad_set(dnas)
        lnz2, _ = e.grad_set_counts(pid, sid2)
        np.testing.assert_array_equal(lnz, lnz2)
    finally:
        e.close()


def test_hit_buffers_too_small(engine):
    """Two embedded copies of a homolog give two hits: with room for one the
    call reports DM_ERANGE and the sizes it needs, and the repeated call with
    those sizes gives the oracle's hits."""
    print("hit buffers too small", flush=True)
    rng = np.random.default_rng(9)
    text, prof, dnas = synthetic_case(52, 80, [2000], homolog_every=1)
    from dummest_amd import synth
    raw = O.parse_hmm(text)[0]
    ins = synth.reverse_translate(rng, np.array(raw.consensus, np.int64) % 20)
    dna = synth.random_dna(rng, 300) + ins + synth.random_dna(rng, 400) + ins + synth.random_dna(rng, 300)
    M, v, p, b, _ = D.parse_hmm_text(text)
    pid = engine.add_profile(M, v, p, b)
    codes = D.decode(dna)
    sid = engine.add_sequences([codes])[0]
    thr = float(np.exp(10.0))
    hits = np.zeros((1, 6), np.float64)
    segs = np.zeros((1, 3), np.int32)
    nh, ns = ctypes.c_int32(), ctypes.c_int32()
    rc = D.lib().dm_pair_hits(engine._h, pid, sid, thr, D._ptr(hits, ctypes.c_double), 1,
                              D._ptr(segs, ctypes.c_int32), 1, ctypes.byref(nh), ctypes.byref(ns))
    assert rc == D.DM_ERANGE
    assert nh.value >= 2 and ns.value >= 2
    got = engine.find_similarities(pid, sid, thr, hit_cap=1, seg_cap=1)
    assert len(got) == nh.value and sum(len(h[5]) for h in got) == ns.value
    ref = O.search(O.Profile("x", M, v, p, b, []), codes, thr)
    assert [h[5] for h in got if h[5]] == [h[5] for h in ref.hits if h[5]]


def test_hip_error_leaves_engine_usable(engine):
    """An allocation larger than the device (64 TiB) fails with DM_EHIP; the
    next scoring call on the same engine matches the oracle."""
    print("HIP error path", flush=True)
    with pytest.raises(D.DummestError) as ex:
        engine.copy_probe(1 << 46)
    assert ex.value.code == D.DM_EHIP
    assert "hipMalloc" in str(ex.value) or "memory" in str(ex.value).lower()
    text, prof, dnas = synthetic_case(61, 70, [500, 800], homolog_every=2)
    M, v, p, b, _ = D.parse_hmm_text(text)
    pid = engine.add_profile(M, v, p, b)
    codes = [D.decode(s) for s in dnas]
    sids = engine.add_sequences(codes)
    mm, me, bm, be = engine.fwdbwd_scores([pid] * 2, sids, want_bwd=True)
    rm, rb = O.fwdbwd_batch(O.Profile("x", M, v, p, b, []), codes)
    for k in range(2):
        assert abs(ln(mm[k], me[k]) - ln(*rm[k])) < TOL_NATS
        assert abs(ln(bm[k], be[k]) - ln(*rb[k])) < TOL_NATS


def test_score_plan_rejects_overlong_sequence(engine):
    """The score pass reads a row strip of a wave's W1 blocks through one
    buffer resource (below 2^31 bytes): a pair whose sequence exceeds
    2^23 / R - G + 1 letters for every configuration (9M letters) is refused
    with DM_EUNSUP when its plan is built, and the engine still scores the
    next pairs as the oracle does."""
    print("overlong score pass", flush=True)
    text, prof, dnas = synthetic_case(67, 60, [400, 700], homolog_every=2)
    M, v, p, b, _ = D.parse_hmm_text(text)
    pid = engine.add_profile(M, v, p, b)
    rng = np.random.default_rng(67)
    longSeq = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, 9_000_000)].tobytes().decode()
    lid = engine.add_sequences([D.decode(longSeq)])[0]
    with pytest.raises(D.DummestError) as ex:
        engine.plan([pid], [lid], mode="score")
    assert ex.value.code == D.DM_EUNSUP
    assert "too long" in str(ex.value)
    codes = [D.decode(s) for s in dnas]
    sids = engine.add_sequences(codes)
    mm, me, bm, be = engine.fwdbwd_scores([pid] * 2, sids, want_bwd=True)
    rm, rb = O.fwdbwd_batch(O.Profile("x", M, v, p, b, []), codes)
    for k in range(2):
        assert abs(ln(mm[k], me[k]) - ln(*rm[k])) < TOL_NATS
        assert abs(ln(bm[k], be[k]) - ln(*rb[k])) < TOL_NATS
