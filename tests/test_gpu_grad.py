"""Gradient of the summed forward log-likelihood on the GPU (grad.hip,
dm_grad) against the oracle: ln Z of every pair against the oracle's FP64
total of all local paths (tests/test_oracle.py pins that total: forward and
backward sums agree), and d F / d logits against FP64 central finite
differences of the oracle's sum of ln Z over every coordinate of a small
profile and over 64 random coordinates of a 300-position one."""
import numpy as np
import pytest

import dummest_amd as D
import oracle as O
from dummest_amd import grad as G
from dummest_amd import synth

import oracle_flows as OF

pytestmark = pytest.mark.gpu
COMP = str.maketrans("ACGT", "TGCA")


def make_case(rng, M, lens, homolog_every=2):
    el = rng.normal(size=(M, 20)) * 1.2
    tl = rng.normal(size=(M + 1, 7)) * 0.3 + np.array([3.0, 0, 0, 0.8, 0, 0.8, 0])
    raw = np.zeros((M + 1, 29))
    MtoI, MtoD, ItoI, DtoD, _ = G.transitions(tl)
    raw[:, 0], raw[:, 1], raw[:, 2], raw[:, 3] = MtoI, ItoI, MtoD, DtoD
    raw[:M, 4:24] = G._softmax(el)
    bg = O.finalize(O.RawProfile("g", M, raw), dtype=np.float64).bg
    p = G._softmax(el)
    seqs = []
    for k, n in enumerate(lens):
        host = synth.random_dna(rng, n)
        if k % homolog_every == 0:
            prot = np.array([rng.choice(20, p=p[r]) for r in range(M)])
            hom = synth.mutate_dna(rng, synth.reverse_translate(rng, prot), indel=0.01)
            if k % 4 == 2:
                hom = hom.translate(COMP)[::-1]
            host = synth.embed(rng, host, hom[: max(3, n - 10)])
        seqs.append(host)
    return el, tl, bg, seqs


def oracle_lnz(el, tl, bg, seqs, rounded=True):
    v, pa = G.profile_from_logits(el, tl, bg)
    b = bg
    if rounded:  # the engine holds the profile in float
        v, pa, b = (x.astype(np.float32).astype(np.float64) for x in (v, pa, bg))
    prof = O.Profile("g", el.shape[0], np.ascontiguousarray(v), np.ascontiguousarray(pa), np.ascontiguousarray(b), [])
    return np.array([O.totals(prof, O.decode(s))[0] for s in seqs])


@pytest.mark.parametrize("M,lens", [(10, [40, 77, 120, 61]), (63, [200, 150]), (64, [180, 90]), (65, [210, 64]),
                                    (130, [300, 120, 256])])
def test_lnz_matches_oracle(engine, M, lens):
    rng = np.random.default_rng(M)
    el, tl, bg, seqs = make_case(rng, M, lens)
    print("grad lnZ case", M, lens, flush=True)
    L = G.SummedLogLikelihood(engine, seqs, bg)
    F, ge, gt, lnz = L.evaluate(el, tl)
    want = oracle_lnz(el, tl, bg, seqs)
    # the engine's null model hands FP32 seeds and scales to the DP (as the
    # reference's FP32 build does): ln Z within the score tolerance, 1e-4 nats
    print("max |ln Z - oracle| =", np.abs(lnz - want).max())
    assert np.abs(lnz - want).max() < 1e-4, (lnz, want)
    assert F == pytest.approx(lnz.sum(), abs=1e-9)


def test_counts_identities(engine):
    rng = np.random.default_rng(5)
    el, tl, bg, seqs = make_case(rng, 20, [90, 140, 60])
    v, pa = G.profile_from_logits(el, tl, bg)
    pid = engine.add_profile(20, v, pa, bg)
    lnz, c = engine.grad_counts(pid, seqs)
    # every match use emits one codon
    assert np.allclose(c[:, 11:].sum(1), c[:, 10], rtol=1e-12, atol=1e-300)
    # weights scale linearly; an empty list gives zeros
    lnz2, c2 = engine.grad_counts(pid, seqs, weights=[2.0, 0.5, 0.0])
    _, ca = engine.grad_counts(pid, seqs[:1])
    _, cb = engine.grad_counts(pid, seqs[1:2])
    assert np.allclose(c2, 2 * ca + 0.5 * cb, rtol=1e-12, atol=1e-300)
    assert np.array_equal(lnz, lnz2)
    # batches of the matrix budget give the same numbers
    engine.set_option("grad_budget", 6 * 8 * 21 * 100)
    try:
        lnz3, c3 = engine.grad_counts(pid, seqs)
    finally:
        engine.set_option("grad_budget", 8 << 30)
    assert np.array_equal(lnz3, lnz) and np.array_equal(c3, c)


def test_counts_match_oracle_matrices(engine):
    # expected counts from the oracle's own forward and backward matrices
    # (tests/oracle_flows.py; tests/test_grad.py pins those flows against
    # finite differences of the oracle's FP64 ln Z): match entry (also by
    # codon), a1, a2, d1, d2 per row
    rng = np.random.default_rng(21)
    M = 40
    el, tl, bg, seqs = make_case(rng, M, [150, 260, 91])
    v, pa = G.profile_from_logits(el, tl, bg)
    v, pa, b = (np.ascontiguousarray(x.astype(np.float32)) for x in (v, pa, bg))
    pid = engine.add_profile(M, v, pa, b)
    prof = O.Profile("g", M, v, pa, b, [])
    for s in seqs:
        print("counts vs oracle matrices, M", M, "n", len(s), flush=True)
        lnz, got = engine.grad_counts(pid, [s])
        want, by_codon, lnz_o = OF.matrix_flows(prof, s)
        # the oracle's matrices are FP32: counts within 1e-4 of the row maximum
        for k, w in want.items():
            err = np.abs(got[:, OF.PARAM_INDEX[k]] - w).max()
            print(f"  {k}: max |GPU - oracle| = {err:.3g} of max {w.max():.3g}")
            assert err <= 1e-4 * max(w.max(), 1e-6), k
        err = np.abs(got[:, 11:] - by_codon).max()
        print(f"  by codon: max |GPU - oracle| = {err:.3g}; ln Z {lnz[0]:.9f} vs {lnz_o:.9f}")
        assert err <= 1e-4 * max(by_codon.max(), 1e-6)
        assert abs(lnz[0] - lnz_o) < 1e-4


def test_gradient_matches_oracle_finite_differences(engine):
    # every coordinate of a small profile, FP64 central differences (h = 1e-4)
    rng = np.random.default_rng(11)
    M = 6
    el, tl, bg, seqs = make_case(rng, M, [45, 70, 33])
    L = G.SummedLogLikelihood(engine, seqs, bg)
    F, ge, gt, _ = L.evaluate(el, tl)
    h = 1e-4
    scale = max(np.abs(ge).max(), np.abs(gt).max())

    def fd(which, r, k):
        a, b = (el.copy(), el.copy()) if which == 0 else (tl.copy(), tl.copy())
        a[r, k] += h
        b[r, k] -= h
        if which == 0:
            return (oracle_lnz(a, tl, bg, seqs, False).sum() - oracle_lnz(b, tl, bg, seqs, False).sum()) / (2 * h)
        return (oracle_lnz(el, a, bg, seqs, False).sum() - oracle_lnz(el, b, bg, seqs, False).sum()) / (2 * h)

    worst = 0.0
    for r in range(M):
        for k in range(20):
            worst = max(worst, abs(fd(0, r, k) - ge[r, k]))
    for r in range(M + 1):
        for k in range(7):
            worst = max(worst, abs(fd(1, r, k) - gt[r, k]))
    print("max |GPU - FD| / max |grad| =", worst / scale)
    assert worst <= 2e-5 * scale


def test_gradient_300_positions_random_coordinates(engine):
    # the differentiation workload's profile size: M = 300, 1-3 kb sequences
    # (half carrying frameshifted homologs); 64 random coordinates, h = 1e-3,
    # differences of the GPU's own FP64 sum of ln Z (pinned to the oracle above)
    rng = np.random.default_rng(300)
    M = 300
    lens = rng.integers(1000, 3001, 12)
    el, tl, bg, seqs = make_case(rng, M, lens)
    L = G.SummedLogLikelihood(engine, seqs, bg)
    F, ge, gt, lnz = L.evaluate(el, tl)
    want = oracle_lnz(el, tl, bg, seqs[:2])
    # FP32 null-model scales (sC = scale^3 rounded once, applied at every codon
    # step) put ln Z about 4e-8 nats per column from the FP64 oracle's
    print("max |ln Z - oracle| =", np.abs(lnz[:2] - want).max(), "at", lens[:2], "letters")
    assert np.abs(lnz[:2] - want).max() < 2e-7 * max(lens[:2]) + 1e-5
    h = 1e-3
    scale = max(np.abs(ge).max(), np.abs(gt).max())
    worst = 0.0
    for _ in range(64):
        which = int(rng.integers(0, 2))
        r = int(rng.integers(0, M if which == 0 else M + 1))
        k = int(rng.integers(0, 20 if which == 0 else 7))
        a, b = (el.copy(), el.copy()) if which == 0 else (tl.copy(), tl.copy())
        a[r, k] += h
        b[r, k] -= h
        fp = L.evaluate(a, tl)[0] if which == 0 else L.evaluate(el, a)[0]
        fm = L.evaluate(b, tl)[0] if which == 0 else L.evaluate(el, b)[0]
        g = ge[r, k] if which == 0 else gt[r, k]
        worst = max(worst, abs((fp - fm) / (2 * h) - g))
    print("max |GPU - FD| / max |grad| =", worst / scale)
    assert worst <= 1e-4 * scale


def test_autograd_function(engine):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(8)
    el, tl, bg, seqs = make_case(rng, 9, [60, 80])
    L = G.SummedLogLikelihood(engine, seqs, bg)
    F, ge, gt, _ = L.evaluate(el, tl)
    Fn = L.function()
    e = torch.tensor(el, dtype=torch.float64, requires_grad=True)
    t = torch.tensor(tl, dtype=torch.float64, requires_grad=True)
    out = 2.5 * Fn.apply(e, t)
    out.backward()
    assert out.item() == pytest.approx(2.5 * F)
    assert np.allclose(e.grad.numpy(), 2.5 * ge) and np.allclose(t.grad.numpy(), 2.5 * gt)


def strong_case(rng, M, host_len, copies):
    """Peaked emissions and the profile's most likely protein, reverse-translated
    and embedded `copies` times without mutations: ln Z far past FP64's range."""
    el = rng.normal(size=(M, 20)) * 5.0
    tl = rng.normal(size=(M + 1, 7)) * 0.3 + np.array([5.0, 0, 0, 0.8, 0, 0.8, 0])
    raw = np.zeros((M + 1, 29))
    MtoI, MtoD, ItoI, DtoD, _ = G.transitions(tl)
    raw[:, 0], raw[:, 1], raw[:, 2], raw[:, 3] = MtoI, ItoI, MtoD, DtoD
    raw[:M, 4:24] = G._softmax(el)
    bg = O.finalize(O.RawProfile("s", M, raw), dtype=np.float64).bg
    hom = synth.reverse_translate(rng, el.argmax(1))
    host = synth.random_dna(rng, host_len)
    for k in range(copies):
        at = (k + 1) * host_len // (copies + 1)
        host = host[:at] + hom + host[at:]
    return el, tl, bg, host


def test_strong_homologs_beyond_fp64_range(engine):
    # ln Z of about 1056 nats (FP64 ends at e^709): two exact copies of a
    # 400-position profile's most likely protein in one sequence, one copy in
    # another, and a random sequence; the strip/block exponents must carry the
    # rise over the first copy, the fall after it and the second rise
    rng = np.random.default_rng(7)
    M = 400
    el, tl, bg, two = strong_case(rng, M, 600, 2)
    one = strong_case(np.random.default_rng(7), M, 900, 1)[3]
    seqs = [two, one, synth.random_dna(rng, 1500)]
    print("strong homologs: M", M, "lengths", [len(s) for s in seqs], flush=True)
    L = G.SummedLogLikelihood(engine, seqs, bg)
    F, ge, gt, lnz = L.evaluate(el, tl)
    want = oracle_lnz(el, tl, bg, seqs)
    print("ln Z", lnz, "oracle", want)
    assert want[0] > 1000 and np.isfinite(lnz).all()
    assert np.abs(lnz - want).max() < 2e-7 * max(len(s) for s in seqs) + 1e-5
    # the two copies share the posterior: about one match use per row in all
    v, pa = G.profile_from_logits(el, tl, bg)
    v, pa, b = (np.ascontiguousarray(x.astype(np.float32)) for x in (v, pa, bg))
    pid = engine.add_profile(M, v, pa, b)
    prof = O.Profile("s", M, v, pa, b, [])
    for s in seqs:
        _, got = engine.grad_counts(pid, [s])
        want_c, by_codon, _ = OF.matrix_flows(prof, s)
        for k, w in want_c.items():
            err = np.abs(got[:, OF.PARAM_INDEX[k]] - w).max()
            print(f"  n {len(s)} {k}: max |GPU - oracle| = {err:.3g} of max {w.max():.3g}")
            # the oracle's FP32 matrices over a 1,200-letter alignment
            assert err <= 1e-3 * max(w.max(), 1e-6), k
        assert np.abs(got[:, 11:] - by_codon).max() <= 1e-3 * max(by_codon.max(), 1e-6)
    # finite differences of the GPU's own sum of ln Z at 16 coordinates
    h = 1e-3
    scale = max(np.abs(ge).max(), np.abs(gt).max())
    worst = 0.0
    for _ in range(16):
        which = int(rng.integers(0, 2))
        r = int(rng.integers(0, M if which == 0 else M + 1))
        k = int(rng.integers(0, 20 if which == 0 else 7))
        a, b2 = (el.copy(), el.copy()) if which == 0 else (tl.copy(), tl.copy())
        a[r, k] += h
        b2[r, k] -= h
        fp = L.evaluate(a, tl)[0] if which == 0 else L.evaluate(el, a)[0]
        fm = L.evaluate(b2, tl)[0] if which == 0 else L.evaluate(el, b2)[0]
        g = ge[r, k] if which == 0 else gt[r, k]
        worst = max(worst, abs((fp - fm) / (2 * h) - g))
    print("max |GPU - FD| / max |grad| =", worst / scale)
    assert worst <= 1e-4 * scale


def test_long_sequences(engine):
    # sequences far past one LDS window: 20,000 letters with two frameshifted
    # homologs (313 column blocks: the count pass reads its exponents in place)
    # and 9,000 letters with one; ln Z vs the oracle's FP64 total, counts vs the
    # flows through the oracle's matrices
    rng = np.random.default_rng(20)
    M = 40
    el, tl, bg, _ = make_case(rng, M, [10])
    p = G._softmax(el)
    seqs = []
    for n, copies in ((20000, 2), (9000, 1)):
        host = synth.random_dna(rng, n)
        for c in range(copies):
            prot = np.array([rng.choice(20, p=p[r]) for r in range(M)])
            hom = synth.mutate_dna(rng, synth.reverse_translate(rng, prot), indel=0.01)
            at = (c + 1) * n // (copies + 1)
            host = host[:at] + hom + host[at + len(hom):]
        seqs.append(host)
    print("long sequences:", [len(s) for s in seqs], flush=True)
    L = G.SummedLogLikelihood(engine, seqs, bg)
    F, ge, gt, lnz = L.evaluate(el, tl)
    want = oracle_lnz(el, tl, bg, seqs)
    print("ln Z", lnz, "oracle", want)
    assert np.abs(lnz - want).max() < 2e-7 * max(len(s) for s in seqs) + 1e-5
    v, pa = G.profile_from_logits(el, tl, bg)
    v, pa, b = (np.ascontiguousarray(x.astype(np.float32)) for x in (v, pa, bg))
    pid = engine.add_profile(M, v, pa, b)
    prof = O.Profile("l", M, v, pa, b, [])
    for s in seqs:
        _, got = engine.grad_counts(pid, [s])
        assert np.allclose(got[:, 11:].sum(1), got[:, 10], rtol=1e-12, atol=1e-300)
        want_c, by_codon, _ = OF.matrix_flows(prof, s)
        for k, w in want_c.items():
            err = np.abs(got[:, OF.PARAM_INDEX[k]] - w).max()
            print(f"  n {len(s)} {k}: max |GPU - oracle| = {err:.3g} of max {w.max():.3g}")
            assert err <= 1e-4 * max(w.max(), 1e-6), k
        assert np.abs(got[:, 11:] - by_codon).max() <= 1e-4 * max(by_codon.max(), 1e-6)


def test_longest_sequence_limit(engine):
    v, pa = G.profile_from_logits(np.zeros((5, 20)), np.zeros((6, 7)), np.full(29, 0.05))
    pid = engine.add_profile(5, v, pa, np.full(29, 0.05))
    with pytest.raises(D.DummestError, match="letters"):
        engine.grad_counts(pid, ["A" * ((1 << 28) + 1)])


def _grad_with_budget(engine, pid, seqs, budget, weights=None):
    engine.set_option("grad_budget", budget)
    try:
        lnz, c = engine.grad_counts(pid, seqs, weights=weights)
        return lnz, c, engine.last_grad_segments
    finally:
        engine.set_option("grad_budget", 8 << 30)


def test_checkpointed_segments_bit_identical(engine):
    """Sequences whose matrices pass the budget are checkpointed (the edge row
    of every 64-row strip kept, each segment of strips computed again): ln Z
    and the counts equal the full-matrix path's bit for bit, at segments of
    one, two and three strips and with a sequence held in full beside them;
    the full path's are checked against the oracle's flows in
    test_counts_match_oracle_matrices and here for the longest sequence."""
    rng = np.random.default_rng(71)
    M = 200  # 4 strips
    lens = [3000, 1500, 700]
    el, tl, bg, seqs = make_case(rng, M, lens, homolog_every=1)
    v, pa = G.profile_from_logits(el, tl, bg)
    v, pa, b = (np.ascontiguousarray(x.astype(np.float32)) for x in (v, pa, bg))
    pid = engine.add_profile(M, v, pa, b)
    wts = [1.0, 0.25, 3.0]
    print("checkpointed gradient: full path", flush=True)
    lnz, c, seg = _grad_with_budget(engine, pid, seqs, 8 << 30, wts)
    assert seg == 0
    strip = lambda n: 48 * (n + 63) * 64  # bytes of one strip's six matrices
    # budget of two strips of the longest: it runs in 2 segments of 2 strips,
    # the 1500-letter one in segments of 3 and 1, the shortest in full
    cases = [(2 * strip(3000), 4), (1, 12), (3 * strip(3000), 2)]
    for budget, want_seg in cases:
        print("checkpointed gradient: budget", budget, flush=True)
        lnz2, c2, seg = _grad_with_budget(engine, pid, seqs, budget, wts)
        assert seg == want_seg, (budget, seg)
        np.testing.assert_array_equal(lnz2, lnz)
        np.testing.assert_array_equal(c2, c)
    prof = O.Profile("k", M, v, pa, b, [])
    _, got, seg = _grad_with_budget(engine, pid, seqs[:1], 1)
    assert seg == 4
    want, by_codon, lnz_o = OF.matrix_flows(prof, seqs[0])
    for k, w in want.items():
        assert np.abs(got[:, OF.PARAM_INDEX[k]] - w).max() <= 1e-4 * max(w.max(), 1e-6), k
    assert np.abs(got[:, 11:] - by_codon).max() <= 1e-4 * max(by_codon.max(), 1e-6)


def test_sequence_past_a_million_letters(engine):
    """A 1.1-million-letter contig (past 2^20 letters) with three embedded
    homologs: held in full (6.8 GB of matrices) and checkpointed one strip at
    a time, the same numbers bit for bit; ln Z against the oracle's FP64
    total."""
    rng = np.random.default_rng(72)
    M = 70  # 2 strips
    n = 1_100_000
    el, tl, bg, _ = make_case(rng, M, [10])
    p = G._softmax(el)
    host = synth.random_dna(rng, n)
    for c in range(3):
        prot = np.array([rng.choice(20, p=p[r]) for r in range(M)])
        hom = synth.mutate_dna(rng, synth.reverse_translate(rng, prot), indel=0.01)
        at = (c + 1) * n // 4
        host = host[:at] + hom + host[at + len(hom):]
    assert len(host) == n
    v, pa = G.profile_from_logits(el, tl, bg)
    v, pa, b = (np.ascontiguousarray(x.astype(np.float32)) for x in (v, pa, bg))
    pid = engine.add_profile(M, v, pa, b)
    print("contig of", n, "letters: full matrices", flush=True)
    lnz, c, seg = _grad_with_budget(engine, pid, [host], 8 << 30)
    assert seg == 0
    print("contig of", n, "letters: checkpointed", flush=True)
    lnz2, c2, seg = _grad_with_budget(engine, pid, [host], 4 << 30)
    assert seg == 2
    np.testing.assert_array_equal(lnz2, lnz)
    np.testing.assert_array_equal(c2, c)
    assert np.allclose(c[:, 11:].sum(1), c[:, 10], rtol=1e-12, atol=1e-300)
    want = oracle_lnz(el, tl, bg, [host])
    print("ln Z", lnz, "oracle", want, flush=True)
    assert abs(lnz[0] - want[0]) < 2e-7 * n + 1e-5
    assert lnz[0] > 20.0  # the homologs are found


def test_unusual_letters_and_short_sequences(engine):
    # lowercase letters are bases; N, R, Y make the codon '?' (code 22) and its
    # match flow goes to the non-ACGT codon bin; sequences of 3-5 letters; an
    # empty list gives zero counts; fewer than 3 letters is refused
    rng = np.random.default_rng(33)
    M = 12
    el, tl, bg, seqs = make_case(rng, M, [90, 70])
    s0 = seqs[0]
    s0 = s0[:10].lower() + "N" + s0[11:40] + "RY" + s0[42:60].lower() + "NNN" + s0[63:]
    seqs = [s0, seqs[1], "ACG", "TGGC", "atgaa"]
    v, pa = G.profile_from_logits(el, tl, bg)
    v, pa, b = (np.ascontiguousarray(x.astype(np.float32)) for x in (v, pa, bg))
    pid = engine.add_profile(M, v, pa, b)
    prof = O.Profile("u", M, v, pa, b, [])
    for s in seqs:
        print("letters case, n", len(s), flush=True)
        lnz, got = engine.grad_counts(pid, [s])
        want, by_codon, lnz_o = OF.matrix_flows(prof, s)
        assert abs(lnz[0] - lnz_o) < 1e-4
        for k, w in want.items():
            assert np.abs(got[:, OF.PARAM_INDEX[k]] - w).max() <= 1e-4 * max(w.max(), 1e-6), (s, k)
        assert np.abs(got[:, 11:] - by_codon).max() <= 1e-4 * max(by_codon.max(), 1e-6)
        assert np.allclose(got[:, 11:].sum(1), got[:, 10], rtol=1e-12, atol=1e-300)
    # the non-ACGT codon bin is used by the first sequence
    _, got = engine.grad_counts(pid, [s0])
    assert got[:, 11 + 64].sum() > 0
    lnz, got = engine.grad_counts(pid, [])
    assert len(lnz) == 0 and not got.any()
    with pytest.raises(D.DummestError, match="at least 3"):
        engine.grad_counts(pid, ["AC"])


def test_registered_set_is_bit_identical(engine):
    # dm_grad_set over registered sequences gives dm_grad's numbers bit for
    # bit, with weights, across profile updates, interleaved with dm_grad calls
    rng = np.random.default_rng(44)
    M = 30
    el, tl, bg, seqs = make_case(rng, M, [120, 333, 64, 250])
    v, pa = G.profile_from_logits(el, tl, bg)
    pid = engine.add_profile(M, v, pa, bg)
    sid = engine.grad_set(seqs)
    lnz_a, c_a = engine.grad_counts(pid, seqs)
    lnz_b, c_b = engine.grad_set_counts(pid, sid)
    assert np.array_equal(lnz_a, lnz_b) and np.array_equal(c_a, c_b)
    w = [1.0, 0.25, 3.0, 0.0]
    assert np.array_equal(engine.grad_counts(pid, seqs, w)[1], engine.grad_set_counts(pid, sid, w)[1])
    v2, pa2 = G.profile_from_logits(el * 0.5, tl, bg)
    engine.set_profile(pid, M, v2, pa2, bg)  # same length: rewritten in place on the device
    lnz_c, c_c = engine.grad_set_counts(pid, sid)
    lnz_d, c_d = engine.grad_counts(pid, seqs)
    assert np.array_equal(lnz_c, lnz_d) and np.array_equal(c_c, c_d) and not np.array_equal(c_c, c_a)
    pid2 = engine.add_profile(M, v2, pa2, bg)  # the same profile through a full upload
    lnz_e, c_e = engine.grad_set_counts(pid2, sid)
    assert np.array_equal(lnz_e, lnz_c) and np.array_equal(c_e, c_c)
    empty = engine.grad_set([])
    lz, ce = engine.grad_set_counts(pid, empty)
    assert len(lz) == 0 and not ce.any()
    engine.grad_set_destroy(sid)
    with pytest.raises(D.DummestError, match="unknown sequence set"):
        engine.grad_set_counts(pid, sid)
