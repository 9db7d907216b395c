"""tantan masking (-m 1 profiles, -m 2 sequences) without a GPU: the oracle's
restatement pinned against the reference's own tantan (oracle/_ref, built from
tantan-wrapper.hh and tantan/tantan.cc) and against a dense-matrix HMM
forward-backward; profile masking of the product's host transform against the
oracle, bit for bit."""
import numpy as np
import pytest

import oracle as O
import dummest_amd as D
from dummest_amd import synth
from tests.common import golden_hmm


def tantan_cases(seed, count):
    """Random, tandem-repeat, mutated-repeat and unusual-letter sequences."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(count):
        prot = t % 2 == 1
        A = 20 if prot else 4
        n = int(rng.integers(0, 700)) if t % 3 else int(rng.integers(0, 40))
        kind = t % 5
        seq = rng.integers(0, A, n)
        if kind and n:
            per = int(rng.integers(1, 60 if prot else 120))
            seq = np.resize(rng.integers(0, A, per), n)
            mut = rng.random(n) < (0.0, 0.02, 0.1, 0.3, 0.5)[kind]
            seq[mut] = rng.integers(0, A, int(mut.sum()))
            k = int(rng.integers(0, n // 2 + 1))
            seq[:k] = rng.integers(0, A, k)
        if prot and t % 7 == 0 and n:
            seq[rng.random(n) < 0.05] = 20 + rng.integers(0, 2)
        out.append((np.ascontiguousarray(seq.astype(np.uint8)), prot))
    return out


def test_oracle_tantan_matches_reference_build():
    """Bit-exact against the reference's tantan compiled with its release flags."""
    if O.ref_tantan(np.zeros(1, np.uint8), False) is None:
        pytest.skip("oracle/_ref not built (needs the reference tree; `make -C oracle ref`)")
    masked = 0
    for seq, prot in tantan_cases(11, 160):
        got = O.tantan(seq, prot)
        ref = O.ref_tantan(seq, prot)
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
        masked += int(np.sum(ref >= 0.5))
    assert masked > 1000  # the cases exercise masking


def dense_repeat_probs(seq, prot):
    """The tantan HMM as dense matrices (background + K offset states, begin in
    the background, transition then emission per letter), unscaled FP64."""
    K = 50 if prot else 100
    rp, re_, decay = 0.005, 0.05, 0.9
    first = (1 - decay) / (1 - decay ** K)
    b2f = rp * first * decay ** np.arange(K)
    T = np.zeros((K + 1, K + 1))
    T[0, 0] = 1 - rp
    T[0, 1:] = b2f
    T[1:, 0] = re_
    T[np.arange(1, K + 1), np.arange(1, K + 1)] = 1 - re_
    if prot:
        B = O.blosum62().astype(float)
        sc = np.full((22, 22), -4.0)
        sc[:20, :20] = B
        ratio = np.exp(0.32403216734804408 * sc)
    else:
        ratio = np.full((22, 22), 1 / 3)
        ratio[np.arange(4), np.arange(4)] = 3.0
    n = len(seq)
    E = np.zeros((n, K + 1))
    E[:, 0] = 1.0
    for p in range(n):
        for i in range(min(p, K)):
            E[p, i + 1] = ratio[seq[p], seq[p - i - 1]]
    alpha = np.zeros((n, K + 1))
    prev = np.zeros(K + 1)
    prev[0] = 1.0
    for p in range(n):
        prev = (prev @ T) * E[p]
        alpha[p] = prev
    end = np.full(K + 1, re_)
    end[0] = 1 - rp
    z = alpha[-1] @ end
    beta = np.zeros((n, K + 1))
    beta[-1] = end
    for p in range(n - 2, -1, -1):
        beta[p] = T @ (E[p + 1] * beta[p + 1])
    return 1 - alpha[:, 0] * beta[:, 0] / z


@pytest.mark.parametrize("prot", [False, True])
def test_oracle_tantan_against_dense_hmm(prot):
    rng = np.random.default_rng(5 + prot)
    A = 20 if prot else 4
    for n in (1, 2, 5, 17, 60, 130, 220):
        for rep in (0, 1):
            seq = rng.integers(0, A, n)
            if rep:
                seq = np.resize(rng.integers(0, A, int(rng.integers(1, 9))), n)
            seq = seq.astype(np.uint8)
            np.testing.assert_allclose(O.tantan(seq, prot), dense_repeat_probs(seq, prot), atol=2e-6)


def test_mask_dna_rules():
    """Masked letters become N; non-ACGT letters count as A for tantan."""
    rng = np.random.default_rng(3)
    dna = "".join("ACGT"[x] for x in rng.integers(0, 4, 300))
    rep = "ACGTTGCA" * 30
    s = dna[:150] + rep + dna[150:]
    masked, probs = O.mask_dna(s, 0.5)
    assert all((m == "N") == (p >= 0.5) for m, p in zip(masked, probs))
    assert masked[160:380].count("N") > 150 and masked[:140].count("N") < 20
    same, p2 = O.mask_dna(s.replace("A", "n"), 0.5)  # non-ACGT -> A for tantan
    np.testing.assert_array_equal(probs, p2)
    low, p3 = O.mask_dna(s.lower(), 0.5)
    np.testing.assert_array_equal(probs, p3)


def repeat_profile_text(seed=21, L=240):
    rng = np.random.default_rng(seed)
    unit = rng.integers(0, 20, 7)
    anc = rng.integers(0, 20, L)
    anc[60:180] = np.resize(unit, 120)
    text, _, _ = synth.make_profile(rng, L, f"rep{seed}", anc=anc)
    return text


@pytest.mark.parametrize("source", ["repeat", "herpes.hmm.gz", "MET-test.hmm.gz"])
def test_profile_masking_matches_oracle(source):
    text = repeat_profile_text() if source == "repeat" else golden_hmm(source)
    total = 0
    for idx, rp in enumerate(O.parse_hmm(text)):
        ref = O.finalize(rp, np.float32, mask_threshold=0.5)
        M, v, p, b, cnt, nm = D.parse_hmm_text(text, idx, mask_threshold=0.5, with_masked=True)
        np.testing.assert_array_equal(v, ref.values)
        np.testing.assert_array_equal(p, ref.params)
        np.testing.assert_array_equal(b, ref.bg)
        assert nm == ref.masked
        total += nm
        if idx >= 3:
            break
    if source == "repeat":
        assert total >= 60  # the tandem block is masked
        unmasked = O.finalize(O.parse_hmm(text)[0], np.float32)
        assert not np.array_equal(unmasked.bg, ref.bg)  # and it changes the background


def test_masking_threshold_edges():
    text = repeat_profile_text(22)
    rp = O.parse_hmm(text)[0]
    for thr in (0.0, 0.3, 1.0):
        ref = O.finalize(rp, np.float32, mask_threshold=thr) if thr > 0 else None
        if thr == 0.0:
            # every position masked: the reference asserts (no unmasked mean)
            with pytest.raises(D.DummestError):
                D.parse_hmm_text(text, 0, mask_threshold=thr)
            continue
        M, v, p, b, cnt, nm = D.parse_hmm_text(text, 0, mask_threshold=thr, with_masked=True)
        np.testing.assert_array_equal(b, ref.bg)
        assert nm == ref.masked
