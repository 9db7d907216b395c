"""The lane-batched SIMD CPU baseline (oracle/simd_baseline.cpp, the bench's
CPU comparator) against the oracle's scalar restatement of the same
forward+backward score path: per pair the best mid score and the best
backward start, within 1e-4 nats (the baseline lets the compiler fuse
multiply-adds, as the reference's own build does; the oracle does not).
Uneven lengths inside a lane batch, batch sizes that are not multiples of the
lane count, several profiles in one call, strong homologs whose rows rescale,
sequences of 3 and 4 letters, and both instruction sets the host runs."""
import numpy as np
import pytest

import oracle as O
from tests.common import ln_es, synthetic_case

TOL_NATS = 1e-4


def variants():
    out = ["v3"]
    if O.simd_variant() == "v4":
        out.append("v4")
    return out


def build_case():
    profs, seqs, pp, ps = [], [], [], []
    for seed, M, lens in ((3, 20, [3, 4, 60, 151, 300, 222, 90]),
                          (4, 150, [150 + 7 * k for k in range(21)]),
                          (5, 400, [1200, 300, 800, 40, 999])):
        _, prof, dnas = synthetic_case(seed, M, lens, homolog_every=2)
        p = len(profs)
        profs.append(prof)
        for d in dnas:
            ps.append(len(seqs))
            pp.append(p)
            seqs.append(O.decode(d))
    # every profile also against the first sequences of the others
    for p in range(len(profs)):
        for s in range(0, len(seqs), 5):
            pp.append(p)
            ps.append(s)
    return profs, seqs, np.array(pp), np.array(ps)


@pytest.fixture(scope="module")
def case():
    profs, seqs, pp, ps = build_case()
    want = {}
    for p in range(len(profs)):
        idx = np.nonzero(pp == p)[0]
        mid, bwd = O.fwdbwd_batch(profs[p], [seqs[s] for s in ps[idx]], threads=4)
        for k, q in enumerate(idx):
            want[q] = (ln_es(*mid[k]), ln_es(*bwd[k]))
    return profs, seqs, pp, ps, want


@pytest.mark.parametrize("variant", variants())
@pytest.mark.parametrize("threads", [1, 3])
def test_scores_match_oracle(case, variant, threads):
    profs, seqs, pp, ps, want = case
    mid, bwd, sec, lanes = O.simd_fwdbwd_pairs(profs, seqs, pp, ps, threads=threads, variant=variant)
    assert lanes == (16 if variant == "v4" else 8)
    assert sec > 0
    worst = 0.0
    for q in range(len(pp)):
        gm, gb = ln_es(*mid[q]), ln_es(*bwd[q])
        wm, wb = want[q]
        worst = max(worst, abs(gm - wm), abs(gb - wb))
        assert abs(gm - wm) < TOL_NATS, (q, gm, wm)
        assert abs(gb - wb) < TOL_NATS, (q, gb, wb)
    print(variant, "threads", threads, "max |d ln| =", worst)
    # the strong homologs pass e^70 (2^101 > 1e30): their rows were rescaled
    assert max(w[0] for w in want.values()) > 70.0


def test_batches_are_independent(case):
    """A pair's scores do not depend on which lane batch it lands in."""
    profs, seqs, pp, ps, want = case
    v = variants()[-1]
    mid, bwd, _, _ = O.simd_fwdbwd_pairs(profs, seqs, pp, ps, threads=2, variant=v)
    sub = np.arange(0, len(pp), 3)
    mid2, bwd2, _, _ = O.simd_fwdbwd_pairs(profs, seqs, pp[sub], ps[sub], threads=2, variant=v)
    np.testing.assert_array_equal(mid2, mid[sub])
    np.testing.assert_array_equal(bwd2, bwd[sub])
    # without the backward start: the same mid scores
    mid3, bwd3, _, _ = O.simd_fwdbwd_pairs(profs, seqs, pp, ps, threads=2, variant=v, want_bwd=False)
    np.testing.assert_array_equal(mid3, mid)
    assert not bwd3.any()
