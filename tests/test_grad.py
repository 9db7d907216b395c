"""The summed-likelihood gradient's host side (dummest_amd/grad.py): the
profile built from logits is finalizeProfile's (dummest-profile.hh:166-224,
through the oracle's restatement), and the chain rule from expected counts
(d ln Z / d ln theta) to the logits matches finite differences of a function
that is linear in the log-parameters with those counts as coefficients (the
first-order form of ln Z). The GPU's counts are checked in test_gpu_grad.py."""
import numpy as np
import pytest

import oracle as O
from dummest_amd import grad as G


def _raw(el, tl):
    M = el.shape[0]
    MtoI, MtoD, ItoI, DtoD, _ = G.transitions(tl)
    raw = np.zeros((M + 1, 29))
    raw[:, 0], raw[:, 1], raw[:, 2], raw[:, 3] = MtoI, ItoI, MtoD, DtoD
    raw[:M, 4:24] = G._softmax(el)
    return raw


@pytest.mark.parametrize("M", [1, 7, 40])
def test_profile_from_logits_is_finalize(M):
    rng = np.random.default_rng(M)
    el, tl = rng.normal(size=(M, 20)), rng.normal(size=(M + 1, 7))
    p = O.finalize(O.RawProfile("t", M, _raw(el, tl)), dtype=np.float64)
    v, pa = G.profile_from_logits(el, tl, p.bg)
    assert np.array_equal(pa, p.params)
    assert np.array_equal(v[:, 4:], p.values[:, 4:])


def _surrogate(c, el, tl, bg):
    # sum over rows of c * ln(param) and c * ln(codon emission)
    v, pa = G.profile_from_logits(el, tl, bg)
    M = el.shape[0]
    f = 0.0
    for k in (0, 3, 6, 9, 10):
        m = pa[:M + 1, k] > 0
        f += (c[:M + 1, k][m] * np.log(pa[:M + 1, k][m])).sum()
    for cod in range(64):
        aa = G.CODON_AA[cod]
        if aa >= 0:
            f += (c[:M, 11 + cod] * np.log(v[:M, 4 + aa])).sum()
    return f


def test_chain_rule_matches_finite_differences():
    rng = np.random.default_rng(3)
    M = 6
    el, tl = rng.normal(size=(M, 20)), rng.normal(size=(M + 1, 7))
    bg = O.finalize(O.RawProfile("t", M, _raw(el, tl)), dtype=np.float64).bg
    c = np.abs(rng.normal(size=(M + 1, 76)))
    c[M, 6:11] = 0.0  # the last row has no delete, delete-extend or match uses
    c[M, 11:] = 0.0
    ge, gt, _ = G.logit_grads(c, el, tl)
    h = 1e-6
    for r in range(M):
        for k in range(20):
            a, b = el.copy(), el.copy()
            a[r, k] += h
            b[r, k] -= h
            fd = (_surrogate(c, a, tl, bg) - _surrogate(c, b, tl, bg)) / (2 * h)
            assert abs(fd - ge[r, k]) < 1e-6 * max(1, abs(fd)), (r, k, fd, ge[r, k])
    for r in range(M + 1):
        for k in range(7):
            a, b = tl.copy(), tl.copy()
            a[r, k] += h
            b[r, k] -= h
            fd = (_surrogate(c, el, a, bg) - _surrogate(c, el, b, bg)) / (2 * h)
            assert abs(fd - gt[r, k]) < 1e-6 * max(1, abs(fd)), (r, k, fd, gt[r, k])


def test_oracle_matrix_flows_are_log_derivatives():
    # The expected counts the GPU's are checked against (tests/oracle_flows.py,
    # flows through the oracle's FP32 forward and backward matrices) equal
    # d ln Z / d ln theta by FP64 central differences of the oracle's total of
    # all local paths: per row for enter, a1, a2, d1, d2, and per amino acid
    # for the emission (the sum of its codons' match counts). A flow with a
    # wrong row, column offset or scale (s1 for s2, say) fails here.
    import oracle_flows as OF

    rng = np.random.default_rng(4)
    M = 8
    el, tl = rng.normal(size=(M, 20)) * 1.5, rng.normal(size=(M + 1, 7)) * 0.3 + [2.0, 0, 0, 0.5, 0, 0.5, 0]
    bg = O.finalize(O.RawProfile("t", M, _raw(el, tl)), dtype=np.float64).bg
    v, pa = G.profile_from_logits(el, tl, bg)
    v, pa, b = (np.ascontiguousarray(x.astype(np.float32)) for x in (v, pa, bg))
    prof32 = O.Profile("t", M, v, pa, b, [])
    p = G._softmax(el)
    prot = np.array([rng.choice(20, p=p[r]) for r in range(M)])
    from dummest_amd import synth
    dna = synth.random_dna(rng, 20) + synth.reverse_translate(rng, prot) + synth.random_dna(rng, 17)
    dna = dna[:40] + "A" + dna[40:]  # one frameshift
    counts, by_codon, lnz = OF.matrix_flows(prof32, dna)
    codes = O.decode(dna)
    h = 1e-5

    def lnz64(vv, pp):
        return O.totals(O.Profile("t", M, vv, pp, b.astype(np.float64), []), codes)[0]

    v64, p64 = v.astype(np.float64), pa.astype(np.float64)
    assert abs(lnz64(v64, p64) - lnz) < 1e-5
    worst = 0.0
    for k, col in OF.PARAM_INDEX.items():
        for i in range(M + 1):
            if p64[i, col] == 0:
                assert counts[k][i] == 0
                continue
            a, c = p64.copy(), p64.copy()
            a[i, col] *= np.exp(h)
            c[i, col] *= np.exp(-h)
            fd = (lnz64(v64, a) - lnz64(v64, c)) / (2 * h)
            worst = max(worst, abs(fd - counts[k][i]))
    n_aa = np.zeros((M + 1, 20))
    for aa in range(20):
        n_aa[:, aa] = by_codon[:, :64][:, G.CODON_AA == aa].sum(1)
    for i in range(M):
        for aa in range(20):
            a, c = v64.copy(), v64.copy()
            a[i, 4 + aa] *= np.exp(h)
            c[i, 4 + aa] *= np.exp(-h)
            fd = (lnz64(a, p64) - lnz64(c, p64)) / (2 * h)
            worst = max(worst, abs(fd - n_aa[i, aa]))
    scale = max(max(x.max() for x in counts.values()), n_aa.max())
    assert scale > 0.5  # the homolog is used
    assert worst <= 1e-4 * scale, worst / scale
