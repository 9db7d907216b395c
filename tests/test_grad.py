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
