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


def test_hits_match_oracle(engine):
    print("hits L=150", flush=True)
    pid, prof, codes, sids = load_case(engine, 300, 150, [1500, 2500, 900], homolog_every=1)
    for c, s in zip(codes, sids):
        ref = O.search(prof, c, 1.0)
        got = engine.find_similarities(pid, s, 1.0)
        assert len(got) == len(ref.hits)
        for g, r in zip(got, ref.hits):
            assert abs(g[0] - r[0]) < TOL_NATS
            assert (g[3], g[4]) == (r[3], r[4])
            assert g[5] == r[5]
