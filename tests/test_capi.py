"""C ABI checks that need no GPU: the library loads, exports every function
include/dummest_amd.h declares, and its host-side pieces (profile transform,
decoding) match the oracle bit for bit."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
import dummest_amd as D
from tests.common import golden_hmm, synthetic_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(dm_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(D.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("fixture", ["herpes.hmm.gz", "MET-test.hmm.gz"])
def test_profile_transform_matches_oracle(fixture):
    text = golden_hmm(fixture)
    for idx, rp in enumerate(O.parse_hmm(text)):
        M, v, p, b, cnt = D.parse_hmm_text(text, idx)
        ref = O.finalize(rp, np.float32)
        assert M == rp.length
        np.testing.assert_array_equal(v, ref.values)
        np.testing.assert_array_equal(p, ref.params)
        np.testing.assert_array_equal(b, ref.bg)


@pytest.mark.parametrize("bg_type", ["G", "A", "M"])
def test_background_types_match_oracle(bg_type):
    text, _, _ = synthetic_case(4, 90, [], 0)
    rp = O.parse_hmm(text)[0]
    M, v, p, b, _ = D.parse_hmm_text(text, 0, bg_type=bg_type)
    ref = O.finalize(rp, np.float32, bg_type=bg_type)
    np.testing.assert_array_equal(b, ref.bg)
    np.testing.assert_array_equal(p, ref.params)


def test_decode_matches_oracle():
    rng = np.random.default_rng(0)
    for n in (3, 10, 1000):
        s = "".join(rng.choice(list("ACGTacgtNRY"), size=n))
        np.testing.assert_array_equal(D.decode(s), O.decode(s))


def test_malformed_profile_is_rejected():
    with pytest.raises(D.DummestError):
        D.parse_hmm_text("NAME x\nHMM A C\n m->m\n  0.1 0.2\n")


def test_calibration_sequences_match_oracle():
    text, p, _ = synthetic_case(3, 50, [], 0)
    bg = np.asarray(p.bg, np.float32)
    for L, border, trial in [(500, 100, 0), (1000, 0, 7), (31, 5, 123)]:
        assert D.calib_sequence(bg, L, border, trial) == O.calib_sequence(bg, L, border, trial)
    rates = (0.03, 0.01, 0.02, 0.004, 0.0005, 0.02)
    assert D.calib_sequence(bg, 800, 50, 3, rates) == O.calib_sequence(bg, 800, 50, 3, rates)


def test_gumbel_fits_match_oracle():
    rng = np.random.default_rng(4)
    x = 5.0 + rng.gumbel(0, 1 / 0.8, size=3000)
    a = D.gumbel_fits(x, 1000)
    b = O.gumbel(x, 1000)
    for k in D.GUMBEL_KEYS:
        assert a[k] == pytest.approx(b[k], rel=1e-12), k
