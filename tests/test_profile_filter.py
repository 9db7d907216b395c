"""The Gaussian letter filter of -d/-D (filterLetterProbabilities,
dummest-profile.hh:64-103, applied by finalizeProfile at :119-121) in the
engine's profile transform, against the oracle bit for bit, and pinned by
closed forms: on a profile whose positions are all alike, every letter's
signal (log probability less the position's mean log) is constant along the
profile, and a Gaussian whose weights sum to one leaves a constant unchanged,
so -d (signal minus its smoothing) flattens every position to the uniform
distribution and -D (which adds the mean back) gives back the profile."""
import gzip
import os

import numpy as np
import pytest

import dummest_amd as D
import oracle as O
from dummest_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SYN = {"A": 4, "G": 4, "P": 4, "T": 4, "V": 4, "R": 6, "L": 6, "S": 6, "I": 3, "M": 1, "W": 1}


def texts():
    yield "herpes", gzip.open(os.path.join(GOLD, "herpes.hmm.gz"), "rt").read(), 0
    met = gzip.open(os.path.join(GOLD, "MET-test.hmm.gz"), "rt").read()
    yield "met0", met, 0
    yield "met1", met, 1
    yield "syn", synth.make_profile(np.random.default_rng(3), 70, "syn")[0], 0


@pytest.mark.parametrize("sd,keep", [(2.0, False), (2.0, True), (5.5, False), (40.0, True), (300.0, False)])
def test_filter_matches_oracle(sd, keep):
    for name, text, idx in texts():
        M, v, p, b, _ = D.parse_hmm_text(text, index=idx, filter_std_dev=sd, keep_nonvarying=keep)
        ref = O.finalize(O.parse_hmm(text)[idx], np.float32, filter_std_dev=sd, keep_nonvarying=keep)
        np.testing.assert_array_equal(v, ref.values, err_msg=name)
        np.testing.assert_array_equal(p, ref.params, err_msg=name)
        np.testing.assert_array_equal(b, ref.bg, err_msg=name)
        plain = D.parse_hmm_text(text, index=idx)[1]
        assert not np.array_equal(v, plain), name  # the filter changed the profile


def letter_probs(values, stop=O.DEFAULT_OPTS[4]):
    """Per-position letter probabilities back from the finalized codon
    emissions ((1 - stop) p / #synonymous codons)."""
    n = np.array([SYN.get(a, 2) for a in synth.AA], np.float64)
    return values[:-1, 4:24].astype(np.float64) * n / (1 - stop)


def flat_profile_text(M, row):
    """M positions with the same emissions and transitions."""
    nl = lambda x: f"{-np.log(x):.5f}"
    lines = ["HMMER3/f [test]", "NAME  flat", f"LENG  {M}", "ALPH  amino",
             "HMM          " + "        ".join(synth.AA),
             "            m->m     m->i     m->d     i->m     i->i     d->m     d->d",
             "  COMPO   " + " ".join(nl(x) for x in row), "          " + " ".join(nl(0.05) for _ in row)]
    trans = "          " + " ".join(nl(x) for x in (0.9, 0.05, 0.05, 0.6, 0.4, 0.7, 0.3))
    lines.append(trans)
    for k in range(M):
        lines.append(f"{k + 1:7d}   " + " ".join(nl(x) for x in row) + f" {k + 1:6d} a - - -")
        lines.append("          " + " ".join(nl(0.05) for _ in row))
        lines.append(trans if k + 1 < M else "          " + " ".join(
            nl(x) for x in (0.95, 0.05)) + " * 0.60000 0.40000 0.00000 *")
    lines.append("//")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("sd", [2.0, 7.0, 30.0])
def test_filter_closed_forms(sd):
    rng = np.random.default_rng(int(sd))
    row = rng.dirichlet(np.ones(20) * 2)
    text = flat_profile_text(60, row)
    base = letter_probs(D.parse_hmm_text(text)[1])
    assert np.ptp(base, axis=0).max() < 1e-6  # every position alike
    hp = letter_probs(D.parse_hmm_text(text, filter_std_dev=sd)[1])
    np.testing.assert_allclose(hp, 1.0 / 20, rtol=2e-6)
    keep = letter_probs(D.parse_hmm_text(text, filter_std_dev=sd, keep_nonvarying=True)[1])
    np.testing.assert_allclose(keep, base, rtol=2e-6)
