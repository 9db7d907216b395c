"""The profile reader (readProfiles, dummest-profile.hh:235-300, with
finalizeProfile): a text of many profiles is parsed and finalized record by
record on host threads; every profile must equal the one parsed from its own
record alone, in file order, with and without profile masking, and malformed
data must give the serial reader's error."""
import ctypes

import numpy as np
import pytest

import dummest_amd as D
from dummest_amd import synth


@pytest.fixture(scope="module")
def records():
    rng = np.random.default_rng(91)
    return [synth.make_profile(rng, int(L), f"r{k}", ndesc=6)[0] for k, L in
            enumerate(rng.integers(20, 160, size=24))]


@pytest.mark.parametrize("mask", [None, 0.5])
def test_many_profiles_equal_single_records(records, mask):
    text = "".join(records)
    for k, rec in enumerate(records):
        a = D.parse_hmm_text(text, index=k, mask_threshold=mask, with_masked=True)
        b = D.parse_hmm_text(rec, index=0, mask_threshold=mask, with_masked=True)
        assert a[4] == len(records) and b[4] == 1
        assert a[0] == b[0]
        for x, y in zip(a[1:4], b[1:4]):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), k
        assert a[5] == b[5]


def _parse_rc(text):
    r = np.array(D.DEFAULT_RATES, np.float64)
    M, cnt = ctypes.c_int32(), ctypes.c_int32()
    return D.lib().dm_parse_hmm_text(text.encode(), ord("G"), r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), -1.0,
                                     0.0, 0, 0, ctypes.byref(M), None, None, None, ctypes.byref(cnt), None), cnt.value


def test_malformed_record_gives_the_serial_error(records):
    # a transition line cut short in the 11th record of 24
    bad = records[10].splitlines(True)
    i = next(k for k, l in enumerate(bad) if l.strip().startswith("1 "))
    bad[i + 2] = "          0.1 0.2\n"
    rc, _ = _parse_rc("".join(records[:10]) + "".join(bad) + "".join(records[11:]))
    assert rc == D.DM_EFORMAT
    # a text ending inside a profile
    rc, _ = _parse_rc("".join(records) + records[0][: len(records[0]) // 2])
    assert rc == D.DM_EFORMAT
    # the complete text reads
    rc, cnt = _parse_rc("".join(records))
    assert rc == D.DM_OK and cnt == len(records)
