"""dm_filter_scores_static, the entry point that allocates no device memory:
through the C ABI, its scores against dm_filter_scores' on the same pairs
(bit for bit: the same kernels), on two calls in a row, the second smaller,
with row-strip and single-strip profiles; inputs beyond its scratch are
refused with DM_EUNSUP. The caller's device buffers (codes in, scores out)
come from the HIP runtime the library itself loads."""
import ctypes

import numpy as np
import pytest

import dummest_amd as D
from tests.common import synthetic_case

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    D.lib()  # loads the HIP runtime libdummest_amd.so links
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def call_static(e, pid, codes):
    h = hip()
    offs = np.zeros(len(codes), np.int64)
    lens = np.array([len(c) for c in codes], np.int32)
    offs[1:] = np.cumsum(lens[:-1])
    flat = np.ascontiguousarray(np.concatenate(codes), np.uint8)
    best = np.full(len(codes), -1.0, np.float32)
    dc, db = ctypes.c_void_p(), ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(dc), flat.nbytes) == 0 and h.hipMalloc(ctypes.byref(db), best.nbytes) == 0
    try:
        assert h.hipMemcpy(dc, flat.ctypes.data, flat.nbytes, H2D) == 0
        assert h.hipMemcpy(db, best.ctypes.data, best.nbytes, H2D) == 0
        rc = D.lib().dm_filter_scores_static(e._h, pid, len(codes), dc, D._ptr(offs, ctypes.c_int64),
                                             D._ptr(lens, ctypes.c_int32), db)
        assert h.hipMemcpy(best.ctypes.data, db, best.nbytes, D2H) == 0
    finally:
        h.hipFree(dc)
        h.hipFree(db)
    return rc, best


def test_static_filter_matches_filter_scores():
    e = D.Engine(0)
    try:
        cases = []
        for seed, L, lens in ((71, 90, [150 + (31 * k) % 800 for k in range(700)]),
                              (72, 1300, [200 + (57 * k) % 1500 for k in range(60)])):
            text, prof, dnas = synthetic_case(seed, L, lens, homolog_every=3)
            M, v, p, b, _ = D.parse_hmm_text(text)
            pid = e.add_profile(M, v, p, b)
            codes = [D.decode(s) for s in dnas]
            sids = e.add_sequences(codes)
            cases.append((pid, codes, sids))
        for pid, codes, sids in cases:
            for n in (len(codes), len(codes) // 3):  # two calls in a row, the second smaller
                print(f"static filter: profile {pid}, {n} sequences", flush=True)
                want = e.find_similarities_backward_only([pid] * n, sids[:n])
                rc, got = call_static(e, pid, codes[:n])
                assert rc == 0, D.lib().dm_engine_error(e._h)
                np.testing.assert_array_equal(got, want)
        # beyond the static scratch (64 MiB): refused, and the engine still works
        pid, codes, sids = cases[0]
        big = [np.zeros(1 << 20, np.uint8) for _ in range(4)]
        rc, _ = call_static(e, pid, big)
        assert rc == D.DM_EUNSUP
        assert "scratch" in D.lib().dm_engine_error(e._h).decode()
        rc, got = call_static(e, pid, codes[:50])
        assert rc == 0
        np.testing.assert_array_equal(got, e.find_similarities_backward_only([pid] * 50, sids[:50]))
    finally:
        e.close()
