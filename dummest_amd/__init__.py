"""dummest_amd -- MI355X (gfx950) implementation of DUMMEST's frameshift-aware
DNA-vs-profile-HMM forward/backward DP.

Python binding over the C ABI of ``libdummest_amd.so`` (declared in
``include/dummest_amd.h``). Names follow the reference's operator interface:

* ``Engine.find_similarities_backward_only`` -- score-only filter DP
  (dummest-dp-backward.hh:300-438)
* ``Engine.fwdbwd_scores`` -- forward+backward best anchored score and
  backward start score (findSimilarities with needAlign=false,
  dummest-search.hh:9-17, as used by estimateK)
* ``Engine.find_similarities`` -- local hits with alignments of one pair
  (findSimilarities with needAlign=true + traceback)

There is no CPU fallback: loading fails loudly when the HIP library is absent.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdummest_amd.so")
WIDTH = 29
ALPHABET = "ACDEFGHIKLMNPQRSTVWYUO?*"
DEFAULT_RATES = (0.0171, 0.0018, 0.0328, 0.0083, 0.0005, 0.046875)
# status codes of the C ABI (include/dummest_amd.h)
DM_OK, DM_EINVAL, DM_EHIP, DM_EFORMAT, DM_EUNSUP, DM_EIO, DM_ERANGE = 0, -1, -2, -3, -4, -5, -6

_lib = None


class DummestError(RuntimeError):
    """A failing call of the C ABI; ``code`` is its DM_E* status (None when
    the failure did not come from a status code)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def lib():
    """Load libdummest_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DummestError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "or `make -C dummest_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _P(t):
    return ctypes.POINTER(t)


def _declare(L):
    i32, i64, f, d, u8, vp = (ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double, ctypes.c_uint8,
                             ctypes.c_void_p)
    L.dm_engine_create.argtypes = [ctypes.c_int, _P(vp)]
    L.dm_engine_destroy.argtypes = [vp]
    L.dm_engine_destroy.restype = None
    L.dm_engine_error.argtypes = [vp]
    L.dm_engine_error.restype = ctypes.c_char_p
    L.dm_set_rates.argtypes = [vp, _P(d)]
    L.dm_read_hmm.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, _P(i32), _P(i32)]
    L.dm_parse_hmm_text.argtypes = [ctypes.c_char_p, ctypes.c_int, _P(d), d, d, ctypes.c_int, i32, _P(i32), _P(f),
                                    _P(f), _P(f), _P(i32), _P(i32)]
    L.dm_set_profile_filter.argtypes = [vp, d, ctypes.c_int]
    L.dm_add_raw_profiles.argtypes = [vp, i32, _P(i32), _P(f), ctypes.c_int, _P(i32)]
    L.dm_add_profile.argtypes = [vp, i32, _P(f), _P(f), _P(f), _P(i32)]
    L.dm_set_profile.argtypes = [vp, i32, i32, _P(f), _P(f), _P(f)]
    L.dm_get_profile.argtypes = [vp, i32, _P(i32), _P(f), _P(f), _P(f)]
    L.dm_profile_count.argtypes = [vp]
    L.dm_decode.argtypes = [ctypes.c_char_p, i64, _P(u8)]
    L.dm_add_sequences.argtypes = [vp, i32, _P(u8), _P(i64), _P(i32), _P(i32)]
    L.dm_clear_sequences.argtypes = [vp]
    L.dm_set_masking.argtypes = [vp, ctypes.c_int, d]
    L.dm_profile_masked.argtypes = [vp, i32]
    L.dm_tantan_dna.argtypes = [vp, i32, ctypes.c_char_p, _P(i64), _P(i32), _P(f)]
    L.dm_add_dna.argtypes = [vp, i32, ctypes.c_char_p, _P(i64), _P(i32), _P(i32)]
    L.dm_filter_scores.argtypes = [vp, i64, _P(i32), _P(i32), _P(f)]
    L.dm_filter_scores_static.argtypes = [vp, i32, i32, vp, _P(i64), _P(i32), vp]
    L.dm_fwdbwd_scores.argtypes = [vp, i64, _P(i32), _P(i32), ctypes.c_int, _P(f), _P(i32), _P(f), _P(i32)]
    L.dm_pair_matrices.argtypes = [vp, i32, i32, _P(f), _P(i32), _P(f), _P(i32)]
    L.dm_pair_null.argtypes = [vp, i32, i32, _P(f), _P(f), _P(f)]
    L.dm_pair_hits.argtypes = [vp, i32, i32, d, _P(d), i32, _P(i32), i32, _P(i32), _P(i32)]
    L.dm_pairs_hits.argtypes = [vp, i64, _P(i32), _P(i32), _P(d), _P(d), i64, _P(i32), i64, _P(i64), _P(i64),
                                _P(i64)]
    L.dm_last_hits_ms.argtypes = [vp]
    L.dm_last_hits_ms.restype = d
    L.dm_plan_create.argtypes = [vp, ctypes.c_int, i64, _P(i32), _P(i32), ctypes.c_int, _P(vp)]
    L.dm_plan_execute.argtypes = [vp, vp, _P(f), _P(f), _P(i32), _P(f), _P(i32)]
    L.dm_plan_capture.argtypes = [vp, vp]
    L.dm_plan_destroy.argtypes = [vp, vp]
    L.dm_plan_destroy.restype = None
    L.dm_kernel_stats.argtypes = [vp, i32, _P(i32), _P(i32), _P(i32), _P(d), _P(i64), _P(i64)]
    L.dm_copy_probe.argtypes = [vp, i64, _P(d)]
    L.dm_build_id.argtypes = []
    L.dm_build_id.restype = ctypes.c_char_p
    L.dm_post_process.argtypes = [i32, _P(f), _P(f), _P(f), i32, _P(u8), _P(f), _P(i32), _P(f), _P(i32), _P(f),
                                  _P(f), _P(d), d, _P(d), i32, _P(i32), i32, _P(i32), _P(i32)]
    L.dm_set_option.argtypes = [vp, ctypes.c_char_p, i64]
    L.dm_calibrate.argtypes = [vp, i32, _P(i32), i32, i32, i32, _P(d), _P(d), _P(d)]
    L.dm_calibrate_anchors.argtypes = [vp, i32, _P(i32), i32, i32, i32, _P(i32), _P(d)]
    L.dm_autotune.argtypes = [vp, i32, _P(d), i32]
    L.dm_set_config_costs.argtypes = [vp, i32, _P(d)]
    L.dm_config_table.argtypes = [vp, i32, _P(i32), _P(i32), _P(i32), _P(d)]
    L.dm_calib_sequence.argtypes = [_P(f), _P(d), i32, i32, i32, ctypes.c_char_p]
    L.dm_gumbel_fits.argtypes = [_P(d), i32, d, _P(d)]
    L.dm_last_dp_ms.argtypes = [vp]
    L.dm_last_dp_ms.restype = d
    L.dm_last_tantan_ms.argtypes = [vp]
    L.dm_last_tantan_ms.restype = d
    L.dm_last_cells.argtypes = [vp]
    L.dm_last_cells.restype = i64
    L.dm_bw_prepare.argtypes = [vp, i32, i32, i32, _P(u8), _P(i64), _P(i32), _P(d)]
    L.dm_bw_estep.argtypes = [vp, _P(d), _P(d), _P(d)]
    L.dm_last_bw_ms.argtypes = [vp]
    L.dm_last_bw_ms.restype = d
    L.dm_grad.argtypes = [vp, i32, i32, ctypes.c_char_p, _P(i64), _P(i32), _P(d), _P(d), _P(d)]
    L.dm_last_grad_ms.argtypes = [vp]
    L.dm_last_grad_ms.restype = d
    L.dm_last_grad_cells.argtypes = [vp]
    L.dm_last_grad_cells.restype = i64
    L.dm_last_grad_segments.argtypes = [vp]
    L.dm_last_grad_segments.restype = i64
    L.dm_grad_set_create.argtypes = [vp, i32, ctypes.c_char_p, _P(i64), _P(i32), _P(i32)]
    L.dm_grad_set.argtypes = [vp, i32, i32, _P(d), _P(d), _P(d)]
    L.dm_grad_set_destroy.argtypes = [vp, i32]
    L.dm_xxh3_128.argtypes = [ctypes.c_char_p, i64, _P(ctypes.c_uint64)]
    L.dm_calibration_key.argtypes = [ctypes.c_char_p, i32, _P(f), ctypes.c_char_p, i32, i32, i32, _P(d), d,
                                     ctypes.c_char_p]
    L.dm_cache_file_path.argtypes = [ctypes.c_char_p, i32]
    L.dm_cache_read.argtypes = [ctypes.c_char_p, i64, ctypes.c_char_p, i32, _P(d)]
    L.dm_cache_read.restype = i64
    L.dm_cache_write.argtypes = [ctypes.c_char_p, i64, ctypes.c_char_p, i32, _P(d)]


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def decode(dna) -> np.ndarray:
    """Codon codes of a DNA string (decodeSequence, dummest-sequence.hh:181-204)."""
    b = dna.encode() if isinstance(dna, str) else bytes(dna)
    out = np.zeros(len(b), np.uint8)
    rc = lib().dm_decode(b, len(b), _ptr(out, ctypes.c_uint8))
    if rc:
        raise DummestError(f"dm_decode failed ({rc})")
    return out


def parse_hmm_text(text: str, index: int = 0, bg_type: str = "G", rates=DEFAULT_RATES, mask_threshold=None,
                   with_masked=False, filter_std_dev=0.0, keep_nonvarying=False):
    """Host-side readProfiles + finalizeProfile of one profile of an HMMER3/f text
    (no device needed): returns (M, values, params, bg, count), plus the number
    of masked positions when with_masked. mask_threshold turns on profile
    masking (-m 1); filter_std_dev the Gaussian letter filter (-d, with
    keep_nonvarying -D)."""
    r = np.array(rates, np.float64)
    M, cnt, nm = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    t = text.encode()
    thr = -1.0 if mask_threshold is None else float(mask_threshold)
    flt = (float(filter_std_dev), int(bool(keep_nonvarying)))
    rc = lib().dm_parse_hmm_text(t, ord(bg_type), _ptr(r, ctypes.c_double), thr, *flt, index, ctypes.byref(M), None,
                                 None, None, ctypes.byref(cnt), None)
    if rc:
        raise DummestError(f"dm_parse_hmm_text failed ({rc})")
    m = M.value
    v = np.zeros((m + 1, WIDTH), np.float32)
    p = np.zeros((m + 2, 11), np.float32)
    b = np.zeros(WIDTH, np.float32)
    f = ctypes.c_float
    rc = lib().dm_parse_hmm_text(t, ord(bg_type), _ptr(r, ctypes.c_double), thr, *flt, index, ctypes.byref(M),
                                 _ptr(v, f), _ptr(p, f), _ptr(b, f), ctypes.byref(cnt), ctypes.byref(nm))
    if rc:
        raise DummestError(f"dm_parse_hmm_text failed ({rc})")
    if with_masked:
        return m, v, p, b, cnt.value, nm.value
    return m, v, p, b, cnt.value


def post_process(M, values, params, bg, codes, w0, W1, prefix_seed, scales, min_prob_ratio,
                 rates=DEFAULT_RATES, hit_cap=4096, seg_cap=1 << 20):
    """Host post-gate processing on given w0/W1 ExpScore matrices (no device)."""
    f, i32, dd = ctypes.c_float, ctypes.c_int32, ctypes.c_double
    arr = lambda a, t: np.ascontiguousarray(a, t)
    v, p, b = arr(values, np.float32), arr(params, np.float32), arr(bg, np.float32)
    c = arr(codes, np.uint8)
    w0m, w0e = arr(w0[0], np.float32), arr(w0[1], np.int32)
    W1m, W1e = arr(W1[0], np.float32), arr(W1[1], np.int32)
    ps, sc = arr(prefix_seed, np.float32), arr(scales, np.float32)
    r = np.array(rates, np.float64)
    hits = np.zeros((hit_cap, 6), np.float64)
    segs = np.zeros((seg_cap, 3), np.int32)
    nh, ns = ctypes.c_int32(), ctypes.c_int32()
    rc = lib().dm_post_process(int(M), _ptr(v, f), _ptr(p, f), _ptr(b, f), len(c), _ptr(c, ctypes.c_uint8),
                               _ptr(w0m, f), _ptr(w0e, i32), _ptr(W1m, f), _ptr(W1e, i32), _ptr(ps, f), _ptr(sc, f),
                               _ptr(r, dd), float(min_prob_ratio), _ptr(hits, dd), hit_cap, _ptr(segs, i32), seg_cap,
                               ctypes.byref(nh), ctypes.byref(ns))
    if rc:
        raise DummestError(f"dm_post_process failed ({rc})")
    return _unpack_hits(hits, segs, nh.value)


def _unpack_hits(hits, segs, nh):
    out = []
    k = 0
    for h in hits[:nh]:
        c = int(h[5])
        out.append((float(h[0]), float(h[1]), int(h[2]), int(h[3]), int(h[4]),
                    [tuple(int(x) for x in r) for r in segs[k:k + c]]))
        k += c
    return out


GUMBEL_KEYS = ("mmL", "mmK", "mmKsimple", "mlL", "mlK", "mlKsimple", "lmL", "lmK")


def calib_sequence(bg, length: int, border: int, trial: int, rates=DEFAULT_RATES) -> bytes:
    """DNA letters of one calibration trial (host only, no device)."""
    b = np.ascontiguousarray(bg, np.float32)
    r = np.array(rates, np.float64)
    buf = ctypes.create_string_buffer(length + border)
    rc = lib().dm_calib_sequence(_ptr(b, ctypes.c_float), _ptr(r, ctypes.c_double), length, border, trial, buf)
    if rc:
        raise DummestError(f"dm_calib_sequence failed ({rc})")
    return buf.raw[:length + border]


def gumbel_fits(scores, seq_len) -> dict:
    """estimateGumbel of a score sample (host only)."""
    s = np.ascontiguousarray(scores, np.float64)
    out = np.zeros(8, np.float64)
    rc = lib().dm_gumbel_fits(_ptr(s, ctypes.c_double), len(s), float(seq_len), _ptr(out, ctypes.c_double))
    if rc:
        raise DummestError(f"dm_gumbel_fits failed ({rc})")
    return dict(zip(GUMBEL_KEYS, out.tolist()))


def xxh3_128(data: bytes) -> int:
    """XXH3 128-bit hash (seed 0) as one integer, high half in the top 64 bits."""
    out = (ctypes.c_uint64 * 2)()
    rc = lib().dm_xxh3_128(data, len(data), out)
    if rc:
        raise DummestError(f"dm_xxh3_128 failed ({rc})")
    return out[0] | (out[1] << 64)


# Columns of dm_grad's counts: the 11 Params of a row, then 65 codon bins.
GRAD_WIDTH = 76

# Field order of a calibration-cache entry (CacheEntry, dummest-calibration.hh:148-181).
CACHE_FIELDS = tuple(f"{m}{a}{q}" for m in ("MM", "ML", "LM") for q in ("L", "K", "Ksimple")
                     for a in ("end", "beg", "mid") if not (m == "LM" and q == "Ksimple")) + (
    "flt_mid_l", "flt_mid_k", "flt_mid_k_simple")


def calibration_key(name: str, M: int, values, seq_len: int, border: int, trials: int, rates=DEFAULT_RATES,
                    tantan_threshold: float = 0.5, binary_hash: str = "") -> str:
    """computeCacheKey (dummest-calibration.hh:201-226) of a finalized profile
    (values (M+1) x 29) and the calibration options."""
    v = np.ascontiguousarray(values, np.float32)
    if v.size != 29 * (M + 1):
        raise ValueError("values must hold (M+1) x 29 floats")
    r = np.array(rates, np.float64)
    out = ctypes.create_string_buffer(33)
    rc = lib().dm_calibration_key(name.encode(), M, _ptr(v, ctypes.c_float), binary_hash.encode(), seq_len, border,
                                  trials, _ptr(r, ctypes.c_double), float(tantan_threshold), out)
    if rc:
        raise DummestError(f"dm_calibration_key failed ({rc})")
    return out.value.decode()


def cache_file_path() -> str:
    """The calibration cache file the reference's search order picks (creates its directory)."""
    out = ctypes.create_string_buffer(4096)
    rc = lib().dm_cache_file_path(out, 4096)
    if rc:
        raise DummestError(f"dm_cache_file_path failed ({rc})")
    return out.value.decode()


def cache_read(path: str):
    """Entries of a calibration cache file, in file order: [(key, 27 doubles)]."""
    n = lib().dm_cache_read(path.encode(), 0, None, 0, None)
    if n < 0:
        raise DummestError(f"dm_cache_read failed ({n})")
    stride = 256
    keys = ctypes.create_string_buffer(max(1, n) * stride)
    vals = np.zeros((max(1, n), 27), np.float64)
    rc = lib().dm_cache_read(path.encode(), n, keys, stride, _ptr(vals, ctypes.c_double))
    if rc != n:
        raise DummestError(f"dm_cache_read failed ({rc})")
    raw = keys.raw
    return [(raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode(), vals[i].copy()) for i in range(n)]


def cache_write(path: str, entries) -> None:
    """Writes [(key, 27 doubles)] as a calibration cache file, in the given order."""
    stride = 256
    n = len(entries)
    keys = ctypes.create_string_buffer(max(1, n) * stride)
    vals = np.zeros((max(1, n), 27), np.float64)
    for i, (k, v) in enumerate(entries):
        kb = k.encode()
        if len(kb) >= stride:
            raise ValueError("key too long")
        ctypes.memmove(ctypes.addressof(keys) + i * stride, kb, len(kb))
        vals[i] = np.asarray(v, np.float64)
    rc = lib().dm_cache_write(path.encode(), n, keys, stride, _ptr(vals, ctypes.c_double))
    if rc:
        raise DummestError(f"dm_cache_write failed ({rc})")


def build_id() -> str:
    """Hash of the kernel sources the loaded library was built from."""
    return lib().dm_build_id().decode()


def es_log(m, e):
    """Natural log of an ExpScore (scoreToLog, dummest-score.hh:202-207)."""
    m = np.asarray(m, np.float64)
    e = np.asarray(e, np.float64)
    with np.errstate(divide="ignore"):
        return np.where(m == 0, -np.inf, np.log(np.where(m == 0, 1.0, m)) + e * math.log(2.0))


class Engine:
    """One HIP stream on one device plus its uploaded profiles and sequences."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().dm_engine_create(device, ctypes.byref(h))
        if rc:
            raise DummestError(f"dm_engine_create failed ({rc}): no usable HIP device {device}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().dm_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            msg = lib().dm_engine_error(self._h).decode(errors="replace")
            raise DummestError(f"{what} failed ({rc}): {msg}", rc)

    # -- configuration ------------------------------------------------------
    def set_rates(self, rates=DEFAULT_RATES):
        r = np.array(rates, np.float64)
        self._check(lib().dm_set_rates(self._h, _ptr(r, ctypes.c_double)), "dm_set_rates")

    def set_option(self, name: str, value: int):
        self._check(lib().dm_set_option(self._h, name.encode(), int(value)), "dm_set_option")

    # -- profiles -----------------------------------------------------------
    def read_hmm(self, path: str, bg_type: str = "G") -> list:
        first, count = ctypes.c_int32(), ctypes.c_int32()
        self._check(lib().dm_read_hmm(self._h, path.encode(), ord(bg_type), ctypes.byref(first), ctypes.byref(count)),
                    "dm_read_hmm")
        return list(range(first.value, first.value + count.value))

    def add_raw_profiles(self, Ms, raws, bg_type: str = "G") -> list:
        """Profiles from reader-order probabilities ((M+1) x 29 each); returns ids."""
        Ms = np.ascontiguousarray(Ms, np.int32)
        raw = np.ascontiguousarray(np.concatenate([np.asarray(r, np.float32).ravel() for r in raws]), np.float32)
        first = ctypes.c_int32()
        self._check(lib().dm_add_raw_profiles(self._h, len(Ms), _ptr(Ms, ctypes.c_int32), _ptr(raw, ctypes.c_float),
                                              ord(bg_type), ctypes.byref(first)), "dm_add_raw_profiles")
        return list(range(first.value, first.value + len(Ms)))

    def add_profile(self, M, values, params, bg) -> int:
        v = np.ascontiguousarray(values, np.float32)
        p = np.ascontiguousarray(params, np.float32)
        b = np.ascontiguousarray(bg, np.float32)
        assert v.size == (M + 1) * WIDTH and p.size == (M + 2) * 11 and b.size == WIDTH
        out = ctypes.c_int32()
        f = ctypes.c_float
        self._check(lib().dm_add_profile(self._h, int(M), _ptr(v, f), _ptr(p, f), _ptr(b, f), ctypes.byref(out)),
                    "dm_add_profile")
        return out.value

    def set_profile(self, pid: int, M, values, params, bg):
        """Replace profile pid with new finalized arrays."""
        v = np.ascontiguousarray(values, np.float32)
        p = np.ascontiguousarray(params, np.float32)
        b = np.ascontiguousarray(bg, np.float32)
        assert v.size == (M + 1) * WIDTH and p.size == (M + 2) * 11 and b.size == WIDTH
        f = ctypes.c_float
        self._check(lib().dm_set_profile(self._h, pid, int(M), _ptr(v, f), _ptr(p, f), _ptr(b, f)), "dm_set_profile")

    def get_profile(self, pid: int):
        M = ctypes.c_int32()
        self._check(lib().dm_get_profile(self._h, pid, ctypes.byref(M), None, None, None), "dm_get_profile")
        m = M.value
        v = np.zeros((m + 1, WIDTH), np.float32)
        p = np.zeros((m + 2, 11), np.float32)
        b = np.zeros(WIDTH, np.float32)
        f = ctypes.c_float
        self._check(lib().dm_get_profile(self._h, pid, ctypes.byref(M), _ptr(v, f), _ptr(p, f), _ptr(b, f)),
                    "dm_get_profile")
        return m, v, p, b

    def profile_length(self, pid: int) -> int:
        M = ctypes.c_int32()
        self._check(lib().dm_get_profile(self._h, pid, ctypes.byref(M), None, None, None), "dm_get_profile")
        return M.value

    # -- sequences ----------------------------------------------------------
    def add_sequences(self, seqs) -> list:
        """Append decoded sequences (arrays of codon codes); returns their ids."""
        lens = np.array([len(s) for s in seqs], np.int32)
        offs = np.zeros(len(seqs), np.int64)
        if len(seqs) > 1:
            offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
        codes = (np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if len(seqs)
                 else np.zeros(1, np.uint8))
        codes = np.ascontiguousarray(codes)
        first = ctypes.c_int32()
        self._check(lib().dm_add_sequences(self._h, len(seqs), _ptr(codes, ctypes.c_uint8),
                                           _ptr(offs, ctypes.c_int64), _ptr(lens, ctypes.c_int32),
                                           ctypes.byref(first)), "dm_add_sequences")
        return list(range(first.value, first.value + len(seqs)))

    def clear_sequences(self):
        self._check(lib().dm_clear_sequences(self._h), "dm_clear_sequences")

    # -- tantan masking ------------------------------------------------------
    def set_masking(self, mask: int, threshold: float = 0.5):
        """-m (1 profiles, 2 sequences, 3 both) and --tantan-threshold."""
        self._check(lib().dm_set_masking(self._h, int(mask), float(threshold)), "dm_set_masking")

    def set_profile_filter(self, std_dev: float = 0.0, keep_nonvarying: bool = False):
        """Gaussian letter filter (-d, or -D with keep_nonvarying) of the profiles
        read or added afterwards; 0 turns it off."""
        self._check(lib().dm_set_profile_filter(self._h, float(std_dev), int(bool(keep_nonvarying))),
                    "dm_set_profile_filter")

    def profile_masked(self, pid: int) -> int:
        """Positions of profile pid masked by tantan."""
        return int(lib().dm_profile_masked(self._h, pid))

    @staticmethod
    def _dna_batch(seqs):
        bs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
        lens = np.array([len(b) for b in bs], np.int32)
        offs = np.zeros(len(bs), np.int64)
        if len(bs) > 1:
            offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
        return b"".join(bs) or b"\0", offs, lens

    def tantan_dna(self, seqs) -> list:
        """tantan repeat probabilities of DNA letters, computed on the device."""
        data, offs, lens = self._dna_batch(seqs)
        out = np.zeros(max(1, int(lens.sum())), np.float32)
        self._check(lib().dm_tantan_dna(self._h, len(seqs), data, _ptr(offs, ctypes.c_int64),
                                        _ptr(lens, ctypes.c_int32), _ptr(out, ctypes.c_float)), "dm_tantan_dna")
        return [out[o:o + n] for o, n in zip(offs, lens)]

    def add_dna(self, seqs) -> list:
        """Append DNA given as letters (masked on the device under mask bit 2,
        then decoded); returns their ids."""
        data, offs, lens = self._dna_batch(seqs)
        first = ctypes.c_int32()
        self._check(lib().dm_add_dna(self._h, len(seqs), data, _ptr(offs, ctypes.c_int64),
                                     _ptr(lens, ctypes.c_int32), ctypes.byref(first)), "dm_add_dna")
        return list(range(first.value, first.value + len(seqs)))

    # -- profile builder -----------------------------------------------------
    def bw_prepare(self, A: int, L: int, seqs, weights):
        """Upload one alignment's ungapped sequences (letter codes 0..A) and
        weights for Baum-Welch against a profile of L positions."""
        lens = np.array([len(x) for x in seqs], np.int32)
        offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64) if len(seqs) else np.zeros(0, np.int64)
        data = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint8) for x in seqs] + [np.zeros(1, np.uint8)]))
        w = np.ascontiguousarray(weights, np.float64)
        self._bw = (A, L, len(seqs))
        self._check(lib().dm_bw_prepare(self._h, A, L, len(seqs), _ptr(data, ctypes.c_uint8),
                                        _ptr(offs, ctypes.c_int64), _ptr(lens, ctypes.c_int32),
                                        _ptr(w, ctypes.c_double)), "dm_bw_prepare")

    def bw_estep(self, probs):
        """One Baum-Welch E-step on the device: (counts (L+1) x (7+A), forward totals)."""
        A, L, n = self._bw
        p = np.ascontiguousarray(probs, np.float64)
        counts = np.zeros((L + 1, 7 + A), np.float64)
        totals = np.zeros(max(1, n), np.float64)
        self._check(lib().dm_bw_estep(self._h, _ptr(p, ctypes.c_double), _ptr(counts, ctypes.c_double),
                                      _ptr(totals, ctypes.c_double)), "dm_bw_estep")
        return counts, totals[:n]

    @property
    def last_bw_ms(self) -> float:
        return lib().dm_last_bw_ms(self._h)

    # -- gradient of the summed log-likelihood --------------------------------
    def grad_counts(self, prof: int, seqs, weights=None):
        """(ln Z per sequence, expected counts (M+1) x 76) of one profile over
        DNA sequences (see dm_grad)."""
        data, offs, lens = self._dna_batch(seqs)
        M = self.profile_length(prof)
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        lnz = np.zeros(max(1, len(seqs)), np.float64)
        counts = np.zeros((M + 1, GRAD_WIDTH), np.float64)
        self._check(lib().dm_grad(self._h, prof, len(seqs), data, _ptr(offs, ctypes.c_int64), _ptr(lens, ctypes.c_int32),
                                  None if w is None else _ptr(w, ctypes.c_double), _ptr(lnz, ctypes.c_double),
                                  _ptr(counts, ctypes.c_double)), "dm_grad")
        return lnz[:len(seqs)], counts

    def grad_set(self, seqs) -> int:
        """Register DNA sequences for repeated gradient evaluations (see
        dm_grad_set_create); returns the set id."""
        data, offs, lens = self._dna_batch(seqs)
        out = ctypes.c_int32()
        self._check(lib().dm_grad_set_create(self._h, len(seqs), data, _ptr(offs, ctypes.c_int64),
                                             _ptr(lens, ctypes.c_int32), ctypes.byref(out)), "dm_grad_set_create")
        self._grad_set_sizes = getattr(self, "_grad_set_sizes", {})
        self._grad_set_sizes[out.value] = len(seqs)
        return out.value

    def grad_set_counts(self, prof: int, set_id: int, weights=None):
        """grad_counts over a registered set (bit-identical results)."""
        n = self._grad_set_sizes[set_id]
        M = self.profile_length(prof)
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        lnz = np.zeros(max(1, n), np.float64)
        counts = np.zeros((M + 1, GRAD_WIDTH), np.float64)
        self._check(lib().dm_grad_set(self._h, prof, set_id, None if w is None else _ptr(w, ctypes.c_double),
                                      _ptr(lnz, ctypes.c_double), _ptr(counts, ctypes.c_double)), "dm_grad_set")
        return lnz[:n], counts

    def grad_set_destroy(self, set_id: int):
        self._check(lib().dm_grad_set_destroy(self._h, set_id), "dm_grad_set_destroy")

    @property
    def last_grad_ms(self) -> float:
        return lib().dm_last_grad_ms(self._h)

    @property
    def last_grad_cells(self) -> int:
        return lib().dm_last_grad_cells(self._h)

    @property
    def last_grad_segments(self) -> int:
        """Strip segments the last gradient call recomputed from checkpoint rows."""
        return lib().dm_last_grad_segments(self._h)

    # -- scoring ------------------------------------------------------------
    @staticmethod
    def _pairs(prof, seq):
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(prof, np.int32), np.shape(seq)), np.int32)
        s = np.ascontiguousarray(seq, np.int32)
        return p.ravel(), s.ravel()

    def find_similarities_backward_only(self, prof, seq) -> np.ndarray:
        """Filter score per pair: max over cells of W1 * prefixSeed, FP32, no rescale."""
        p, s = self._pairs(prof, seq)
        out = np.zeros(len(p), np.float32)
        i32 = ctypes.c_int32
        self._check(lib().dm_filter_scores(self._h, len(p), _ptr(p, i32), _ptr(s, i32), _ptr(out, ctypes.c_float)),
                    "dm_filter_scores")
        return out

    def fwdbwd_scores(self, prof, seq, want_bwd=True):
        """(mid_m, mid_e, bwd_m, bwd_e): best anchored score and backward start score."""
        p, s = self._pairs(prof, seq)
        n = len(p)
        mm = np.zeros(n, np.float32)
        me = np.zeros(n, np.int32)
        bm = np.zeros(n, np.float32)
        be = np.zeros(n, np.int32)
        i32, f = ctypes.c_int32, ctypes.c_float
        self._check(lib().dm_fwdbwd_scores(self._h, n, _ptr(p, i32), _ptr(s, i32), int(want_bwd), _ptr(mm, f),
                                           _ptr(me, i32), _ptr(bm, f), _ptr(be, i32)), "dm_fwdbwd_scores")
        return mm, me, bm, be

    def pair_matrices(self, prof: int, seq: int, n: int):
        M = self.profile_length(prof)
        w0m = np.zeros((M + 1, n), np.float32)
        w0e = np.zeros((M + 1, n), np.int32)
        W1m = np.zeros((M + 1, n), np.float32)
        W1e = np.zeros((M + 1, n), np.int32)
        i32, f = ctypes.c_int32, ctypes.c_float
        self._check(lib().dm_pair_matrices(self._h, prof, seq, _ptr(w0m, f), _ptr(w0e, i32), _ptr(W1m, f),
                                           _ptr(W1e, i32)), "dm_pair_matrices")
        return (w0m, w0e), (W1m, W1e)

    def pair_null(self, prof: int, seq: int, n: int):
        ps = np.zeros(n, np.float32)
        ss = np.zeros(n, np.float32)
        sc = np.zeros(3, np.float32)
        f = ctypes.c_float
        self._check(lib().dm_pair_null(self._h, prof, seq, _ptr(ps, f), _ptr(ss, f), _ptr(sc, f)), "dm_pair_null")
        return ps, ss, sc

    def find_similarities_batch(self, prof, seq, min_prob_ratio, hit_cap=4096, seg_cap=1 << 16):
        """Hits of many pairs in one call (dm_pairs_hits): per pair, the list
        find_similarities returns for it. min_prob_ratio: one value or one per pair."""
        p, s = self._pairs(prof, seq)
        n = len(p)
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(min_prob_ratio, np.float64), (n,)))
        first = np.zeros(n + 1, np.int64)
        nh, ns = ctypes.c_int64(), ctypes.c_int64()
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        while True:
            hits = np.zeros((max(1, hit_cap), 6), np.float64)
            segs = np.zeros((max(1, seg_cap), 3), np.int32)
            rc = lib().dm_pairs_hits(self._h, n, _ptr(p, i32), _ptr(s, i32), _ptr(r, ctypes.c_double),
                                     _ptr(hits, ctypes.c_double), hit_cap, _ptr(segs, i32), seg_cap,
                                     _ptr(first, i64), ctypes.byref(nh), ctypes.byref(ns))
            if rc != DM_ERANGE:
                break
            hit_cap, seg_cap = max(hit_cap, nh.value), max(seg_cap, ns.value)
        self._check(rc, "dm_pairs_hits")
        out, k = [], 0
        for q in range(n):
            lst = []
            for h in hits[first[q]:first[q + 1]]:
                c = int(h[5])
                lst.append((float(h[0]), float(h[1]), int(h[2]), int(h[3]), int(h[4]),
                            [tuple(int(x) for x in segs[k + g]) for g in range(c)]))
                k += c
            out.append(lst)
        return out

    @property
    def last_hits_ms(self) -> float:
        return lib().dm_last_hits_ms(self._h)

    def find_similarities(self, prof: int, seq: int, min_prob_ratio: float, hit_cap=4096, seg_cap=1 << 20):
        """Hits of one pair: list of (lnratio, m, e, anchor1, anchor2, [(start1, start2, length)])."""
        nh, ns = ctypes.c_int32(), ctypes.c_int32()
        i32 = ctypes.c_int32
        while True:
            hits = np.zeros((hit_cap, 6), np.float64)
            segs = np.zeros((seg_cap, 3), np.int32)
            rc = lib().dm_pair_hits(self._h, prof, seq, float(min_prob_ratio), _ptr(hits, ctypes.c_double),
                                    hit_cap, _ptr(segs, i32), seg_cap, ctypes.byref(nh), ctypes.byref(ns))
            if rc != DM_ERANGE:
                break
            # the buffers were too small: the call reported the sizes it needs
            hit_cap, seg_cap = max(hit_cap, nh.value), max(seg_cap, ns.value)
        self._check(rc, "dm_pair_hits")
        out = []
        k = 0
        for h in hits[:nh.value]:
            c = int(h[5])
            out.append((float(h[0]), float(h[1]), int(h[2]), int(h[3]), int(h[4]),
                        [tuple(int(x) for x in r) for r in segs[k:k + c]]))
            k += c
        return out

    def plan(self, prof, seq, mode="score", want_bwd=False):
        """Build a device-resident work queue for the pairs; returns a Plan."""
        p, s = self._pairs(prof, seq)
        h = ctypes.c_void_p()
        i32 = ctypes.c_int32
        m = 1 if mode == "score" else 0
        self._check(lib().dm_plan_create(self._h, m, len(p), _ptr(p, i32), _ptr(s, i32), int(want_bwd),
                                         ctypes.byref(h)), "dm_plan_create")
        return Plan(self, h, len(p), m, want_bwd)

    def calibrate(self, prof, seq_len=1000, border=100, trials=1000):
        """estimateK for each profile id in `prof` (dummest-calibration.hh:338-600).

        Returns (mid_scores, filter_scores, fits): the natural-log scores
        (nprof x trials) and per profile a dict of the mid-anchored and
        filter Gumbel fits (lambda, K, ... as estimateGumbel computes them)."""
        p = np.ascontiguousarray(np.atleast_1d(prof), np.int32)
        n = len(p)
        mid = np.zeros((n, trials), np.float64)
        flt = np.zeros((n, trials), np.float64)
        fits = np.zeros((n, 16), np.float64)
        dd = ctypes.c_double
        self._check(lib().dm_calibrate(self._h, n, _ptr(p, ctypes.c_int32), int(seq_len), int(border), int(trials),
                                       _ptr(mid, dd), _ptr(flt, dd), _ptr(fits, dd)), "dm_calibrate")
        out = [{"mid": dict(zip(GUMBEL_KEYS, f[:8].tolist())), "filter": dict(zip(GUMBEL_KEYS, f[8:].tolist()))}
               for f in fits]
        return mid, flt, out

    def calibrate_anchors(self, prof, seq_len=1000, border=100, trials=1000):
        """The per-trial table of estimateK's verbose output
        (dummest-calibration.hh:511-519) for the same trials as calibrate():
        (anchors, mid_scores): anchors (nprof x trials x 2) = profile position
        and sequence position of each trial's best mid-anchored score, and that
        score in nats (nprof x trials)."""
        p = np.ascontiguousarray(np.atleast_1d(prof), np.int32)
        n = len(p)
        anchors = np.zeros((n, trials, 2), np.int32)
        mid = np.zeros((n, trials), np.float64)
        self._check(lib().dm_calibrate_anchors(self._h, n, _ptr(p, ctypes.c_int32), int(seq_len), int(border),
                                               int(trials), _ptr(anchors, ctypes.c_int32),
                                               _ptr(mid, ctypes.c_double)), "dm_calibrate_anchors")
        return anchors, mid

    def autotune(self, seq_len=300):
        """Time every kernel configuration on the device; later plans pick the
        configuration with the least measured cost per pair. Returns the table."""
        ns = np.zeros(64, np.float64)
        n = lib().dm_autotune(self._h, int(seq_len), _ptr(ns, ctypes.c_double), 64)
        self._check(n if n < 0 else 0, "dm_autotune")
        return self.config_table()

    def set_config_costs(self, ns_per_step):
        """Load a cost table saved from autotune() (same device and build)."""
        ns = np.ascontiguousarray(ns_per_step, np.float64)
        self._check(lib().dm_set_config_costs(self._h, len(ns), _ptr(ns, ctypes.c_double)), "dm_set_config_costs")

    def config_table(self):
        """[(G, R, strips, ns_per_wave_step)] in configuration order."""
        cap = 64
        G, R, S = (np.zeros(cap, np.int32) for _ in range(3))
        ns = np.zeros(cap, np.float64)
        i32 = ctypes.c_int32
        n = lib().dm_config_table(self._h, cap, _ptr(G, i32), _ptr(R, i32), _ptr(S, i32), _ptr(ns, ctypes.c_double))
        return [(int(G[k]), int(R[k]), bool(S[k]), float(ns[k])) for k in range(min(n, cap))]

    def kernel_stats(self):
        """[(G, R, strips, ms, cells, launches)] per kernel configuration of the last execution."""
        cap = 64
        G = np.zeros(cap, np.int32)
        R = np.zeros(cap, np.int32)
        S = np.zeros(cap, np.int32)
        ms = np.zeros(cap, np.float64)
        cells = np.zeros(cap, np.int64)
        lc = np.zeros(cap, np.int64)
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        n = lib().dm_kernel_stats(self._h, cap, _ptr(G, i32), _ptr(R, i32), _ptr(S, i32), _ptr(ms, ctypes.c_double),
                                  _ptr(cells, i64), _ptr(lc, i64))
        return [(int(G[k]), int(R[k]), bool(S[k]), float(ms[k]), int(cells[k]), int(lc[k]))
                for k in range(min(n, cap))]

    def copy_probe(self, nbytes: int) -> float:
        """Device ms of one plain copy of nbytes (traffic-counter reference)."""
        ms = ctypes.c_double()
        self._check(lib().dm_copy_probe(self._h, int(nbytes), ctypes.byref(ms)), "dm_copy_probe")
        return ms.value

    @property
    def last_dp_ms(self) -> float:
        return lib().dm_last_dp_ms(self._h)

    @property
    def last_tantan_ms(self) -> float:
        return lib().dm_last_tantan_ms(self._h)

    @property
    def last_cells(self) -> int:
        return lib().dm_last_cells(self._h)


class Plan:
    """A device-resident work queue (dm_plan); execute() runs it."""

    def __init__(self, engine, handle, npairs, mode, want_bwd):
        self.engine, self._h, self.npairs, self.mode, self.want_bwd = engine, handle, npairs, mode, want_bwd

    def execute(self):
        n = self.npairs
        f, i32 = ctypes.c_float, ctypes.c_int32
        e = self.engine
        if self.mode == 0:
            best = np.zeros(n, np.float32)
            e._check(lib().dm_plan_execute(e._h, self._h, _ptr(best, f), None, None, None, None), "dm_plan_execute")
            return best
        mm = np.zeros(n, np.float32)
        me = np.zeros(n, np.int32)
        bm = np.zeros(n, np.float32)
        be = np.zeros(n, np.int32)
        e._check(lib().dm_plan_execute(e._h, self._h, None, _ptr(mm, f), _ptr(me, i32), _ptr(bm, f), _ptr(be, i32)),
                 "dm_plan_execute")
        return mm, me, bm, be

    def capture(self):
        """Record the plan's launches into a hipGraph (single-chunk plans)."""
        self.engine._check(lib().dm_plan_capture(self.engine._h, self._h), "dm_plan_capture")

    def close(self):
        if self._h:
            lib().dm_plan_destroy(self.engine._h, self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
