"""ORACLE -- test infrastructure only.

Python side of the CPU restatement of DUMMEST's search path (oracle.cpp).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.

* ``parse_hmm`` restates ``readProfiles`` (dummest-profile.hh:235-316): the
  HMMER3/f text state machine and the "-ln p" conversion of
  ``probFromText`` (dummest-profile.hh:14-19).
* everything else binds ``_build/liboracle.so`` (built by ``make`` here).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
WIDTH = 29
ALPHABET = "ACDEFGHIKLMNPQRSTVWYUO?*"
# dummest-core.hh:66-71 defaults: insert1 insert2 delete1 delete2 stop bg-stop
DEFAULT_OPTS = (0.0171, 0.0018, 0.0328, 0.0083, 0.0005, 0.046875)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _p(t):
    return ctypes.POINTER(t)


def _declare(L):
    f, d, i, u8, c = ctypes.c_float, ctypes.c_double, ctypes.c_int, ctypes.c_uint8, ctypes.c_char
    L.or_finalize_f32.argtypes = [i, _p(f), _p(f), _p(f), _p(d), i, _p(f), d, d, i]
    L.or_finalize_f64.argtypes = [i, _p(d), _p(d), _p(d), _p(d), i, _p(f), d, d, i]
    L.or_tantan.argtypes = [_p(u8), i, i, _p(f)]
    L.or_blosum62.argtypes = [_p(ctypes.c_int8)]
    L.or_mask_dna.argtypes = [ctypes.c_char_p, i, d, ctypes.c_char_p, _p(f)]
    L.or_decode.argtypes = [ctypes.c_char_p, i, _p(u8)]
    L.or_bw_estep.argtypes = [i, i, _p(d), i, _p(u8), _p(i), _p(d), _p(d), _p(d)]
    for suf, t in (("f32", f), ("f64", d)):
        getattr(L, "or_null_" + suf).argtypes = [_p(t), _p(u8), i, _p(d), _p(t), _p(t), _p(t),
                                                 _p(f), _p(i), _p(f), _p(i)]
        getattr(L, "or_filter_" + suf).argtypes = [i, _p(t), _p(t), _p(t), _p(d), _p(u8), i]
        getattr(L, "or_filter_" + suf).restype = d
        getattr(L, "or_search_" + suf).argtypes = [i, _p(t), _p(t), _p(t), _p(d), _p(u8), i, d, i, i, i, d,
                                                   _p(d), _p(i), i, i, _p(f), _p(f), _p(f)]
        getattr(L, "or_search_" + suf).restype = i
    for suf, t in (("f32", f), ("f64", d)):
        getattr(L, "or_totals_" + suf).argtypes = [i, _p(t), _p(t), _p(t), _p(d), _p(u8), i, i, d, _p(d)]
    L.or_matrices_f32.argtypes = [i, _p(f), _p(f), _p(f), _p(d), _p(u8), i, _p(f), _p(i), _p(f), _p(i), _p(f), _p(f)]
    L.or_filter_batch_f32.argtypes = [i, _p(f), _p(f), _p(f), _p(d), _p(u8), _p(ctypes.c_int64), _p(i), i,
                                      _p(f), i]
    L.or_fwdbwd_batch_f32.argtypes = [i, _p(f), _p(f), _p(f), _p(d), _p(u8), _p(ctypes.c_int64), _p(i), i,
                                      _p(f), _p(f), i]
    L.or_gumbel.argtypes = [_p(d), i, d, _p(d)]
    L.or_set_mutation.argtypes = [i]
    L.or_set_mutation.restype = None
    L.or_calib_sequence_f32.argtypes = [_p(f), _p(d), i, i, i, ctypes.c_char_p]
    for name in ("add", "badd", "mul", "bmul"):
        getattr(L, "or_es_" + name).argtypes = [f, i, f, i, _p(f), _p(i)]
    L.or_es_normalize.argtypes = [d, i, _p(f), _p(i)]
    L.or_es_scale.argtypes = [f, i, d, _p(f), _p(i)]
    L.or_es_bscale.argtypes = [f, i, d, _p(f), _p(i)]
    L.or_es_promote.argtypes = [f, i, _p(f), _p(i)]
    L.or_es_log.argtypes = [f, i]
    L.or_es_log.restype = d
    L.or_es_to_rowscale.argtypes = [f, i, i, _p(f)]
    L.or_es_from_rowscale.argtypes = [f, i, _p(f), _p(i)]


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _ct(dtype):
    return ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double


# ---------------------------------------------------------------------------
# HMMER3/f reader, dummest-profile.hh:235-316
# ---------------------------------------------------------------------------
@dataclass
class RawProfile:
    name: str
    length: int
    raw: np.ndarray            # (M+1, 29) float64, values as read
    consensus: list = field(default_factory=list)


def _prob(word: str) -> float:
    # probFromText, dummest-profile.hh:14-19
    if word.startswith("*"):
        return 0.0
    try:
        d = float(word)
    except ValueError:
        d = _strtod_prefix(word)
    return math.exp(-d)


def _strtod_prefix(word):
    best = 0.0
    for k in range(len(word), 0, -1):
        try:
            return float(word[:k])
        except ValueError:
            continue
    return best


def parse_hmm(text: str) -> list[RawProfile]:
    profiles = []
    state = 0
    name = None
    rows: list[list[float]] = []
    cons: list[int] = []
    length = 0
    for line in text.splitlines():
        words = line.split()
        word = words[0] if words else ""
        if state == 0:
            if word == "NAME":
                name = words[1] if len(words) > 1 else ""
            elif word == "HMM":
                state = 1
        elif state == 1:
            state = 2
        elif state == 2:
            if word != "COMPO":
                state = 3
        elif state == 3:
            if len(words) < 7:
                raise ValueError("bad transition line")
            mtoi = _prob(words[1])
            mtod = _prob(words[2])
            itoi = _prob(words[4])
            dtod = _prob(words[6])
            if mtoi > 1 or mtod > 1 or itoi > 1 or dtod > 1:
                raise ValueError("bad probability")
            rows.append([mtoi, itoi, mtod, dtod])
            state = 4
        elif state == 4:
            if word == "//":
                if length < 2:
                    raise ValueError("profile too short")
                rows[-1].extend([0.0] * (WIDTH - 4))
                raw = np.array(rows, dtype=np.float64)
                profiles.append(RawProfile(name, length, raw, cons))
                rows, cons, length, name = [], [], 0, None
                state = 0
            else:
                probs = []
                for w in words[1:]:
                    if "." not in w:
                        break
                    p = _prob(w)
                    if p > 1:
                        raise ValueError("bad probability")
                    probs.append(p)
                if len(probs) != 20:
                    raise ValueError("only 20-letter protein profiles are supported")
                rows[-1].extend(probs + [0.0] * 5)
                length += 1
                letter = probs.index(max(probs))
                cons.append(letter)
                state = 2
    if state != 0:
        raise ValueError("truncated profile file")
    return profiles


@dataclass
class Profile:
    name: str
    M: int
    values: np.ndarray   # (M+1, 29) Float, finalized
    params: np.ndarray   # (M+2, 11) Float
    bg: np.ndarray       # (29,) Float
    consensus: list
    masked: int = 0      # positions masked by tantan (profile masking)


def finalize(rp: RawProfile, dtype=np.float32, opts=DEFAULT_OPTS, bg_type="G", mask_threshold=None,
             filter_std_dev=0.0, keep_nonvarying=False) -> Profile:
    """finalizeProfile (dummest-profile.hh:105-233); mask_threshold turns on
    the profile masking of -m 1 (tantan on the consensus, :122-125),
    filter_std_dev the Gaussian letter filter of -d / -D (:64-103, 119-121),
    which takes masking's place."""
    dt = np.dtype(dtype)
    vals = np.ascontiguousarray(rp.raw.astype(dt))
    params = np.zeros((rp.length + 2, 11), dtype=dt)
    bg = np.zeros(WIDTH, dtype=dt)
    o = np.array(opts, dtype=np.float64)
    fn = lib().or_finalize_f32 if dt == np.float32 else lib().or_finalize_f64
    t = _ct(dt)
    rep = None
    if mask_threshold is not None and not filter_std_dev:
        rep = tantan(np.array(rp.consensus, np.uint8), True)
    rp_ptr = _ptr(rep, ctypes.c_float) if rep is not None else None
    ok = fn(rp.length, _ptr(vals, t), _ptr(params, t), _ptr(bg, t), _ptr(o, ctypes.c_double), ord(bg_type),
            rp_ptr, float(mask_threshold or 0.0), float(filter_std_dev), int(bool(keep_nonvarying)))
    if not ok:
        raise ValueError("finalizeProfile rejected the profile")
    prof = Profile(rp.name, rp.length, vals, params, bg, list(rp.consensus))
    prof.masked = 0 if rep is None else int(np.sum(rep >= np.float32(mask_threshold)))
    return prof


def blosum62() -> np.ndarray:
    """The oracle's BLOSUM62 table (tantan-wrapper.hh:16-36 order: ACDEFGHIKLMNPQRSTVWY)."""
    out = np.zeros((20, 20), np.int8)
    lib().or_blosum62(_ptr(out, ctypes.c_int8))
    return out


def tantan(seq: np.ndarray, is_protein: bool) -> np.ndarray:
    """Repeat probability per letter (calcTantanProbabilities, tantan-wrapper.hh:13-64)."""
    c = np.ascontiguousarray(seq, dtype=np.uint8)
    out = np.zeros(max(1, len(c)), np.float32)
    lib().or_tantan(_ptr(c, ctypes.c_uint8), len(c), int(is_protein), _ptr(out, ctypes.c_float))
    return out[:len(c)]


def mask_dna(dna: str | bytes, threshold=0.5):
    """makeMaskedSequence (dummest-sequence.hh:206-236) on DNA letters: the
    masked letters become 'N'. Returns (masked letters, probabilities)."""
    b = dna.encode() if isinstance(dna, str) else bytes(dna)
    out = ctypes.create_string_buffer(len(b) + 1)
    probs = np.zeros(max(1, len(b)), np.float32)
    lib().or_mask_dna(b, len(b), float(threshold), out, _ptr(probs, ctypes.c_float))
    return out.raw[:len(b)].decode(), probs[:len(b)]


REF_TANTAN = os.path.join(HERE, "_ref", "libtantan_ref.so")


def ref_tantan(seq: np.ndarray, is_protein: bool):
    """The reference's own tantan (oracle/_ref, built from its sources by
    `make ref`), or None when that library was not built."""
    if not os.path.exists(REF_TANTAN):
        return None
    L = ctypes.CDLL(REF_TANTAN)
    L.ref_tantan_probs.argtypes = [_p(ctypes.c_uint8), ctypes.c_int, ctypes.c_int, _p(ctypes.c_float)]
    c = np.ascontiguousarray(seq, dtype=np.uint8)
    out = np.zeros(max(1, len(c)), np.float32)
    L.ref_tantan_probs(_ptr(c, ctypes.c_uint8), len(c), int(is_protein), _ptr(out, ctypes.c_float))
    return out[:len(c)]


def bw_estep(L: int, A: int, probs, letters, lengths, weights):
    """One Baum-Welch E-step of dummest-build (dummest-build.cc:561-736): the
    expected counts ((L+1) x (7+A), delete starts 0) and each sequence's
    forward total."""
    p = np.ascontiguousarray(probs, np.float64)
    assert p.shape == (L + 1, 7 + A)
    lt = np.concatenate([np.ascontiguousarray(letters, np.uint8), np.zeros(1, np.uint8)])
    ln = np.ascontiguousarray(lengths, np.int32)
    w = np.ascontiguousarray(weights, np.float64)
    counts = np.zeros((L + 1, 7 + A), np.float64)
    totals = np.zeros(max(1, len(ln)), np.float64)
    lib().or_bw_estep(L, A, _ptr(p, ctypes.c_double), len(ln), _ptr(lt, ctypes.c_uint8), _ptr(ln, ctypes.c_int),
                      _ptr(w, ctypes.c_double), _ptr(counts, ctypes.c_double), _ptr(totals, ctypes.c_double))
    return counts, totals[:len(ln)]


def cache_key(name: str, M: int, values, seq_len: int, border: int, trials: int, opts=DEFAULT_OPTS,
              tantan_threshold=0.5, binary_hash: str = "") -> str:
    """computeCacheKey (dummest-calibration.hh:201-226): the XXH3-128 hash of
    the binary hash text (left out under DUMMEST_CACHE_IGNORE_BINARY_HASH),
    the profile name, width and length (int), the finalized values ((M+1) x
    width floats), the background letter frequencies (the end row's 20
    letters), sequence length, border and trial count (int), then the float
    options INSERT1 INSERT2 DELETE1 DELETE2, BACKGROUND_FRAMESHIFT_RATE (=
    INSERT1 + DELETE2) and _2 (= INSERT2 + DELETE1) summed in float
    (dummest-core.hh:118-124), STOP, BG_STOP, TANTAN_MASK_THRESHOLD; printed
    high 64 bits first (Hash128::to_string, :134-143). The hash is the
    xxhash package's XXH3 (libxxhash 0.8.2; the reference vendors xxHash 0.8
    as include/xxhash.h)."""
    import struct

    import xxhash
    v = np.ascontiguousarray(values, np.float32).reshape(M + 1, WIDTH)
    f32 = np.float32
    i1, i2, d1, d2 = (f32(x) for x in opts[:4])
    b = binary_hash.encode() + name.encode() + struct.pack("<ii", WIDTH, M) + v.tobytes()
    b += v[M, 4:24].tobytes() + struct.pack("<iii", seq_len, border, trials)
    b += np.array([i1, i2, d1, d2, i1 + d2, i2 + d1, opts[4], opts[5], tantan_threshold], np.float32).tobytes()
    return "%032x" % xxhash.xxh3_128(b).intdigest()


REF_CACHE = os.path.join(HERE, "_ref", "libcache_ref.so")


def ref_cache():
    """The reference's vendored XXH3 and cereal (oracle/_ref/libcache_ref.so,
    built by `make ref`), or None when that library was not built."""
    if not os.path.exists(REF_CACHE):
        return None
    L = ctypes.CDLL(REF_CACHE)
    L.ref_xxh3_128_pieces.argtypes = [ctypes.c_char_p, _p(ctypes.c_int64), ctypes.c_int, _p(ctypes.c_uint64)]
    L.ref_cache_write.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int,
                                  _p(ctypes.c_double)]
    L.ref_cache_read.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_char_p, ctypes.c_int,
                                 _p(ctypes.c_double)]
    L.ref_cache_read.restype = ctypes.c_int64
    return L


def ref_xxh3_pieces(data: bytes, sizes) -> int:
    """XXH3_128bits of data fed in pieces of the given sizes (the Hash128
    add() sequence), by the reference's vendored xxhash.h."""
    L = ref_cache()
    sz = np.ascontiguousarray(sizes, np.int64)
    assert int(sz.sum()) == len(data)
    out = (ctypes.c_uint64 * 2)()
    L.ref_xxh3_128_pieces(data, _ptr(sz, ctypes.c_int64), len(sz), out)
    return out[0] | (out[1] << 64)


def ref_cache_write(path: str, entries) -> None:
    """Writes {key: 27 doubles} through the reference's vendored cereal."""
    L = ref_cache()
    n, stride = len(entries), 256
    keys = ctypes.create_string_buffer(max(1, n) * stride)
    vals = np.zeros((max(1, n), 27), np.float64)
    for i, (k, v) in enumerate(entries):
        ctypes.memmove(ctypes.addressof(keys) + i * stride, k.encode(), len(k))
        vals[i] = v
    assert L.ref_cache_write(path.encode(), n, keys, stride, _ptr(vals, ctypes.c_double)) == 0


def ref_cache_read(path: str):
    """{key: 27 doubles} read through the reference's vendored cereal, or None
    when cereal throws."""
    L = ref_cache()
    n = L.ref_cache_read(path.encode(), 0, None, 1, None)
    if n < 0:
        return None
    stride = 256
    keys = ctypes.create_string_buffer(max(1, n) * stride)
    vals = np.zeros((max(1, n), 27), np.float64)
    L.ref_cache_read(path.encode(), n, keys, stride, _ptr(vals, ctypes.c_double))
    raw = keys.raw
    return {raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode(): vals[i].copy() for i in range(n)}


def set_mutation(m: int) -> None:
    """Plant a known mistake in the oracle's DP (0 = none), for tests that
    show their check fails on it: 1 backward 1-nt/2-nt insertion scales
    swapped, 2 forward match emission read one column late, 3 filter match
    step scaled as two letters, 4 backward codon-insertion background read one
    column late."""
    lib().or_set_mutation(int(m))


def decode(dna: str | bytes) -> np.ndarray:
    b = dna.encode() if isinstance(dna, str) else dna
    out = np.zeros(len(b), dtype=np.uint8)
    lib().or_decode(b, len(b), _ptr(out, ctypes.c_uint8))
    return out


def null_model(bg: np.ndarray, codes: np.ndarray, opts=DEFAULT_OPTS):
    dt = bg.dtype
    t = _ct(dt)
    n = len(codes)
    ps = np.zeros(n, dt)
    ss = np.zeros(n, dt)
    sc = np.zeros(3, dt)
    pm = np.zeros(n + 4, np.float32)
    pe = np.zeros(n + 4, np.int32)
    sm = np.zeros(n + 4, np.float32)
    se = np.zeros(n + 4, np.int32)
    o = np.array(opts, np.float64)
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    fn = lib().or_null_f32 if dt == np.float32 else lib().or_null_f64
    fn(_ptr(bg, t), _ptr(c, ctypes.c_uint8), n, _ptr(o, ctypes.c_double), _ptr(ps, t), _ptr(ss, t), _ptr(sc, t),
       _ptr(pm, ctypes.c_float), _ptr(pe, ctypes.c_int), _ptr(sm, ctypes.c_float), _ptr(se, ctypes.c_int))
    return dict(prefix_seed=ps, suffix_seed=ss, scales=sc, prefix=(pm, pe), suffix=(sm, se))


def filter_score(p: Profile, codes: np.ndarray, opts=DEFAULT_OPTS) -> float:
    dt = p.values.dtype
    t = _ct(dt)
    o = np.array(opts, np.float64)
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    fn = lib().or_filter_f32 if dt == np.float32 else lib().or_filter_f64
    return fn(p.M, _ptr(p.values, t), _ptr(p.params, t), _ptr(p.bg, t), _ptr(o, ctypes.c_double),
              _ptr(c, ctypes.c_uint8), len(c))


@dataclass
class SearchResult:
    hits: list          # (lnratio, m, e, anchor1, anchor2, segments[(s1,s2,len)])
    global_best: tuple  # (m, e)
    bwd_best: tuple     # (m, e)
    col_best: np.ndarray  # (n, 3): m, e, row


def es_log(m, e):
    return -math.inf if m == 0 else math.log(m) + e * 0.6931471805599453


def search(p: Profile, codes: np.ndarray, min_prob_ratio: float, need_align=True, want_bwd=True,
           opts=DEFAULT_OPTS, rescale_bits=64, thresh=None, hit_cap=4096, seg_cap=200000) -> SearchResult:
    dt = p.values.dtype
    t = _ct(dt)
    if thresh is None:
        thresh = 1e30 if dt == np.float32 else 1e150
    o = np.array(opts, np.float64)
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    n = len(c)
    hits = np.zeros((hit_cap, 6), np.float64)
    segs = np.zeros((seg_cap, 3), np.int32)
    gb = np.zeros(2, np.float32)
    bb = np.zeros(2, np.float32)
    cb = np.zeros((n, 3), np.float32)
    fn = lib().or_search_f32 if dt == np.float32 else lib().or_search_f64
    nh = fn(p.M, _ptr(p.values, t), _ptr(p.params, t), _ptr(p.bg, t), _ptr(o, ctypes.c_double),
            _ptr(c, ctypes.c_uint8), n, float(min_prob_ratio), int(need_align), int(want_bwd), rescale_bits,
            float(thresh), _ptr(hits, ctypes.c_double), _ptr(segs, ctypes.c_int), hit_cap, seg_cap,
            _ptr(gb, ctypes.c_float), _ptr(bb, ctypes.c_float), _ptr(cb, ctypes.c_float))
    if nh < 0:
        raise RuntimeError("oracle output capacity exceeded")
    out = []
    k = 0
    for h in hits[:nh]:
        ns = int(h[5])
        out.append((h[0], float(h[1]), int(h[2]), int(h[3]), int(h[4]),
                    [tuple(int(x) for x in s) for s in segs[k:k + ns]]))
        k += ns
    return SearchResult(out, (float(gb[0]), int(gb[1])), (float(bb[0]), int(bb[1])), cb)


def totals(p: Profile, codes: np.ndarray, opts=DEFAULT_OPTS, rescale_bits=64, thresh=None):
    """(ln backward total, ln forward total) of all local paths."""
    dt = p.values.dtype
    t = _ct(dt)
    if thresh is None:
        thresh = 1e30 if dt == np.float32 else 1e150
    o = np.array(opts, np.float64)
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    out = np.zeros(2, np.float64)
    fn = lib().or_totals_f32 if dt == np.float32 else lib().or_totals_f64
    fn(p.M, _ptr(p.values, t), _ptr(p.params, t), _ptr(p.bg, t), _ptr(o, ctypes.c_double), _ptr(c, ctypes.c_uint8),
       len(c), rescale_bits, float(thresh), _ptr(out, ctypes.c_double))
    return float(out[0]), float(out[1])


def matrices(p: Profile, codes: np.ndarray, opts=DEFAULT_OPTS):
    """w0 and W1 of one pair as ExpScore matrices ((M+1) x n each), prefixSeed, scales."""
    assert p.values.dtype == np.float32
    n = len(codes)
    shape = (p.M + 1, n)
    w0m, W1m = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
    w0e, W1e = np.zeros(shape, np.int32), np.zeros(shape, np.int32)
    ps = np.zeros(n, np.float32)
    sc = np.zeros(3, np.float32)
    o = np.array(opts, np.float64)
    c = np.ascontiguousarray(codes, dtype=np.uint8)
    f, i = ctypes.c_float, ctypes.c_int
    lib().or_matrices_f32(p.M, _ptr(p.values, f), _ptr(p.params, f), _ptr(p.bg, f), _ptr(o, ctypes.c_double),
                          _ptr(c, ctypes.c_uint8), n, _ptr(w0m, f), _ptr(w0e, i), _ptr(W1m, f), _ptr(W1e, i),
                          _ptr(ps, f), _ptr(sc, f))
    return (w0m, w0e), (W1m, W1e), ps, sc


def filter_batch(p: Profile, seqs: list, opts=DEFAULT_OPTS, threads=None) -> np.ndarray:
    codes, offs, lens = pack(seqs)
    out = np.zeros(len(seqs), np.float32)
    o = np.array(opts, np.float64)
    f = ctypes.c_float
    lib().or_filter_batch_f32(p.M, _ptr(p.values, f), _ptr(p.params, f), _ptr(p.bg, f), _ptr(o, ctypes.c_double),
                              _ptr(codes, ctypes.c_uint8), _ptr(offs, ctypes.c_int64), _ptr(lens, ctypes.c_int),
                              len(seqs), _ptr(out, f), threads or os.cpu_count())
    return out


def fwdbwd_batch(p: Profile, seqs: list, opts=DEFAULT_OPTS, threads=None):
    codes, offs, lens = pack(seqs)
    mid = np.zeros((len(seqs), 2), np.float32)
    bwd = np.zeros((len(seqs), 2), np.float32)
    o = np.array(opts, np.float64)
    f = ctypes.c_float
    lib().or_fwdbwd_batch_f32(p.M, _ptr(p.values, f), _ptr(p.params, f), _ptr(p.bg, f), _ptr(o, ctypes.c_double),
                              _ptr(codes, ctypes.c_uint8), _ptr(offs, ctypes.c_int64), _ptr(lens, ctypes.c_int),
                              len(seqs), _ptr(mid, f), _ptr(bwd, f), threads or os.cpu_count())
    return mid, bwd


def pack(seqs):
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(len(seqs), np.int64)
    if len(seqs):
        offs[1:] = np.cumsum(lens[:-1])
    codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if seqs else np.zeros(0, np.uint8)
    return np.ascontiguousarray(codes), offs, lens


def gumbel(scores, seq_len):
    s = np.array(scores, np.float64)
    out = np.zeros(8, np.float64)
    lib().or_gumbel(_ptr(s, ctypes.c_double), len(s), float(seq_len), _ptr(out, ctypes.c_double))
    keys = ("mmL", "mmK", "mmKsimple", "mlL", "mlK", "mlKsimple", "lmL", "lmK")
    return dict(zip(keys, out.tolist()))


def calib_sequence(bg: np.ndarray, length: int, border: int, trial: int, opts=DEFAULT_OPTS) -> bytes:
    """Random frameshifting codon DNA of one calibration trial (letters)."""
    freqs = np.ascontiguousarray(bg[4:24].astype(np.float32))
    o = np.array(opts, np.float64)
    buf = ctypes.create_string_buffer(length + border)
    lib().or_calib_sequence_f32(_ptr(freqs, ctypes.c_float), _ptr(o, ctypes.c_double), length, border, trial, buf)
    return buf.raw[:length + border]


# ExpScore primitive wrappers (contract tests)
def _es2(fn, *args):
    om, oe = ctypes.c_float(), ctypes.c_int()
    fn(*args, ctypes.byref(om), ctypes.byref(oe))
    return om.value, oe.value


def es_normalize(m, e):
    return _es2(lib().or_es_normalize, m, e)


def es_add(a, b):
    return _es2(lib().or_es_add, a[0], a[1], b[0], b[1])


def es_batch_add(a, b):
    return _es2(lib().or_es_badd, a[0], a[1], b[0], b[1])


def es_mul(a, b):
    return _es2(lib().or_es_mul, a[0], a[1], b[0], b[1])


def es_batch_mul(a, b):
    return _es2(lib().or_es_bmul, a[0], a[1], b[0], b[1])


def es_scale(a, f):
    return _es2(lib().or_es_scale, a[0], a[1], f)


def es_batch_scale(a, f):
    return _es2(lib().or_es_bscale, a[0], a[1], f)


def es_promote(v, cum):
    return _es2(lib().or_es_promote, v, cum)


# -- the lane-batched SIMD CPU baseline (simd_baseline.cpp; measurement only) --
_simd = {}
_SIMD_FLAGS_V4 = ("avx512f", "avx512dq", "avx512bw", "avx512vl")


def cpu_flags() -> set:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def simd_variant(prefer: str | None = None) -> str:
    """'v4' (AVX-512, 16 lanes) where the host has it, else 'v3' (AVX2, 8 lanes)."""
    if prefer:
        return prefer
    return "v4" if all(f in cpu_flags() for f in _SIMD_FLAGS_V4) else "v3"


def simd_lib(variant: str | None = None):
    v = simd_variant(variant)
    if v not in _simd:
        path = os.path.join(HERE, "_build", f"libsimd_{v}.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        i, f, d = ctypes.c_int, ctypes.c_float, ctypes.c_double
        L.sb_lanes.restype = i
        L.sb_fwdbwd_pairs.argtypes = [i, _p(i), _p(ctypes.c_int64), _p(ctypes.c_int64), _p(f), _p(f), _p(f), _p(d),
                                      _p(ctypes.c_uint8), _p(ctypes.c_int64), _p(i), i, _p(i), _p(i), _p(f), _p(f), i, i]
        L.sb_fwdbwd_pairs.restype = d
        _simd[v] = L
    return _simd[v]


def simd_fwdbwd_pairs(profiles: list, seqs: list, pair_prof, pair_seq, opts=DEFAULT_OPTS, threads=None,
                      variant: str | None = None, want_bwd=True):
    """Forward+backward scores of (profile, sequence) pairs by the SIMD
    baseline: (mid (npairs, 2), bwd (npairs, 2), seconds of batched work,
    lanes). profiles: Profile objects; seqs: code arrays; pair_prof /
    pair_seq: indices into them."""
    L = simd_lib(variant)
    Ms = np.array([p.M for p in profiles], np.int32)
    vals = [np.ascontiguousarray(p.values, np.float32).ravel() for p in profiles]
    pars = [np.ascontiguousarray(p.params, np.float32).ravel() for p in profiles]
    val_off = np.zeros(len(profiles), np.int64)
    par_off = np.zeros(len(profiles), np.int64)
    if len(profiles) > 1:
        val_off[1:] = np.cumsum([v.size for v in vals[:-1]])
        par_off[1:] = np.cumsum([q.size for q in pars[:-1]])
    values = np.concatenate(vals) if vals else np.zeros(1, np.float32)
    params = np.concatenate(pars) if pars else np.zeros(1, np.float32)
    bgs = np.ascontiguousarray(np.concatenate([np.asarray(p.bg, np.float32)[:29] for p in profiles]))
    codes, offs, lens = pack(seqs)
    pp = np.ascontiguousarray(pair_prof, np.int32)
    ps = np.ascontiguousarray(pair_seq, np.int32)
    mid = np.zeros((len(pp), 2), np.float32)
    bwd = np.zeros((len(pp), 2), np.float32)
    o = np.array(opts, np.float64)
    f = ctypes.c_float
    i = ctypes.c_int
    sec = L.sb_fwdbwd_pairs(len(profiles), _ptr(Ms, i), _ptr(val_off, ctypes.c_int64), _ptr(par_off, ctypes.c_int64),
                            _ptr(values, f), _ptr(params, f), _ptr(bgs, f), _ptr(o, ctypes.c_double),
                            _ptr(codes if codes.size else np.zeros(1, np.uint8), ctypes.c_uint8),
                            _ptr(offs, ctypes.c_int64), _ptr(lens, i), len(pp), _ptr(pp, i), _ptr(ps, i),
                            _ptr(mid, f), _ptr(bwd, f), int(bool(want_bwd)), int(threads or os.cpu_count() or 1))
    return mid, bwd, sec, L.sb_lanes()
