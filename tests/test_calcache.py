"""Calibration cache shared with the reference engine (ProfileCache,
dummest-calibration.hh:106-336): the XXH3-128 hash, the key of a profile, the
cache file's layout and location, and the executable's load/save behaviour
(no GPU: calibration itself is in test_gpu_calcache.py).

Pinning: the hash against the xxhash package (libxxhash 0.8.2) and against
the reference's vendored xxhash.h (oracle/_ref, streaming the way Hash128 adds
its fields); the file layout against the reference's vendored cereal
(oracle/_ref); the key's field order against the oracle's restatement of
computeCacheKey plus a check that every field reaches the key.
"""
import os
import struct
import subprocess

import numpy as np
import pytest
import xxhash

import dummest_amd as D
import oracle as O
from dummest_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dummest_amd", "dummest-amd")
needs_ref = pytest.mark.skipif(O.ref_cache() is None, reason="oracle/_ref/libcache_ref.so not built (make ref)")

# every XXH3 code path: 0, 1-3, 4-8, 9-16, 17-128 (each 32-byte round count),
# 129-240, and the striped loop (< 1 block, block edges, several blocks)
LENGTHS = (list(range(0, 20)) + [31, 32, 33, 63, 64, 65, 96, 97, 127, 128, 129, 130, 159, 160, 161, 191, 192, 200,
                                  239, 240, 241, 255, 256, 257, 1023, 1024, 1025, 1088, 2048, 2049, 4095, 4160,
                                  65536, 65537, 262147])


def _bytes(rng, n):
    return rng.integers(0, 256, n, dtype=np.uint8).tobytes()


def test_xxh3_matches_xxhash_every_length_class():
    rng = np.random.default_rng(11)
    for n in LENGTHS:
        for _ in range(3):
            b = _bytes(rng, n)
            assert D.xxh3_128(b) == xxhash.xxh3_128(b).intdigest(), n


def test_xxh3_known_answers():
    # the empty input and a short text, from the xxhash package
    assert D.xxh3_128(b"") == xxhash.xxh3_128(b"").intdigest()
    assert D.xxh3_128(b"abc") == xxhash.xxh3_128(b"abc").intdigest()


@needs_ref
def test_xxh3_matches_vendored_streaming():
    # the reference's Hash128 feeds its fields in separate updates
    rng = np.random.default_rng(12)
    for n in [0, 5, 40, 200, 300, 1500, 5000, 70000]:
        b = _bytes(rng, n)
        cuts = np.sort(rng.integers(0, n + 1, rng.integers(0, 12)))
        sizes = np.diff(np.concatenate([[0], cuts, [n]]))
        assert O.ref_xxh3_pieces(b, sizes) == D.xxh3_128(b), n


def _profile(seed=3, L=60):
    rng = np.random.default_rng(seed)
    text = synth.make_profile(rng, L, "kp%d" % seed, ndesc=10)[0]
    M, values, params, bg, _ = D.parse_hmm_text(text)
    return M, values


def test_key_matches_oracle_composition():
    M, values = _profile()
    for bh in ["", "0123456789abcdef0123456789abcdef"]:
        for (L, b, t, thr) in [(5000, 100, 1000, 0.5), (777, 0, 3, 0.25)]:
            rates = (0.02, 0.003, 0.01, 0.004, 0.001, 0.05)
            got = D.calibration_key("kp3", M, values, L, b, t, rates, thr, binary_hash=bh)
            want = O.cache_key("kp3", M, values, L, b, t, rates, thr, binary_hash=bh)
            assert got == want
            assert len(got) == 32 and int(got, 16) >= 0


def test_key_depends_on_every_field():
    M, values = _profile()
    base = dict(name="kp3", M=M, values=values, seq_len=5000, border=100, trials=1000,
                rates=D.DEFAULT_RATES, tantan_threshold=0.5, binary_hash="ab" * 16)
    k0 = D.calibration_key(**base)
    changed = []
    v2 = values.copy()
    v2[M // 2, 10] = np.nextafter(v2[M // 2, 10], np.float32(1))
    v3 = values.copy()
    v3[M, 7] = np.nextafter(v3[M, 7], np.float32(1))  # a background letter frequency
    variants = [dict(name="kp4"), dict(values=v2), dict(values=v3), dict(seq_len=5001), dict(border=101),
                dict(trials=999), dict(tantan_threshold=0.4), dict(binary_hash="")]
    for i in range(6):
        r = list(D.DEFAULT_RATES)
        r[i] *= 1.5
        variants.append(dict(rates=tuple(r)))
    for v in variants:
        k = D.calibration_key(**{**base, **v})
        changed.append(k != k0)
    assert all(changed), changed
    # the rates enter as floats: a change below float resolution keeps the key
    r = list(D.DEFAULT_RATES)
    r[0] = float(np.float32(r[0])) * (1 + 1e-12)
    assert D.calibration_key(**{**base, "rates": tuple(r)}) == k0


def test_key_prefix_is_binary_hash_then_name():
    # without a binary hash the first hashed bytes are the name, then width 29
    M, values = _profile()
    k = D.calibration_key("n", M, values, 10, 0, 2, D.DEFAULT_RATES, 0.5)
    opts = np.array(D.DEFAULT_RATES, np.float64)
    f = [np.float32(x) for x in opts[:4]]
    tail = np.array(f + [f[0] + f[3], f[1] + f[2], opts[4], opts[5], 0.5], np.float32).tobytes()
    v = np.ascontiguousarray(values, np.float32)
    blob = b"n" + struct.pack("<ii", 29, M) + v.tobytes() + v[M, 4:24].tobytes() + struct.pack("<iii", 10, 0, 2) + tail
    assert k == "%032x" % xxhash.xxh3_128(blob).intdigest()


def _entries(rng, n):
    return [("%032x" % int(rng.integers(0, 2**62)), rng.normal(size=27)) for _ in range(n)]


def test_cache_file_layout(tmp_path):
    rng = np.random.default_rng(5)
    ents = _entries(rng, 3)
    p = str(tmp_path / "cache.bin")
    D.cache_write(p, ents)
    raw = open(p, "rb").read()
    # uint64 count, then per entry uint64 key length, the key, 27 doubles
    want = struct.pack("<Q", 3) + b"".join(struct.pack("<Q", len(k)) + k.encode() + np.asarray(v, "<f8").tobytes()
                                           for k, v in ents)
    assert raw == want
    back = D.cache_read(p)
    assert [k for k, _ in back] == [k for k, _ in ents]
    for (_, a), (_, b) in zip(back, ents):
        assert np.array_equal(a, b)


@needs_ref
def test_cache_file_interop_with_vendored_cereal(tmp_path):
    rng = np.random.default_rng(6)
    ents = _entries(rng, 5)
    ents[1][1][3] = np.nan
    ents[2][1][0] = -np.inf
    # ours -> cereal
    p1 = str(tmp_path / "a.bin")
    D.cache_write(p1, ents)
    got = O.ref_cache_read(p1)
    assert set(got) == {k for k, _ in ents}
    for k, v in ents:
        assert np.array_equal(got[k], v, equal_nan=True)
    # cereal -> ours
    p2 = str(tmp_path / "b.bin")
    O.ref_cache_write(p2, ents)
    back = D.cache_read(p2)
    assert {k for k, _ in back} == {k for k, _ in ents}
    d = dict(ents)
    for k, v in back:
        assert np.array_equal(v, d[k], equal_nan=True)
    # same entries in cereal's order: the same bytes
    p3 = str(tmp_path / "c.bin")
    D.cache_write(p3, back)
    assert open(p3, "rb").read() == open(p2, "rb").read()
    # an empty map
    O.ref_cache_write(str(tmp_path / "e.bin"), [])
    assert D.cache_read(str(tmp_path / "e.bin")) == []
    assert open(str(tmp_path / "e.bin"), "rb").read() == b"\0" * 8


def test_cache_file_missing_empty_and_truncated(tmp_path):
    assert D.cache_read(str(tmp_path / "none.bin")) == []
    (tmp_path / "empty.bin").write_bytes(b"")
    assert D.cache_read(str(tmp_path / "empty.bin")) == []
    rng = np.random.default_rng(7)
    p = str(tmp_path / "t.bin")
    D.cache_write(p, _entries(rng, 2))
    raw = open(p, "rb").read()
    for cut in [3, 8 + 4, 8 + 8 + 10, len(raw) - 1]:
        (tmp_path / "cut.bin").write_bytes(raw[:cut])
        with pytest.raises(D.DummestError):
            D.cache_read(str(tmp_path / "cut.bin"))
    # a key length beyond the file
    (tmp_path / "big.bin").write_bytes(struct.pack("<QQ", 1, 1 << 60) + b"x" * 40)
    with pytest.raises(D.DummestError):
        D.cache_read(str(tmp_path / "big.bin"))


def test_cache_path_search_order(tmp_path, monkeypatch):
    for v in ["DUMMEST_CACHE_DIR", "XDG_CACHE_HOME", "HOME", "TMPDIR"]:
        monkeypatch.delenv(v, raising=False)
    monkeypatch.chdir(tmp_path)
    d = tmp_path / "explicit"
    monkeypatch.setenv("DUMMEST_CACHE_DIR", str(d))
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path / "xdg"))
    monkeypatch.setenv("HOME", str(tmp_path / "home"))
    assert D.cache_file_path() == str(d / "cache.bin") and d.is_dir()
    # an unusable directory (a file is in the way) falls through to the next
    (tmp_path / "blocker").write_text("x")
    monkeypatch.setenv("DUMMEST_CACHE_DIR", str(tmp_path / "blocker" / "sub"))
    assert D.cache_file_path() == str(tmp_path / "xdg" / ".dummest-cache" / "cache.bin")
    monkeypatch.setenv("DUMMEST_CACHE_DIR", "")  # empty counts as unset
    monkeypatch.setenv("XDG_CACHE_HOME", "")
    assert D.cache_file_path() == str(tmp_path / "home" / ".dummest-cache" / "cache.bin")
    monkeypatch.delenv("HOME")
    assert D.cache_file_path() == str(tmp_path / ".dummest-cache" / "cache.bin")


def _run(args, cache_dir, env_extra=None):
    env = dict(os.environ, DUMMEST_CACHE_DIR=str(cache_dir), **(env_extra or {}))
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60, env=env)


def test_executable_saves_cache_at_exit(tmp_path):
    # the cache is written when the process ends, with the file name printed
    # (the reference's global ProfileCache saves from its destructor)
    r = _run(["-m", "1", str(tmp_path / "missing.hmm")], tmp_path / "c")
    assert r.returncode == 1
    assert r.stdout.strip().splitlines()[-1] == '# Cache file: "%s"' % (tmp_path / "c" / "cache.bin")
    assert (tmp_path / "c" / "cache.bin").read_bytes() == b"\0" * 8
    # existing entries survive a run that adds none
    rng = np.random.default_rng(8)
    ents = _entries(rng, 4)
    D.cache_write(str(tmp_path / "c" / "cache.bin"), ents)
    _run(["-m", "1", str(tmp_path / "missing.hmm")], tmp_path / "c")
    back = dict(D.cache_read(str(tmp_path / "c" / "cache.bin")))
    assert set(back) == {k for k, _ in ents}


def test_executable_drops_a_malformed_cache(tmp_path):
    (tmp_path / "c").mkdir()
    (tmp_path / "c" / "cache.bin").write_bytes(struct.pack("<Q", 5) + b"\1\2\3")
    r = _run([str(tmp_path / "missing.hmm")], tmp_path / "c")
    assert "# Cache Load Error" in r.stderr
    assert (tmp_path / "c" / "cache.bin").read_bytes() == b"\0" * 8
