"""The engine executable with the calibration cache (estimateK's cache path,
dummest-calibration.hh:338-360 and :584-660; --warmup, dummest.cc:400-425):
a second run finds every profile in the cache and prints the same statistics
and records without calibrating; -S always calibrates; entries written by
another program under the reference's key are used as they are."""
import os
import re
import subprocess

import numpy as np
import pytest

import dummest_amd as D
from dummest_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dummest_amd", "dummest-amd")
LIB = os.path.join(ROOT, "dummest_amd", "libdummest_amd.so")
OPTS = ["-t", "40", "-l", "600", "-b", "20"]


def run(args, cache_dir, **env_extra):
    env = dict(os.environ, DUMMEST_CACHE_DIR=str(cache_dir), **env_extra)
    print("dummest-amd", *args, flush=True)
    r = subprocess.run([CLI, *OPTS, *args], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.fixture(scope="module")
def hmm(tmp_path_factory):
    # profiles A, B and A again (the third has A's key)
    rng = np.random.default_rng(21)
    a = synth.make_profile(rng, 70, "profA", ndesc=12)[0]
    b = synth.make_profile(rng, 45, "profB", ndesc=12)[0]
    p = tmp_path_factory.mktemp("calcache") / "abc.hmm"
    p.write_text(a + b + a)
    return p


def binary_hash():
    data = open(CLI, "rb").read() + open(LIB, "rb").read()
    return "%032x" % D.xxh3_128(data)


def test_second_run_uses_the_cache(hmm, tmp_path):
    out1 = run([str(hmm)], tmp_path)
    out2 = run([str(hmm)], tmp_path)
    # run 1: A and B calibrated, the repeated A found in the cache
    assert out1.count("# sum is") == 2 and out1.count("# Warning: using cached results") == 1
    # run 2: everything cached; otherwise the same text
    assert "# sum is" not in out2 and out2.count("# Warning: using cached results") == 3
    # (a calibrated profile also prints its trial table, which a cached one has not)
    assert re.sub(r"# sum is .*\n#trial.*\n#\tprofPos.*\n(?:\d+\t.*\n)*", "# Warning: using cached results\n",
                  out1) == out2
    lines = out1.splitlines()
    cf = '# Cache file: "%s"' % (tmp_path / "cache.bin")
    # printed once, by the first profile's lookup, after its background line
    assert lines.count(cf) == 1
    i = lines.index(cf)
    assert lines[i - 1].startswith("# Background letter probabilities:")
    assert lines.index("# Profile name: profA") < i < lines.index("# Profile name: profB")
    # the file holds the two distinct keys, computed as the reference does
    ents = dict(D.cache_read(str(tmp_path / "cache.bin")))
    bh = binary_hash()
    want = set()
    for k in range(2):
        M, values, params, bg, _ = D.parse_hmm_text(hmm.read_text(), index=k, mask_threshold=0.5)
        want.add(D.calibration_key(["profA", "profB"][k], M, values, 600, 20, 40, D.DEFAULT_RATES, 0.5,
                                   binary_hash=bh))
    assert set(ents) == want
    # the printed statistics are the stored doubles
    lam = [float(x) for x in re.findall(r"# Lambda: (\S+)", out1)]
    stored = sorted(ents.values(), key=lambda v: v[2])
    assert len(lam) == 6
    assert sorted({lam[0], lam[2]}) == pytest.approx(sorted(v[2] for v in stored), rel=1e-5)
    assert sorted({lam[1], lam[3]}) == pytest.approx(sorted(v[24] for v in stored), rel=1e-5)
    assert lam[0:2] == lam[4:6]


def test_warmup_counts(hmm, tmp_path):
    out1 = run(["--warmup", str(hmm)], tmp_path)
    assert re.findall(r"# Warmup: (\S+) (cache hit|calibrated)", out1) == [
        ("profA", "calibrated"), ("profB", "calibrated"), ("profA", "cache hit")]
    assert "# Warmup: 3 profiles, 1 cache hits, 2 calibrated" in out1
    out2 = run(["--warmup", str(hmm)], tmp_path)
    assert "# Warmup: 3 profiles, 3 cache hits, 0 calibrated" in out2


def test_scores_file_always_calibrates(hmm, tmp_path):
    run([str(hmm)], tmp_path)
    out = run(["-S", str(tmp_path / "s.tsv"), str(hmm)], tmp_path)
    assert out.count("# sum is") == 3 and "using cached results" not in out
    lines = out.splitlines()
    cf = '# Cache file: "%s"' % (tmp_path / "cache.bin")
    # with -S the first use of the cache is the store after calibrating, which
    # follows the first profile's trial table (40 trials after '# sum is' and
    # the two header lines)
    i = lines.index(cf)
    assert lines[i - 1].startswith("40\t") and lines[i - 43].startswith("# sum is")
    rows = open(tmp_path / "s.tsv").read().splitlines()
    assert len(rows) == 1 + 3 * 40 * 4


def test_entries_from_another_writer_are_used(hmm, tmp_path):
    # an entry under the key without the binary hash (as any build of the
    # reference writes it with DUMMEST_CACHE_IGNORE_BINARY_HASH=1)
    M, values, params, bg, _ = D.parse_hmm_text(hmm.read_text(), index=1, mask_threshold=0.5)
    key = D.calibration_key("profB", M, values, 600, 20, 40, D.DEFAULT_RATES, 0.5)
    v = np.arange(1, 28, dtype=np.float64) / 64
    D.cache_write(str(tmp_path / "cache.bin"), [(key, v)])
    out = run(["-e", "0", str(hmm)], tmp_path, DUMMEST_CACHE_IGNORE_BINARY_HASH="1")
    blocks = out.split("# Profile name: ")
    b = blocks[2]
    assert b.startswith("profB") and "using cached results" in b and "# sum is" not in b
    # verbosity 3: the full table of the entry, then the filter copy
    assert "#lamMM\t%g\t%g\t%g" % (v[0], v[1], v[2]) in b
    assert "#kLM\t%g\t%g\t%g" % (v[21], v[22], v[23]) in b
    assert "#lamMM\t%g\t%g\t%g" % (v[24], v[24], v[24]) in b
    assert re.findall(r"# Lambda: (\S+)", b) == ["%g" % v[2], "%g" % v[24]]
    ents = dict(D.cache_read(str(tmp_path / "cache.bin")))
    assert np.array_equal(ents[key], v) and len(ents) == 2


def test_search_records_same_with_cached_calibration(tmp_path):
    rng = np.random.default_rng(5)
    text, anc, descs = synth.make_profile(rng, 80, "srch", ndesc=15)
    reads = []
    for k in range(6):
        host = synth.random_dna(rng, 900)
        if k % 2 == 0:
            host = synth.embed(rng, host, synth.reverse_translate(rng, descs[k]))
        reads.append(">r%d\n%s\n" % (k, host))
    (tmp_path / "p.hmm").write_text(text)
    (tmp_path / "s.fa").write_text("".join(reads))
    args = ["-e", "1", str(tmp_path / "p.hmm"), str(tmp_path / "s.fa")]
    out1 = run(args, tmp_path)
    out2 = run(args, tmp_path)
    recs1 = [ln for ln in out1.splitlines() if ln and not ln.startswith("#")]
    recs2 = [ln for ln in out2.splitlines() if ln and not ln.startswith("#")]
    assert recs1, "the case should produce hits"
    assert recs1 == recs2
    assert "# sum is" in out1 and "# sum is" not in out2


def test_trial_table(hmm, tmp_path):
    """Without sequences the engine prints each trial's best anchor and
    mid-anchored score after '# sum is' (estimateK's verbose trial loop,
    dummest-calibration.hh:392-396, 511-519), the end- and start-anchored
    columns repeating the anchor with -inf: every line against the oracle's
    report for the same trial of the same finalized profile."""
    import oracle as O
    out = run([str(hmm)], tmp_path)
    lines = out.splitlines()
    for k, name in enumerate(["profA", "profB"]):
        M, values, params, bg, _ = D.parse_hmm_text(hmm.read_text(), index=k, mask_threshold=0.5)
        prof = O.Profile(name, M, values, params, bg, [])
        i = lines.index("# Profile name: " + name)
        j = next(x for x in range(i, len(lines)) if lines[x].startswith("# sum is"))
        assert lines[j + 1] == "#trial\tend-anchored\t\tstart-anchored\t\tmid-anchored"
        assert lines[j + 2] == "#\tprofPos\tseqPos\tscore\tprofPos\tseqPos\tscore\tprofPos\tseqPos\tscore"
        for t in range(40):
            f = lines[j + 3 + t].split("\t")
            codes = O.decode(O.calib_sequence(bg, 600, 20, t))
            ln, m, e, a1, a2, _ = O.search(prof, codes, -2.0, need_align=False).hits[2]
            assert f[0] == str(t + 1) and f[3] == "-inf" and f[6] == "-inf", f
            assert f[1:3] == f[4:6] == f[7:9] == [str(a1), str(a2)], (name, t, f)
            assert float(f[9]) == pytest.approx(ln / np.log(2), abs=2e-6), (name, t, f)
        assert lines[j + 43].startswith("# Cache file") or lines[j + 43].startswith("#\tend-")
