"""tantan masking on the GPU against the oracle, and the reference's own
fingerprints for the default -m 3 (tests/test_overflow.sh:66-71).

The device computes the repeat probabilities of DNA in FP64 with the
reference's summation order and fused multiply-adds, so they are compared bit
for bit; masks, decoded codes and profile transforms are exact. The fingerprint
test runs the engine executable as bin/dummest --max does (bin/dummest:479-485,
_write_max_fasta :237-277): both strands of the contig, pipeline headers, -N =
twice the contig length, default calibration (1000 trials of 5000 + 100
letters) and default masking; it checks the top hit's score within the
reference's tolerance, its E-value against the reference's pattern, and its
spans exactly.
"""
import gzip
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O
import dummest_amd as D
from tests.common import GOLDEN, read_fasta, synthetic_case
from tests.test_tantan import repeat_profile_text

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dummest_amd", "dummest-amd")


def dna_cases(seed):
    rng = np.random.default_rng(seed)
    lengths = [0, 1, 2, 3, 4, 5, 15, 16, 17, 31, 32, 33, 99, 100, 101, 102, 103, 104, 105, 117, 250, 1000, 3001]
    out = []
    for k, n in enumerate(lengths * 2):
        s = "".join("ACGT"[x] for x in rng.integers(0, 4, n))
        if k % 3 and n > 20:
            per = int(rng.integers(1, 40))
            unit = "".join("ACGT"[x] for x in rng.integers(0, 4, per))
            a = int(rng.integers(0, n // 2))
            b = min(n, a + int(rng.integers(10, n)))
            block = (unit * (n // per + 2))[:b - a]
            s = s[:a] + block + s[b:]
        if k % 5 == 0 and n > 10:
            s = s[:5] + "nRyN" + s[9:].lower()  # other letters count as A, either case
        out.append(s)
    return out


def test_tantan_dna_bit_exact(engine):
    seqs = dna_cases(3)
    print(f"tantan dna: {len(seqs)} sequences", flush=True)
    got = engine.tantan_dna(seqs)
    masked = 0
    for s, g in zip(seqs, got):
        _, ref = O.mask_dna(s, 0.5)
        np.testing.assert_array_equal(g.view(np.uint32), ref.view(np.uint32))
        masked += int(np.sum(ref >= 0.5))
    assert masked > 500


def test_add_dna_masks_like_reference(engine):
    """dm_add_dna under mask bit 2 = decode(makeMaskedSequence(dna)): the DP
    scores of the two sequence sets are identical."""
    text, prof, dnas = synthetic_case(808, 90, [400, 700, 1200, 300], homolog_every=1)
    for k in range(len(dnas)):
        dnas[k] = dnas[k][:150] + "GAT" * 25 + dnas[k][225:]
    with tempfile.NamedTemporaryFile("w", suffix=".hmm", delete=False) as f:
        f.write(text)
        path = f.name
    try:
        pid = engine.read_hmm(path)[0]
    finally:
        os.unlink(path)
    engine.set_masking(2, 0.5)
    try:
        masked_ids = engine.add_dna(dnas)
    finally:
        engine.set_masking(0, 0.5)
    ref_ids = engine.add_sequences([O.decode(O.mask_dna(d, 0.5)[0]) for d in dnas])
    plain_ids = engine.add_dna(dnas)
    a = engine.fwdbwd_scores(pid, masked_ids, want_bwd=True)
    b = engine.fwdbwd_scores(pid, ref_ids, want_bwd=True)
    c = engine.fwdbwd_scores(pid, plain_ids, want_bwd=True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert any(not np.array_equal(x, y) for x, y in zip(a, c))  # the masking changed something


def test_profile_masking_in_engine(engine):
    text = repeat_profile_text(31)
    with tempfile.NamedTemporaryFile("w", suffix=".hmm", delete=False) as f:
        f.write(text)
        path = f.name
    engine.set_masking(1, 0.5)
    try:
        pid = engine.read_hmm(path)[0]
    finally:
        engine.set_masking(0, 0.5)
        os.unlink(path)
    ref = O.finalize(O.parse_hmm(text)[0], np.float32, mask_threshold=0.5)
    M, v, p, b = engine.get_profile(pid)
    np.testing.assert_array_equal(v, ref.values)
    np.testing.assert_array_equal(p, ref.params)
    np.testing.assert_array_equal(b, ref.bg)
    assert engine.profile_masked(pid) == ref.masked and ref.masked > 0


def revcomp(s):
    return s.translate(str.maketrans("ACGTacgt", "TGCAtgca"))[::-1]


# name | fasta | expected bits | tolerance | E pattern | spans (tests/test_overflow.sh:66-71, FP32 cases)
FINGERPRINTS = [
    ("medium", "medium.fa", 754.34, 1.0, r"[0-9]\.[0-9]+e-216", "TRAIN.Herpes_MCP,428,504,+;medium0,1994,1510,+"),
    ("herpes", "herpes.fa", 2242.35, 6.0, r"0", "TRAIN.Herpes_MCP,1,1358,+;herpes0,3003,4072,+"),
]


@pytest.mark.parametrize("case", FINGERPRINTS, ids=[c[0] for c in FINGERPRINTS])
def test_reference_fingerprints(tmp_path, case):
    name, fa, score, tol, epat, spans = case
    print(f"fingerprint {name}", flush=True)
    hmm = tmp_path / "herpes.hmm"
    hmm.write_bytes(gzip.open(os.path.join(GOLDEN, "herpes.hmm.gz")).read())
    chrom, seq = read_fasta(os.path.join(GOLDEN, fa))[0]
    L = len(seq)
    prof = "TRAIN.Herpes_MCP"
    head = f">{chrom}/1-{L} length={L} profile={prof} "
    fasta = f"{head}plus_strand\n{seq}\n{head}minus_strand_revcomp\n{revcomp(seq)}\n"
    r = subprocess.run([CLI, str(hmm), "-", "-T", "1", "--max", "-e", "10", "-N", str(2 * L)], input=fasta,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    assert "# Masking simple regions in profiles" in r.stdout and "# Masking simple regions in sequences" in r.stdout
    recs = []
    lines = r.stdout.splitlines()
    for k, line in enumerate(lines):
        m = re.match(r"a score=(\S+) E=(\S+) anchor=(\S+)", line)
        if m:
            s = [tuple(lines[k + j].split()[1:5]) for j in (1, 2)]
            recs.append((float(m.group(1)), m.group(2), m.group(3), s))
    assert recs, r.stdout[-2000:]
    top = max(recs, key=lambda x: x[0])
    print(f"{name}: score {top[0]} E {top[1]} anchor {top[2]} spans {top[3]}", flush=True)
    assert abs(top[0] - score) <= tol
    assert ";".join(",".join(t) for t in top[3]) == spans
    assert re.fullmatch(epat, top[1]), top[1]
