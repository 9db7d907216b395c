"""Streaming around the engine (dummest.cc:473-606): the executable driven the
way bin/dummest's --max mode drives the reference engine (bin/dummest:484-497:
`hmm - -T n --max -e E -N total` with the contig x profile x strand stream of
_write_max_fasta, bin/dummest:236-276, on stdin). Chunking (--batch) must not
change a record; the pipeline header's profile= field decides which profile a
record is searched with; window coordinates map hits back into the contig."""
import os
import re
import subprocess

import numpy as np
import pytest

from dummest_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dummest_amd", "dummest-amd")
CAL = ["-t", "60", "-l", "800", "-b", "20"]
COMP = str.maketrans("ACGTacgt", "TGCAtgca")


def max_stream(contigs, profile_names):
    """The byte stream bin/dummest writes in --max mode (its _write_max_fasta)."""
    out = []
    for chrom, seq in contigs.items():
        L = len(seq)
        rc = seq.translate(COMP)[::-1]
        for prof in sorted(profile_names):
            head = f">{chrom}/1-{L} length={L} profile={prof} "
            out += [head + "plus_strand\n", seq + "\n", head + "minus_strand_revcomp\n", rc + "\n"]
    return "".join(out)


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    rng = np.random.default_rng(44)
    ta, _, da = synth.make_profile(rng, 90, "protA", ndesc=15)
    tb, _, db = synth.make_profile(rng, 60, "protB", ndesc=15)
    d = tmp_path_factory.mktemp("stream")
    (d / "ab.hmm").write_text(ta + tb)
    c1 = synth.embed(rng, synth.random_dna(rng, 3000), synth.reverse_translate(rng, da[0]))
    hb = synth.reverse_translate(rng, db[1])
    c2 = synth.embed(rng, synth.random_dna(rng, 2500), hb.translate(COMP)[::-1])  # on the minus strand
    c3 = synth.random_dna(rng, 1200)
    contigs = {"chr1": c1, "chr2": c2, "chr3": c3}
    return d / "ab.hmm", contigs


def run(args, stdin):
    print("dummest-amd", *args, flush=True)
    r = subprocess.run([CLI, *CAL, *args], input=stdin, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    return r.stdout


def records(out):
    return [ln for ln in out.splitlines() if ln and not ln.startswith("#")]


def test_max_mode_stream_chunking_invariant(case):
    hmm, contigs = case
    stream = max_stream(contigs, ["protA", "protB"])
    tot = 2 * sum(len(s) for s in contigs.values())
    base = ["-T", "1", "--max", "-e", "10", "-N", str(tot)]
    whole = run([str(hmm), "-", *base], stream)
    assert f"# Total sequence length: {tot}" in whole
    recs = records(whole)
    assert recs, "the case should produce hits"
    for batch in ["1", "2", "5"]:
        assert records(run([str(hmm), "-", *base, "--batch", batch], stream)) == recs, batch
    # the same stream from a file
    f = hmm.parent / "stream.fa"
    f.write_text(stream)
    assert records(run([str(hmm), str(f), *base], "")) == recs


def test_profile_field_routes_and_coordinates_map_back(case):
    hmm, contigs = case
    stream = max_stream(contigs, ["protA", "protB"])
    tot = 2 * sum(len(s) for s in contigs.values())
    out = run([str(hmm), "-", "-T", "1", "--max", "-e", "1e-3", "-N", str(tot)], stream)
    recs = records(out)
    # "a score=..." then the profile "s" line and the sequence "s" line
    blocks = [recs[i:i + 3] for i in range(0, len(recs), 3)]
    assert blocks
    seen = set()
    for a, sp, sq in blocks:
        assert a.startswith("a ")
        _, pname, pstart, plen, _, _, _ = sp.split(None, 6)
        f = sq.split()
        name, start, alen, strand, total = f[1], int(f[2]), int(f[3]), f[4], int(f[5])
        chrom = name.split("/")[0] if "/" in name else name
        assert chrom in contigs and total == len(contigs[chrom])
        assert 0 <= start and start + alen <= total
        seen.add((pname, chrom, strand))
        # the aligned DNA is the contig's (or its reverse complement's) letters
        letters = re.sub(r"[^A-Za-z]", "", f[6]).upper()
        src = contigs[chrom] if strand == "+" else contigs[chrom].translate(COMP)[::-1]
        assert src[start:start + alen] == letters
    # protA's homolog is on chr1's plus strand, protB's on chr2's minus strand
    assert ("protA", "chr1", "+") in seen and ("protB", "chr2", "-") in seen
    assert not any(c == "chr3" for _, c, _ in seen)


def test_profile_field_selects_the_profile(case):
    # chr1 holds a protA homolog: streamed with profile=protB only, it has no hits
    hmm, contigs = case
    one = {"chr1": contigs["chr1"]}
    tot = 2 * len(contigs["chr1"])
    args = [str(hmm), "-", "-T", "1", "--max", "-e", "1e-3", "-N", str(tot)]
    with_a = records(run(args, max_stream(one, ["protA"])))
    with_b = records(run(args, max_stream(one, ["protB"])))
    assert with_a and all(r.split()[1] == "protA" for r in with_a[1::3])
    assert with_b == []
