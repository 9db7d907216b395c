"""Synthetic workloads shaped like the benchmark configurations.

Profiles: an ancestral protein (uniform over the 20 amino acids) and 50
descendants with 30% substitutions (uniform off-diagonal) and 2% indels of
geometric length (p = 0.5). Match emissions are the descendants' amino-acid
counts at each ancestral position with a +1 pseudocount per letter; the
transitions are the descendants' indel event frequencies with small
pseudocounts. The result is written as HMMER3/f text, the format the engine
reads (readProfiles, dummest-profile.hh:235-329). This stands in for
dummest-build's Baum-Welch fit (dummest-build.cc:748-834), which needs
Dirichlet-mixture priors; only the profile values differ, the DP work per
(profile, sequence) pair is the same.

DNA: i.i.d. bases at a chosen GC fraction; homologs are reverse-translated
descendants (uniform synonymous codons) carrying 1-nt insertions and deletions.
"""
from __future__ import annotations

import math

import numpy as np

AA = "ACDEFGHIKLMNPQRSTVWY"
CODONS = {
    "A": ["GCT", "GCC", "GCA", "GCG"], "R": ["CGT", "CGC", "CGA", "CGG", "AGA", "AGG"], "N": ["AAT", "AAC"],
    "D": ["GAT", "GAC"], "C": ["TGT", "TGC"], "Q": ["CAA", "CAG"], "E": ["GAA", "GAG"],
    "G": ["GGT", "GGC", "GGA", "GGG"], "H": ["CAT", "CAC"], "I": ["ATT", "ATC", "ATA"],
    "L": ["TTA", "TTG", "CTT", "CTC", "CTA", "CTG"], "K": ["AAA", "AAG"], "M": ["ATG"], "F": ["TTT", "TTC"],
    "P": ["CCT", "CCC", "CCA", "CCG"], "S": ["TCT", "TCC", "TCA", "TCG", "AGT", "AGC"],
    "T": ["ACT", "ACC", "ACA", "ACG"], "W": ["TGG"], "Y": ["TAT", "TAC"], "V": ["GTT", "GTC", "GTA", "GTG"],
}


def _nl(p):
    return "*" if p <= 0 else f"{-math.log(p):.5f}"


def evolve(rng, anc: np.ndarray, sub=0.30, indel=0.02):
    """One descendant: per-position substitutions, then indels of geometric length."""
    L = len(anc)
    d = anc.copy()
    s = rng.random(L) < sub
    off = rng.integers(1, 20, size=L)
    d[s] = (anc[s] + off[s]) % 20
    out = []
    aligned = []  # ancestral index per emitted residue (-1 = insertion)
    i = 0
    ins_ev = np.zeros(L + 1, np.int32)
    del_ev = np.zeros(L + 1, np.int32)
    while i < L:
        r = rng.random()
        if r < indel / 2:  # insertion after position i-1
            k = rng.geometric(0.5)
            out.extend(rng.integers(0, 20, size=k).tolist())
            aligned.extend([-1] * k)
            ins_ev[i] += 1
        elif r < indel:  # deletion starting at i
            k = rng.geometric(0.5)
            del_ev[i] += 1
            i += k
            continue
        out.append(int(d[i]))
        aligned.append(i)
        i += 1
    return np.array(out, np.int64), np.array(aligned, np.int64), ins_ev, del_ev


def make_profile(rng, L: int, name: str, ndesc=50, anc=None):
    """(HMMER3 text, ancestral protein as letter indices, descendants). anc
    (letter indices, length L) fixes the ancestor, e.g. a tandem repeat."""
    anc = rng.integers(0, 20, size=L) if anc is None else np.asarray(anc)
    counts = np.ones((L, 20))
    ins = np.zeros(L + 1)
    dels = np.zeros(L + 1)
    descs = []
    for _ in range(ndesc):
        seq, al, ie, de = evolve(rng, anc)
        descs.append(seq)
        m = al >= 0
        np.add.at(counts, (al[m], seq[m]), 1)
        ins += ie
        dels += de
    em = counts / counts.sum(1, keepdims=True)
    bg = np.full(20, 0.05)
    lines = ["HMMER3/f [synthetic]", f"NAME  {name}", f"LENG  {L}", "ALPH  amino",
             "HMM          " + "        ".join(AA),
             "            m->m     m->i     m->d     i->m     i->i     d->m     d->d"]
    lines.append("  COMPO   " + " ".join(_nl(x) for x in em.mean(0)))
    lines.append("          " + " ".join(_nl(x) for x in bg))

    def trans(node):
        mi = (ins[node] + 0.5) / (ndesc + 1.5)
        md = (dels[node] + 0.5) / (ndesc + 1.5) if node < L else 0.0
        mm = 1.0 - mi - md
        ii, dd = 0.5, (0.5 if node < L else 0.0)
        return "          " + " ".join(_nl(x) for x in (mm, mi, md, 1 - ii, ii, 1 - dd if node < L else 1.0, dd))

    lines.append(trans(0))
    for k in range(L):
        lines.append(f"{k + 1:7d}   " + " ".join(_nl(x) for x in em[k]) + f" {k + 1:6d} {AA[anc[k]].lower()} - - -")
        lines.append("          " + " ".join(_nl(x) for x in bg))
        lines.append(trans(k + 1))
    lines.append("//")
    return "\n".join(lines) + "\n", anc, descs


def simulate_msa(rng, L: int = 200, ndesc: int = 50, name: str = "sim", sub=0.30, indel=0.02):
    """Stockholm text of `ndesc` descendants of a random ancestor of L amino
    acids (uniform over the 20), each with `sub` uniform off-diagonal
    substitutions and `indel` indels of geometric length (p = 0.5), aligned to
    the ancestor: one column per ancestral position (a gap where it was
    deleted) and, before each, as many insertion columns as the longest
    insertion there (gaps in the descendants without one). The columns of
    ancestral positions are filled by most descendants and become the
    profile's positions when dummest-build keeps columns with a residue
    fraction of at least 0.5.
    Returns (text, ancestor as letter indices, descendants as letter indices)."""
    anc = rng.integers(0, 20, size=L)
    descs = [evolve(rng, anc, sub=sub, indel=indel) for _ in range(ndesc)]
    # insertion runs before ancestral position i (i = L: after the last)
    ins_before = np.zeros(L + 1, np.int64)
    per_desc = []
    for seq, al, _, _ in descs:
        runs = {}
        cur = []
        nxt = 0
        cols = {}
        for aa, a in zip(seq, al):
            if a < 0:
                cur.append(int(aa))
            else:
                if cur:
                    runs[int(a)] = cur
                    cur = []
                cols[int(a)] = int(aa)
                nxt = int(a) + 1
        if cur:
            runs[L if nxt >= L else nxt] = runs.get(L if nxt >= L else nxt, []) + cur
        for i, r in runs.items():
            ins_before[i] = max(ins_before[i], len(r))
        per_desc.append((runs, cols))
    rows = []
    for runs, cols in per_desc:
        out = []
        for i in range(L + 1):
            r = runs.get(i, [])
            out.append("".join(AA[x].lower() for x in r) + "." * (ins_before[i] - len(r)))
            if i < L:
                out.append(AA[cols[i]] if i in cols else "-")
        rows.append("".join(out))
    lines = ["# STOCKHOLM 1.0", "", f"#=GF ID {name}", ""]
    for k, r in enumerate(rows):
        lines.append(f"{name}_{k:03d} {r}")
    lines.append("//")
    return "\n".join(lines) + "\n", anc, [d[0] for d in descs]


def reverse_translate(rng, prot: np.ndarray) -> str:
    return "".join(CODONS[AA[a]][rng.integers(len(CODONS[AA[a]]))] for a in prot)


def random_dna(rng, n: int, gc=0.5) -> str:
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return "".join(np.array(list("ACGT"))[rng.choice(4, size=n, p=p)])


def mutate_dna(rng, dna: str, indel=0.01, sub=0.0) -> str:
    out = []
    for c in dna:
        r = rng.random()
        if r < indel / 2:
            continue  # 1-nt deletion
        if r < indel:
            out.append("ACGT"[rng.integers(4)])  # 1-nt insertion
        if sub and rng.random() < sub:
            c = "ACGT"[rng.integers(4)]
        out.append(c)
    return "".join(out)


def embed(rng, host: str, insert: str) -> str:
    if len(insert) >= len(host):
        return insert[:len(host)]
    at = int(rng.integers(0, len(host) - len(insert) + 1))
    return host[:at] + insert + host[at + len(insert):]


def lognormal_lengths(rng, n, median=150, sigma=0.8, lo=30, hi=2000):
    return np.clip(np.round(np.exp(np.log(median) + sigma * rng.standard_normal(n))), lo, hi).astype(int)


# ---------------------------------------------------------------------------
# Vectorized generators for the large workloads.
# ---------------------------------------------------------------------------
def raw_profile(rng, L: int, ndesc=50, sub=0.30, indel=0.02):
    """Reader-order probabilities ((L+1) x 29) of a profile built from `ndesc`
    simulated descendants of a random ancestor, and the ancestor.

    Same model as make_profile, vectorized over descendants: substitutions are
    uniform off-diagonal, deletions start at rate indel/2 per position with
    geometric lengths, insertions at rate indel/2 (they only count towards
    m->i)."""
    anc = rng.integers(0, 20, size=L)
    d = np.broadcast_to(anc, (ndesc, L)).copy()
    s = rng.random((ndesc, L)) < sub
    d[s] = (d[s] + rng.integers(1, 20, size=int(s.sum()))) % 20
    deleted = np.zeros((ndesc, L), bool)
    starts = np.argwhere(rng.random((ndesc, L)) < indel / 2)
    lens = rng.geometric(0.5, size=len(starts))
    del_ev = np.zeros(L + 1)
    for (r, c), k in zip(starts, lens):
        deleted[r, c:c + k] = True
        del_ev[c] += 1
    ins_ev = np.bincount(np.nonzero(rng.random((ndesc, L + 1)) < indel / 2)[1], minlength=L + 1).astype(float)
    counts = np.ones((L, 20))
    rows, cols = np.nonzero(~deleted)
    np.add.at(counts, (cols, d[rows, cols]), 1)
    em = counts / counts.sum(1, keepdims=True)
    raw = np.zeros((L + 1, 29), np.float64)
    mi = (ins_ev + 0.5) / (ndesc + 1.5)
    md = (del_ev + 0.5) / (ndesc + 1.5)
    md[L] = 0.0
    raw[:, 0] = mi       # m->i
    raw[:, 1] = 0.5      # i->i
    raw[:, 2] = md       # m->d
    raw[:, 3] = 0.5      # d->d
    raw[L, 3] = 0.0
    raw[:L, 4:24] = em   # row r holds node r+1's emissions
    return raw.astype(np.float32), anc


_CODON_TABLE = np.array([[ord(c) for c in CODONS[a][k % len(CODONS[a])]] for a in AA for k in range(6)],
                        np.uint8).reshape(20, 6, 3)
_NCOD = np.array([len(CODONS[a]) for a in AA])


def reverse_translate_np(rng, prot: np.ndarray) -> np.ndarray:
    """DNA bytes of a protein with uniform synonymous codons."""
    k = (rng.random(len(prot)) * _NCOD[prot]).astype(int)
    return _CODON_TABLE[prot, k].reshape(-1)


def random_dna_np(rng, n: int, gc=0.5) -> np.ndarray:
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return np.frombuffer(b"ACGT", np.uint8)[rng.choice(4, size=n, p=p)]


def mutate_np(rng, dna: np.ndarray, indel=0.01, sub=0.0) -> np.ndarray:
    """1-nt deletions and insertions at rate indel (half each) and substitutions."""
    bases = np.frombuffer(b"ACGT", np.uint8)
    r = rng.random(len(dna))
    keep = r >= indel / 2
    ins = (r >= indel / 2) & (r < indel)
    out = dna.copy()
    if sub:
        s = rng.random(len(dna)) < sub
        out[s] = bases[rng.integers(0, 4, size=int(s.sum()))]
    reps = keep.astype(int) + ins.astype(int)
    res = np.repeat(out, reps)
    # the inserted copy becomes a random base
    pos = np.cumsum(reps)[ins] - 1
    res[pos] = bases[rng.integers(0, 4, size=len(pos))]
    return res


def workload_c4(seed=1004, n_profiles=5000, n_reads=100000, pair_rate=0.004, max_len=2000, source_pairs=False):
    """Uneven reads against many profiles (the fwd+bwd-at-scale configuration).

    profiles: lengths lognormal (median 150, sigma 0.8, clipped to [30, 2000]
    and here also to `max_len`); reads: uniform 150-300 bp, half from a
    profile's ancestor (reverse-translated, then 1% 1-nt indels and 0.5%
    substitutions), half random (GC 0.5); pairs: each (read, profile) kept
    with probability `pair_rate`. source_pairs adds, after those, the pair of
    every homologous read with the profile it came from (reduced-size parity
    runs, where the random mask alone would pair few reads with their source).
    Returns dict(Ms, raws, reads (uint8 DNA arrays), prof, seq)."""
    rng = np.random.default_rng(seed)
    Ms = np.minimum(lognormal_lengths(rng, n_profiles), max_len)
    raws, ancs = [], []
    for M in Ms:
        r, a = raw_profile(rng, int(M))
        raws.append(r)
        ancs.append(a)
    lens = rng.integers(150, 301, size=n_reads)
    homolog = rng.random(n_reads) < 0.5
    src = rng.integers(0, n_profiles, size=n_reads)
    reads = []
    for k in range(n_reads):
        n = int(lens[k])
        if homolog[k]:
            anc = ancs[src[k]]
            naa = min(len(anc), n // 3 + 2)
            st = int(rng.integers(0, len(anc) - naa + 1))
            dna = mutate_np(rng, reverse_translate_np(rng, anc[st:st + naa]), indel=0.01, sub=0.005)
            if len(dna) < n:
                dna = np.concatenate([dna, random_dna_np(rng, n - len(dna))])
            reads.append(dna[:n])
        else:
            reads.append(random_dna_np(rng, n))
    per = rng.binomial(n_profiles, pair_rate, size=n_reads)
    seq = np.repeat(np.arange(n_reads, dtype=np.int32), per)
    prof = rng.integers(0, n_profiles, size=len(seq)).astype(np.int32)
    if source_pairs:
        h = np.nonzero(homolog)[0].astype(np.int32)
        seq = np.concatenate([seq, h])
        prof = np.concatenate([prof, src[h].astype(np.int32)])
    return dict(Ms=Ms, raws=raws, reads=reads, prof=prof, seq=seq, homolog=homolog, src=src)


def workload_c1(seed=1001, n_profiles=1000, n_windows=10000, homolog_frac=0.2, max_len=2000):
    """Profiles x windows, all pairs (the forward-only score configuration).
    windows: uniform 1-10 kb, GC uniform 0.35-0.65, 20% carrying an embedded
    frameshifted homolog of a random profile's ancestor."""
    rng = np.random.default_rng(seed)
    Ms = np.minimum(lognormal_lengths(rng, n_profiles), max_len)
    raws, ancs = [], []
    for M in Ms:
        r, a = raw_profile(rng, int(M))
        raws.append(r)
        ancs.append(a)
    reads = []
    for k in range(n_windows):
        n = int(rng.integers(1000, 10001))
        dna = random_dna_np(rng, n, float(rng.uniform(0.35, 0.65)))
        if rng.random() < homolog_frac:
            anc = ancs[int(rng.integers(n_profiles))]
            ins = mutate_np(rng, reverse_translate_np(rng, anc), indel=0.01)[: n // 2]
            at = int(rng.integers(0, n - len(ins) + 1))
            dna[at:at + len(ins)] = ins
        reads.append(dna)
    return dict(Ms=Ms, raws=raws, reads=reads)


def workload_c3(seed=1003, M=300, n_seqs=4096, homolog_frac=0.5, frameshift=0.01):
    """One profile of M positions (built as raw_profile builds them) and DNA
    sequences of length uniform 1-3 kb: a fraction carries the profile's
    ancestor reverse-translated with 1-nt insertions and deletions at rate
    `frameshift`, embedded at a random place; the rest is random (GC 0.5).
    Also the logits at which the gradient is taken: the log of the profile's
    match emissions (M x 20) and of its transitions per state ((M+1) x 7:
    m->m, m->i, m->d, i->m, i->i, d->m, d->d).
    Returns dict(M, raw, seqs (str), emis_logits, trans_logits)."""
    rng = np.random.default_rng(seed)
    raw, anc = raw_profile(rng, M)
    seqs = []
    for _ in range(n_seqs):
        n = int(rng.integers(1000, 3001))
        dna = random_dna_np(rng, n)
        if rng.random() < homolog_frac:
            ins = mutate_np(rng, reverse_translate_np(rng, anc), indel=frameshift)[:n]
            at = int(rng.integers(0, n - len(ins) + 1))
            dna[at:at + len(ins)] = ins
        seqs.append(dna.tobytes().decode())
    tiny = 1e-6
    MtoI, ItoI, MtoD, DtoD = (np.maximum(raw[:, k], tiny) for k in range(4))
    trans = np.log(np.stack([np.maximum(1 - MtoI - MtoD, tiny), MtoI, MtoD,
                             np.maximum(1 - ItoI, tiny), ItoI, np.maximum(1 - DtoD, tiny), DtoD], 1))
    emis = np.log(np.maximum(raw[:M, 4:24], tiny))
    return dict(M=M, raw=raw, seqs=seqs, emis_logits=emis, trans_logits=trans)
