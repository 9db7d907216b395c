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


def make_profile(rng, L: int, name: str, ndesc=50):
    """(HMMER3 text, ancestral protein as letter indices, descendants)."""
    anc = rng.integers(0, 20, size=L)
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
