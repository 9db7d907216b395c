"""Shared fixtures for the parity tests: seeded synthetic profiles and DNA."""
import gzip
import os

import numpy as np

import oracle as O
from dummest_amd import synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_fasta(path):
    out, name, buf = [], None, []
    for line in open(path):
        line = line.strip()
        if line.startswith(">"):
            if name is not None:
                out.append((name, "".join(buf)))
            name, buf = line[1:].split()[0], []
        elif line:
            buf.append(line)
    if name is not None:
        out.append((name, "".join(buf)))
    return out


def golden_hmm(name):
    return gzip.open(os.path.join(GOLDEN, name), "rt").read()


def synthetic_case(seed, L, lengths, homolog_every=2):
    """Profile text, oracle profile (float32), DNA strings (some with homologs)."""
    rng = np.random.default_rng(seed)
    text, anc, descs = synth.make_profile(rng, L, f"syn{seed}")
    prof = O.finalize(O.parse_hmm(text)[0], np.float32)
    dnas = []
    for k, n in enumerate(lengths):
        host = synth.random_dna(rng, n, 0.5)
        if homolog_every and k % homolog_every == 0:
            d = descs[k % len(descs)]
            ins = synth.mutate_dna(rng, synth.reverse_translate(rng, d), indel=0.01)
            host = synth.embed(rng, host, ins)
        dnas.append(host)
    return text, prof, dnas


def ln_es(m, e):
    return O.es_log(float(m), int(e))
