"""Expected counts of one (profile, DNA) pair from the oracle's forward and
backward matrices.

The oracle's matrices are the reference's recurrences with its per-row
rescaling, returned as FP32 ExpScores (forwardPass dummest-dp-forward.hh:9-250,
backwardPass dummest-dp-backward.hh:444-615). The posterior use of a
transition from cell (i, j) into a cell that the backward holds is
w0(i, j) x coefficient x W1(target) / Z, as forwardPass forms the match
posterior X(i, j) (dummest-dp-forward.hh:126-156), with
Z = sum_j prefixSeed(j) W1(i, j). Five parameters have flows that need only w0
and W1: match entry (enter, with the codon emitted at j+1), the 1- and
2-letter inserts (alpha'1, alpha'2) and the 1- and 2-letter deletes (delta'1,
delta'2)."""
import numpy as np

import oracle as O

PARAM_INDEX = {"a1": 1, "a2": 2, "d1": 7, "d2": 8, "enter": 10}
_BASE = {"A": 0, "C": 1, "G": 2, "T": 3}


def matrix_flows(prof: "O.Profile", dna: str):
    """(per-row counts {name: (M+1,)}, match counts by codon (M+1, 65), ln Z)."""
    M = prof.M
    codes = O.decode(dna)
    n = len(codes)
    (w0m, w0e), (W1m, W1e), ps, sc = O.matrices(prof, codes)
    s1, s2, sC = (float(x) for x in sc)
    # every product keeps its exponent apart (scores may pass FP64's range):
    # Z = Zm 2^ZE; a flow / Z = ldexp(mantissas x coefficient, exponents - ZE) / Zm
    live = (W1m > 0) & (ps[None, :] > 0)
    ZE = int(W1e[live].max()) if live.any() else 0
    Zm = float((np.ldexp(W1m.astype(np.float64), W1e - ZE) * ps.astype(np.float64)).sum())
    P = prof.params.astype(np.float64)[: M + 1]
    w0m, W1m = w0m.astype(np.float64), W1m.astype(np.float64)

    def shifted(A, di, dj, fill=0):  # A(i + di, j + dj), fill outside the matrix
        out = np.full_like(A, fill)
        out[: M + 1 - di, : n - dj] = A[di:, dj:]
        return out

    def flow(coef, di, dj):
        return np.ldexp(w0m * coef * shifted(W1m, di, dj), w0e + shifted(W1e, di, dj) - ZE) / Zm

    code1 = np.zeros(n, np.int64)
    code1[: n - 1] = codes[1:]  # code(j + 1)
    emit = prof.values[:, 4 + code1].astype(np.float64)  # e_i(code(j + 1))
    flows = {
        "enter": flow((P[:, 10] * sC)[:, None] * emit, 1, 3),
        "a1": flow((P[:, 1] * s1)[:, None], 0, 1),
        "a2": flow((P[:, 2] * s2)[:, None], 0, 2),
        "d1": flow((P[:, 7] * s2)[:, None], 1, 2),
        "d2": flow((P[:, 8] * s1)[:, None], 1, 1),
    }
    counts = {k: f.sum(1) for k, f in flows.items()}
    # codon at j+1 as 16 b0 + 4 b1 + b2; 64 for a letter outside ACGT or past the end
    cod = np.full(n, 64)
    for j in range(n - 3):
        b = [_BASE.get(dna[j + k].upper(), -1) for k in (1, 2, 3)]
        if min(b) >= 0:
            cod[j] = 16 * b[0] + 4 * b[1] + b[2]
    by_codon = np.zeros((M + 1, 65))
    for c in np.unique(cod):
        by_codon[:, c] = flows["enter"][:, cod == c].sum(1)
    return counts, by_codon, float(np.log(Zm) + ZE * np.log(2.0))
