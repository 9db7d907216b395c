"""The oracle's DP recurrences pinned by explicit enumeration of alignment
paths (tiny profiles and sequences, FP64).

An alignment of DNA positions 0..n-1 to profile rows 0..M is a path of moves
between row/column states. Node W(i, j) means "row i, nucleotides up to and
including j accounted for". A path starts at W(i, j) with the prefix null mass
prefixSeed[j] (nucleotides 0..j unaligned), makes moves, and ends at some
W(i', j') with suffixSeed[j'] (nucleotides j'+1..n-1 unaligned). The moves,
as the reference's forward recurrence states them (dummest-dp-forward.hh:123,
158-165, 218-222; the parameter order alpha'[3] beta'[3] delta'[3] eps' enter
of dummest-core.hh:233-239):

  W(i,j) -> W(i+1,j+3)  match: enter_i, emission of the codon j+1..j+3
  W(i,j) -> W(i+1,j+2)  delete1 frameshift: delta'_i[1]
  W(i,j) -> W(i+1,j+1)  delete2 frameshift: delta'_i[2]
  W(i,j) -> W(i,j+1)    insert1 frameshift: alpha'_i[1]
  W(i,j) -> W(i,j+2)    insert2 frameshift: alpha'_i[2]
  W(i,j) -> Z(i,j+3)    open a codon insertion: alpha'_i[0], background of the codon j+1..j+3
  Z(i,j) -> Z(i,j+3)    extend it: beta'_i[0], background of the codon j+1..j+3
  Z(i,j) -> W(i,j)      close it
  W(i,j) -> D(i,j)      open a deletion: delta'_i[0]
  D(i,j) -> D(i+1,j)    extend it: eps'_{i+1}
  D(i,j) -> W(i+1,j)    close it

Two rules fix the rest, independently of how the recurrences are coded: a
move that consumes k nucleotides carries the null model's per-base scale to
the power k (s1 = scale/4 with the 1/4 of a random letter, s2 = scale^2/16,
sC = scale^3), so every complete path carries scale^n exactly; a codon's
emission or background is that of the codon starting at its first consumed
nucleotide. The path masses give:
  * the total Z = sum over paths (forward total sum w0 suffixSeed and
    backward total sum prefixSeed W1 of dummest-dp-backward.hh/-forward.hh);
  * the filter score max over start cells of the mass of paths starting there
    (max W1 prefixSeed, dummest-dp-backward.hh:300-438);
  * the best anchored score max over cells of the mass of paths through it
    (max w0 W1, dummest-dp-forward.hh:174-206), and per column the best such
    cell and its row.
Each is compared with the oracle in FP64; mistakes planted in the oracle
(scales swapped, a column offset) must make the comparison fail.
"""
import math

import numpy as np
import pytest

import oracle as O

OPTS = O.DEFAULT_OPTS


def random_case(rng, M, n):
    """A finalized-profile-shaped random model: every parameter positive (so
    every move is live), emissions and background positive on codes 0..23 and
    zero on the padding code 24."""
    values = np.zeros((M + 1, 29), np.float64)
    values[:, 4:28] = rng.uniform(0.01, 0.12, (M + 1, 24))
    params = np.zeros((M + 2, 11), np.float64)
    params[:M + 1] = rng.uniform(0.02, 0.6, (M + 1, 11))
    bg = np.zeros(29, np.float64)
    bg[4:28] = rng.uniform(0.01, 0.08, 24)
    codes = rng.integers(0, 24, n).astype(np.uint8)
    return O.Profile("bf", M, values, params, bg, []), codes


def enumerate_paths(prof, codes, null):
    """Every complete path, its mass; returns (total, per-start mass, per-W-node
    through mass)."""
    M, n = prof.M, len(codes)
    ps, ss = null["prefix_seed"], null["suffix_seed"]
    s1, s2, sC = (float(x) for x in null["scales"])
    kscale = {1: s1, 2: s2, 3: sC}
    par = prof.params
    em = prof.values[:, 4:]
    bgc = prof.bg[4:]
    start = np.zeros((M + 1, n))
    through = np.zeros((M + 1, n))
    total = 0.0
    # depth-first over (node, mass so far, W nodes visited)
    for i0 in range(M + 1):
        for j0 in range(n):
            stack = [(("W", i0, j0), float(ps[j0]), ((i0, j0),))]
            while stack:
                (kind, i, j), m, visited = stack.pop()
                if kind == "W":
                    done = m * float(ss[j])
                    total += done
                    start[i0, j0] += done
                    for (a, b) in visited:
                        through[a, b] += done
                    for k, a in ((1, 1), (2, 2)):  # insert1, insert2 frameshifts
                        if j + k < n:
                            stack.append((("W", i, j + k), m * par[i, a] * kscale[k], visited + ((i, j + k),)))
                    if j + 3 < n:  # open a codon insertion
                        stack.append((("Z", i, j + 3), m * par[i, 0] * kscale[3] * bgc[codes[j + 1]], visited))
                    stack.append((("D", i, j), m * par[i, 6], visited))  # open a deletion
                    if i + 1 <= M:
                        if j + 3 < n:  # match
                            stack.append((("W", i + 1, j + 3), m * par[i, 10] * kscale[3] * em[i, codes[j + 1]],
                                          visited + ((i + 1, j + 3),)))
                        for k, a in ((2, 7), (1, 8)):  # delete1 (2 nt), delete2 (1 nt) frameshifts
                            if j + k < n:
                                stack.append((("W", i + 1, j + k), m * par[i, a] * kscale[k], visited + ((i + 1, j + k),)))
                elif kind == "Z":
                    stack.append((("W", i, j), m, visited + ((i, j),)))
                    if j + 3 < n:
                        stack.append((("Z", i, j + 3), m * par[i, 3] * kscale[3] * bgc[codes[j + 1]], visited))
                else:  # D
                    if i + 1 <= M:
                        stack.append((("W", i + 1, j), m, visited + ((i + 1, j),)))
                        stack.append((("D", i + 1, j), m * par[i + 1, 9], visited))
    return total, start, through


def oracle_values(prof, codes):
    bwd, fwd = O.totals(prof, codes, rescale_bits=64)
    flt = O.filter_score(prof, codes)
    r = O.search(prof, codes, -2.0, need_align=False, want_bwd=True)
    col = np.array([O.es_log(float(m), int(e)) for m, e, _ in r.col_best])
    return dict(bwd=bwd, fwd=fwd, filter=math.log(flt), mid=O.es_log(*r.global_best),
                bwd_best=O.es_log(*r.bwd_best), col=col, col_row=r.col_best[:, 2].astype(int))


CASES = [(1, 5, 11), (2, 7, 12), (3, 9, 13), (2, 9, 14)]


@pytest.fixture(scope="module", params=CASES, ids=[f"M{m}-n{n}" for m, n, _ in CASES])
def case(request):
    M, n, seed = request.param
    prof, codes = random_case(np.random.default_rng(seed), M, n)
    null = O.null_model(prof.bg, codes, OPTS)
    total, start, through = enumerate_paths(prof, codes, null)
    return prof, codes, null, total, start, through


def test_paths_pin_the_oracle(case):
    prof, codes, null, total, start, through = case
    o = oracle_values(prof, codes)
    lz = math.log(total)
    assert o["fwd"] == pytest.approx(lz, abs=1e-10)
    assert o["bwd"] == pytest.approx(lz, abs=1e-10)
    assert o["filter"] == pytest.approx(math.log(start.max()), abs=1e-10)
    # the backward start score of the search is the same maximum, as an ExpScore
    assert o["bwd_best"] == pytest.approx(math.log(start.max()), abs=1e-7)
    # the best anchored score is formed from float mantissas (ExpScore)
    assert o["mid"] == pytest.approx(math.log(through.max()), abs=1e-7)
    # per column: the best anchor (dummest-dp-forward.hh:196-206) and its row
    np.testing.assert_allclose(o["col"], np.log(through.max(0)), atol=1e-7, rtol=0)
    np.testing.assert_array_equal(o["col_row"], through.argmax(0))


def test_paths_carry_scale_to_the_n(case):
    """Scaling the null model's per-base scale by t (s1 t, s2 t^2, sC t^3,
    prefixSeed[j] t^(j+1), suffixSeed[j] t^(n-1-j)) scales every path, hence
    the total, by t^n exactly: the move table's scale powers match the
    nucleotides the moves consume."""
    prof, codes, null, total, _, _ = case
    n = len(codes)
    t = 2.0
    j = np.arange(n)
    scaled = dict(prefix_seed=null["prefix_seed"] * t ** (j + 1), suffix_seed=null["suffix_seed"] * t ** (n - 1 - j),
                  scales=null["scales"] * np.array([t, t * t, t ** 3]))
    total2, _, _ = enumerate_paths(prof, codes, scaled)
    assert total2 == pytest.approx(total * t ** n, rel=1e-13)


@pytest.mark.parametrize("mutation,key", [(1, "bwd"), (2, "fwd"), (3, "filter"), (4, "bwd"), (1, "mid"), (2, "mid")])
def test_planted_mistakes_fail(case, mutation, key):
    prof, codes, null, total, start, through = case
    want = {"fwd": math.log(total), "bwd": math.log(total), "filter": math.log(start.max()),
            "mid": np.log(through.max(0))}[key]
    O.set_mutation(mutation)
    try:
        got = oracle_values(prof, codes)["col" if key == "mid" else key]
    finally:
        O.set_mutation(0)
    # "mid": the per-column best anchors (some column must move)
    assert np.max(np.abs(np.asarray(got) - want)) > 1e-6, (mutation, key, got, want)
