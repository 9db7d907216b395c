"""The benchmark workloads, at reduced size, through the engine's batched path
against the oracle.

* uneven reads vs profiles (forward+backward at scale): reads of 150-300 bp
  (half homologs with 1% 1-nt indels and 0.5% substitutions) against profiles
  of lognormal length (median 150, clipped 30-2000, so the row-strip kernels
  run too), pairs from a random mask plus every homologous read with its
  source profile; the autotuned, length-sorted, persistent plan that bench.py
  times (dm_plan_create mode 1 + dm_plan_execute). Every pair's best anchored
  score and backward start score against the oracle within 1e-4 nats, and the
  hits (posterior-decoded alignments) of every pair past the gate identical to
  the oracle's (findFinalSimilaritiesBatched, dummest-search.hh:76-285;
  forwardPass epilogue dummest-dp-forward.hh:456-545).
* profiles x DNA windows (the score-only pass): windows of 1-10 kb, GC
  0.35-0.65, 20% carrying a frameshifted homolog, every pair, through a
  filter plan (dm_plan_create mode 0) against the oracle's
  findSimilaritiesBackwardOnly (dummest-dp-backward.hh:300-438) within 1e-4
  nats; scores that overflow FP32 (no rescale in the filter) are +inf on both
  sides.
* one profile built by dummest-amd-build from a simulated alignment of 50
  descendants, against 5 kb targets (a quarter carrying a reverse-translated
  descendant with 1-nt indels), through the engine executable with and
  without --max: the records against the oracle's.
* the plain form (one launch per pair) and a captured plan: scores
  bit-identical to the batched plan's; a plan cut into many chunks: results in
  the caller's order, bit-identical to the single-chunk plan's.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import dummest_amd as D
import oracle as O
from dummest_amd import synth
from tests.test_gpu_cli import contigs_of, fmt_records, key, parse_records
from tests.test_gpu_parity import check_hits, oracle_cells, reported

pytestmark = pytest.mark.gpu

TOL_NATS = 1e-4
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "dummest_amd", "dummest-amd")
BUILD = os.path.join(ROOT, "dummest_amd", "dummest-amd-build")


def ln(m, e):
    return O.es_log(float(m), int(e))


@pytest.fixture(scope="module")
def reads_case():
    """Engine with the reduced read-vs-profile workload loaded and autotuned."""
    w = synth.workload_c4(seed=1004, n_profiles=200, n_reads=1500, pair_rate=0.005, source_pairs=True)
    e = D.Engine(0)
    pids = e.add_raw_profiles([int(m) for m in w["Ms"]], w["raws"])
    codes = [D.decode(r.tobytes()) for r in w["reads"]]
    sids = e.add_sequences(codes)
    prof = np.array([pids[p] for p in w["prof"]], np.int32)
    seq = np.array([sids[s] for s in w["seq"]], np.int32)
    e.autotune(300)
    profs = {}
    for p in sorted(set(int(x) for x in w["prof"])):
        M, v, pr, b = e.get_profile(pids[p])
        profs[p] = O.Profile(f"p{p}", M, v, pr, b, [])
    yield dict(w=w, e=e, pids=pids, codes=codes, prof=prof, seq=seq, profs=profs)
    e.close()


def oracle_scores(case):
    """Oracle best anchored and backward start scores of every pair (ln)."""
    w, codes, profs = case["w"], case["codes"], case["profs"]
    mid = np.zeros(len(w["prof"]))
    bwd = np.zeros(len(w["prof"]))
    by = {}
    for k, p in enumerate(w["prof"]):
        by.setdefault(int(p), []).append(k)
    for p, ks in by.items():
        rm, rb = O.fwdbwd_batch(profs[p], [codes[w["seq"][k]] for k in ks])
        for k, a, b in zip(ks, rm, rb):
            mid[k] = ln(*a)
            bwd[k] = ln(*b)
    return mid, bwd


def test_uneven_reads_batched_plan(reads_case):
    c = reads_case
    e, w = c["e"], c["w"]
    npairs = len(c["prof"])
    cells = sum((int(w["Ms"][p]) + 1) * len(c["codes"][s]) for p, s in zip(w["prof"], w["seq"]))
    print(f"uneven reads: {npairs} pairs, {cells:.3g} cells, {len(set(w['prof']))} profiles, "
          f"M up to {max(int(w['Ms'][p]) for p in w['prof'])}", flush=True)
    plan = e.plan(c["prof"], c["seq"], "score", want_bwd=True)
    mm, me, bm, be = plan.execute()
    plan.close()
    stats = e.kernel_stats()
    assert len(stats) >= 4, "the mix should use several kernel configurations"
    assert any(s[2] for s in stats), "profiles past 512 rows should run on row-strip kernels"
    mid, bwd = oracle_scores(c)
    c["oracle_scores"] = (mid, bwd)
    d_mid = np.array([abs(ln(mm[k], me[k]) - mid[k]) for k in range(npairs)])
    d_bwd = np.array([abs(ln(bm[k], be[k]) - bwd[k]) for k in range(npairs)])
    print(f"max |d mid| {d_mid.max():.2e}, max |d bwd| {d_bwd.max():.2e} nats", flush=True)
    assert d_mid.max() < TOL_NATS and d_bwd.max() < TOL_NATS
    c["mid_scores"] = (mm, me)


def test_strong_hits_every_strip_configuration(reads_case):
    """Homologous reads of the profiles longer than 512 rows (strong hits, ln
    ratio above 100) under every row-strip kernel: each strip's forward
    starts a new anchor offset far below the hit found in an earlier strip
    (158 nats sits 128 binary orders above it in this workload), which once
    let a clamped running maximum replace the true one."""
    c = reads_case
    e, w = c["e"], c["w"]
    ks = [k for k in range(len(c["prof"]))
          if int(w["Ms"][w["prof"][k]]) > 512 and w["homolog"][w["seq"][k]] and w["src"][w["seq"][k]] == w["prof"][k]]
    assert len(ks) >= 3
    prof, seq = c["prof"][ks], c["seq"][ks]
    mid = []
    for k in ks:
        r = O.search(c["profs"][int(w["prof"][k])], c["codes"][int(w["seq"][k])], -2.0, need_align=False)
        mid.append(ln(*r.global_best))
    assert max(mid) > 100
    table = e.config_table()
    try:
        for idx, (G, R, strips, _) in enumerate(table):
            if not strips:
                continue
            e.set_option("config", idx)
            mm, me, _, _ = e.fwdbwd_scores(prof, seq, want_bwd=False)
            for j in range(len(ks)):
                assert abs(ln(mm[j], me[j]) - mid[j]) < TOL_NATS, (G, R, ks[j], ln(mm[j], me[j]), mid[j])
    finally:
        e.set_option("config", -1)


def test_uneven_reads_hits(reads_case):
    """Every pair past the gate (best anchored score >= e^20, about the
    threshold an E-value of 1e-3 sets for these sizes): hits identical to the
    oracle's."""
    c = reads_case
    e, w = c["e"], c["w"]
    if "mid_scores" not in c:
        plan = e.plan(c["prof"], c["seq"], "score")
        mm, me, _, _ = plan.execute()
        plan.close()
    else:
        mm, me = c["mid_scores"]
    thr_ln = 20.0
    gated = [k for k in range(len(c["prof"])) if mm[k] > 0 and ln(mm[k], me[k]) >= thr_ln]
    print(f"uneven reads: {len(gated)} gated pairs", flush=True)
    assert len(gated) >= 300, "the homologous reads should pass the gate"
    thr = float(np.exp(thr_ln))
    nrec = 0
    for k in gated:
        p, s = int(w["prof"][k]), int(w["seq"][k])
        got = e.find_similarities(int(c["prof"][k]), int(c["seq"][k]), thr)
        ref = O.search(c["profs"][p], c["codes"][s], thr)
        check_hits(got, ref.hits, oracle_cells(c["profs"][p], c["codes"][s]))
        nrec += len(reported(ref.hits))
    assert nrec >= len(gated)


def test_plain_and_captured_plans(reads_case):
    """The plain form (one DP launch per pair) and a captured plan (hipGraph
    replay, twice) give the batched plan's scores bit for bit, on a sample of
    the pairs (every configuration the sample needs)."""
    c = reads_case
    e = c["e"]
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(len(c["prof"]), size=300, replace=False))
    prof, seq = c["prof"][pick], c["seq"][pick]
    base = e.plan(prof, seq, "score", want_bwd=True)
    ref = base.execute()
    base.close()
    cap = e.plan(prof, seq, "score", want_bwd=True)
    cap.capture()
    for _ in range(2):
        got = cap.execute()
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
    cap.close()
    e.set_option("plain", 1)
    try:
        got = e.fwdbwd_scores(prof, seq, want_bwd=True)
    finally:
        e.set_option("plain", 0)
    # the plain form runs one 64-lane group per pair (G = 64 configurations):
    # the grouping of rows differs, so scores agree within the DP tolerance
    for k in range(len(pick)):
        assert abs(ln(got[0][k], got[1][k]) - ln(ref[0][k], ref[1][k])) < TOL_NATS
        assert abs(ln(got[2][k], got[3][k]) - ln(ref[2][k], ref[3][k])) < TOL_NATS


def test_chunked_plan_results_in_caller_order(reads_case):
    """A plan cut into many work-queue chunks (chunk_pairs = 97): every chunk's
    kernels write their pairs' results at the pairs' indices in the caller's
    list, and one copy per execution returns them; the scores equal the
    single-chunk plan's bit for bit (a pair's result does not depend on the
    pairs it shares a unit with), over two executions of each plan (the
    second reads W1 scratch the first left behind)."""
    c = reads_case
    e = c["e"]
    rng = np.random.default_rng(11)
    pick = rng.permutation(len(c["prof"]))[:1200]  # caller order unrelated to the queue order
    prof, seq = c["prof"][pick], c["seq"][pick]
    one = e.plan(prof, seq, "score", want_bwd=True)
    ref = one.execute()
    np.testing.assert_array_equal(ref[0], one.execute()[0])
    one.close()
    e.set_option("chunk_pairs", 97)
    try:
        many = e.plan(prof, seq, "score", want_bwd=True)
        for _ in range(2):
            got = many.execute()
            for a, b in zip(got, ref):
                np.testing.assert_array_equal(a, b)
        many.close()
        fplan = e.plan(prof, seq, "filter")
        fgot = fplan.execute()
        fplan.close()
    finally:
        e.set_option("chunk_pairs", 1 << 22)
    fplan = e.plan(prof, seq, "filter")
    fref = fplan.execute()
    fplan.close()
    np.testing.assert_array_equal(fgot, fref)
    mid, _ = c.get("oracle_scores") or oracle_scores(c)
    for j, k in enumerate(pick[:200]):
        assert abs(ln(ref[0][j], ref[1][j]) - mid[k]) < TOL_NATS


def test_windows_filter_plan():
    w = synth.workload_c1(seed=1001, n_profiles=12, n_windows=60)
    e = D.Engine(0)
    try:
        pids = e.add_raw_profiles([int(m) for m in w["Ms"]], w["raws"])
        codes = [D.decode(r.tobytes()) for r in w["reads"]]
        sids = e.add_sequences(codes)
        P, S = np.meshgrid(np.arange(len(pids)), np.arange(len(sids)), indexing="ij")
        prof = np.array([pids[p] for p in P.ravel()], np.int32)
        seq = np.array([sids[s] for s in S.ravel()], np.int32)
        cells = sum((int(w["Ms"][p]) + 1) * len(codes[s]) for p, s in zip(P.ravel(), S.ravel()))
        print(f"windows: {len(prof)} pairs, {cells:.3g} cells", flush=True)
        plan = e.plan(prof, seq, "filter")
        best = plan.execute()
        plan.close()
        n_inf = 0
        for p in range(len(pids)):
            M, v, pr, b = e.get_profile(pids[p])
            op = O.Profile(f"p{p}", M, v, pr, b, [])
            ref = O.filter_batch(op, codes)
            for s in range(len(sids)):
                g, r = float(best[p * len(sids) + s]), float(ref[s])
                if not np.isfinite(r):
                    n_inf += 1
                    assert g == r, (p, s, g, r)
                    continue
                assert g > 0 and abs(math.log(g) - math.log(r)) < TOL_NATS, (p, s, g, r)
        print(f"windows: {n_inf} scores overflow to +inf", flush=True)
        assert n_inf >= 1, "embedded homologs should overflow the no-rescale filter"
    finally:
        e.close()


def build_profile(tmp_path, rng, L=200):
    text, anc, descs = synth.simulate_msa(rng, L, 50, name="sim")
    (tmp_path / "sim.sto").write_text(text)
    r = subprocess.run([BUILD, str(tmp_path / "sim.sto")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    (tmp_path / "sim.hmm").write_text(r.stdout)
    return r.stdout, descs


@pytest.mark.parametrize("max_mode", [False, True])
def test_built_profile_5kb_targets(tmp_path, max_mode):
    print(f"built profile vs 5 kb targets, --max {max_mode}", flush=True)
    rng = np.random.default_rng(1000)
    hmm, descs = build_profile(tmp_path, rng)
    prof = O.finalize(O.parse_hmm(hmm)[0], np.float32)
    assert prof.M >= 180
    fasta = []
    for k in range(16):
        dna = synth.random_dna(rng, 5000, 0.5)
        if k % 4 == 0:
            ins = synth.mutate_dna(rng, synth.reverse_translate(rng, descs[k]), indel=0.01)
            dna = synth.embed(rng, dna, ins)
        fasta.append(f">t{k}/1-5000 length=5000 profile=sim plus_strand\n")
        fasta.extend(dna[i:i + 80] + "\n" for i in range(0, len(dna), 80))
    fasta_text = "".join(fasta)
    (tmp_path / "t.fa").write_text(fasta_text)
    trials, length, border, evalue = 48, 600, 30, 1.0
    args = [CLI, "-m", "0", "-t", str(trials), "-l", str(length), "-b", str(border), "-e", str(evalue)]
    if max_mode:
        args.append("--max")
    r = subprocess.run(args + [str(tmp_path / "sim.hmm"), str(tmp_path / "t.fa")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    got = parse_records(r.stdout)
    seqs = [O.decode(O.calib_sequence(prof.bg, length, border, t)) for t in range(trials)]
    mid, bwd = O.fwdbwd_batch(prof, seqs)
    gm = O.gumbel([ln(*x) for x in mid], length)
    gf = O.gumbel([ln(*x) for x in bwd], length)
    lam, K, lamf, Kf = gm["mmL"], gm["mmK"], gf["mmL"], gf["mmK"]
    contigs = contigs_of(fasta_text)
    total = sum(len(c[2]) for c in contigs)
    consensus = list(O.parse_hmm(hmm)[0].consensus)
    want = []
    running = 0
    for c in contigs:
        running += len(c[2])
        codes = O.decode(c[2])
        ratio = float(np.float32((K * running / evalue) ** (1 / lam)))
        if not max_mode:  # the filter wave (dummest-search.hh:146-179); --max skips it
            best = O.filter_score(prof, codes)
            if not best > 0:
                continue
            lp = math.log(best)
            if not (Kf * total / math.exp(lp * lamf) <= evalue or not math.isfinite(lp)):
                continue
        res = O.search(prof, codes, ratio)
        for rec in fmt_records("sim", prof.M, consensus, c, res.hits, K, lam, total, evalue):
            rec["codes"], rec["cstart"] = codes, c[1]
            want.append(rec)
    near = lambda rec: abs(rec["E"] - evalue) < 0.01 * evalue
    gk = {key(x): x for x in got if not near(x)}
    wk = {key(x): x for x in want if not near(x)}
    print(f"records: {len(gk)} (oracle {len(wk)})", flush=True)
    assert len(wk) >= 4, "every embedded descendant should be found"
    assert set(gk) == set(wk)
    for k2, wrec in wk.items():
        g = gk[k2]
        assert abs(g["bits"] - wrec["bits"]) < 1e-4 / math.log(2) + 1e-6
        assert g["E"] == pytest.approx(wrec["E"], rel=1e-3)
        if (g["a1"], g["a2"]) != (wrec["a1"], wrec["a2"]):
            (w0m, w0e), (W1m, W1e), _, _ = O.matrices(prof, wrec["codes"])
            i, j = g["a1"], g["a2"] - wrec["cstart"]
            cell = ln(w0m[i, j], w0e[i, j]) + ln(W1m[i, j], W1e[i, j])
            assert abs(cell - wrec["bits"] * math.log(2)) < 1e-4
