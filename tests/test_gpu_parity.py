"""GPU parity: the gfx950 BiWFA (through the C ABI) against the golden WFA2 outputs and
the CPU oracle, bit-exact on score and CIGAR.

Run on an MI355X:  python -u -m pytest tests/test_gpu_parity.py -m gpu -v
"""
import numpy as np
import pytest

from conftest import PENALTY_SETS, golden_alg

pytestmark = pytest.mark.gpu

bw = pytest.importorskip("biwfa_gpu_pii2025_amd")


def _say(msg):
    print("[gpu-parity] " + msg, flush=True)


@pytest.fixture(scope="module")
def aligners():
    cache = {}

    def get(pen, score_only=False, **kw):
        key = (pen, score_only, tuple(sorted(kw.items())))
        if key not in cache:
            cache[key] = bw.WavefrontAligner(penalties=pen, score_only=score_only, **kw)
        return cache[key]
    yield get
    for a in cache.values():
        a.close()


def _compare(res, ref_scores, ref_cigars, label):
    n = len(ref_scores)
    bad = [i for i in range(n) if int(res.score[i]) != int(ref_scores[i]) or res.cigar(i) != ref_cigars[i]]
    st = set(int(s) for s in res.status)
    assert not bad, "%s: %d/%d pairs differ, first %s: gpu (%d, %s) vs ref (%d, %s); status %s" % (
        label, len(bad), n, bad[0], res.score[bad[0]], res.cigar(bad[0])[:80],
        ref_scores[bad[0]], ref_cigars[bad[0]][:80], st)
    assert st == {0}, "%s: unexpected status %s" % (label, st)


@pytest.mark.parametrize("tag", list(PENALTY_SETS))
def test_utest_fixture_full(aligners, utest_pairs, tag):
    """305 WFA2 utest pairs vs test.biwfa.affine<tag>.alg (score + CIGAR)."""
    pen = PENALTY_SETS[tag]
    _say("utest full penalties=%s" % (pen,))
    gold = golden_alg(tag)
    res = aligners(pen).align_batch(bw.pack_pairs(utest_pairs))
    _compare(res, [g[0] for g in gold], [g[1] for g in gold], "utest%s" % tag)


@pytest.mark.parametrize("tag", list(PENALTY_SETS))
def test_utest_fixture_score_only(aligners, utest_pairs, tag):
    pen = PENALTY_SETS[tag]
    _say("utest score-only penalties=%s" % (pen,))
    gold = golden_alg(tag, score_only=True)
    res = aligners(pen, score_only=True).align_batch(bw.pack_pairs(utest_pairs))
    assert [int(s) for s in res.score] == [g[0] for g in gold]
    assert set(int(s) for s in res.status) == {0}


@pytest.mark.parametrize("name,n", [("C1", 400), ("C2", 20000), ("C3", 12), ("C4", 300), ("C5", 4)])
def test_configs_vs_oracle(aligners, oracle, name, n):
    """Seeded samples of every workload vs the CPU oracle (same inputs).  C5 keeps its
    edit model (15% events, 10% long indel runs) at 20 kbp so the oracle finishes in seconds."""
    import dataclasses
    cfg = bw.CONFIGS[name]
    if name == "C5":
        cfg = dataclasses.replace(cfg, len_min=20_000, len_max=20_000, num_edits=3_000)
    _say("config %s n=%d" % (name, n))
    ds = bw.generate_dataset(cfg, num_pairs=n)
    pairs = ds.pairs()
    ref = oracle.align(pairs, threads=16)
    res = aligners((0, 4, 6, 2)).align_batch(ds.pack())
    _compare(res, ref.alg_score, [ref.cigar(i) for i in range(n)], name)
    assert np.array_equal(res.wf_score, ref.wf_score)


def test_c5_real_shape_vs_wfa2_golden(aligners):
    """The first 4 pairs of the divergent-long workload at their real shape (100 kbp,
    15,000 edit events with 10% long indel runs, s ~ 1.9e5, ~11 recursion levels,
    every pair in the widest tier) against WFA2-lib's own output
    (tests/golden/c5, written by tests/golden/make_c5_golden.py): score, CIGAR and
    the level-0 breakpoint score."""
    import gzip
    import json
    import os
    d = os.path.join(os.path.dirname(__file__), "golden", "c5")
    with gzip.open(os.path.join(d, "c5_first4.alg.gz"), "rt") as fh:
        gold = [line.rstrip("\n").split("\t") for line in fh if line.strip()]
    meta = json.load(open(os.path.join(d, "c5_first4.json")))["pairs"]
    ds = bw.generate_dataset(bw.CONFIGS["C5"], num_pairs=len(gold))
    assert [int(x) for x in ds.t_len] == [m["text_length"] for m in meta]
    _say("C5 real shape: %d pairs of 100 kbp" % len(gold))
    res = aligners((0, 4, 6, 2)).align_batch(ds.pack())
    _compare(res, [int(g[0]) for g in gold], [g[1] for g in gold], "C5-100k")
    assert [int(v) for v in res.wf_score] == [-m["wfa2_cigar_score"] for m in meta]


def test_single_pair_api(aligners, utest_pairs):
    """bw_align / accessors on individual pairs (wavefront_align surface)."""
    gold = golden_alg("")
    a = aligners((0, 4, 6, 2))
    for i in (0, 1, 2, 50, 150, 304):
        p, t = utest_pairs[i]
        assert a.align(p, t) == 0
        assert a.score == gold[i][0]
        assert a.cigar == gold[i][1]
        ops = a.cigar_ops
        assert len(ops) == sum(int(r) >> 4 for r in a.sam_cigar())
        assert bw.runs_to_ops(a.sam_cigar()) == ops


@pytest.mark.parametrize("tag", ["", ".p2"])
def test_seq_file_to_alg_file(aligners, tmp_path, tag):
    """align_benchmark's path end to end: the library reads wfa.utest.seq, aligns
    every pair and writes the .alg file, which must equal WFA2's golden byte for
    byte (full and score-only)."""
    import os
    from conftest import UTEST
    pen = PENALTY_SETS[tag]
    with bw.SeqFile(os.path.join(UTEST, "wfa.utest.seq")) as f:
        packed = f.packed()
    for score_only, name in ((False, "test.biwfa.affine%s.alg"), (True, "test.biwfa.score.affine%s.alg")):
        name = name % tag
        _say("seq -> alg %s" % name)
        res = aligners(pen, score_only=score_only).align_batch(packed)
        out = tmp_path / name
        bw.write_alg(str(out), res, score_only=score_only)
        assert out.read_bytes() == open(os.path.join(UTEST, name), "rb").read(), name


@pytest.mark.parametrize("tag", ["", ".p5"])
def test_single_pair_api_every_utest_pair(utest_pairs, oracle, tag):
    """bw_align over all 305 utest pairs on one aligner (its device batch is reused
    from call to call) against test.biwfa.affine<tag>.alg; then a pair longer than
    the reserved capacity (the batch grows), a tuning change (the batch is planned
    again) and the short pairs once more."""
    import time
    pen = PENALTY_SETS[tag]
    gold = golden_alg(tag)
    with bw.WavefrontAligner(penalties=pen) as a:
        assert a.align(*utest_pairs[0]) == 0        # first call: reserves the batch
        us = {False: [], True: []}
        for i, (p, t) in enumerate(utest_pairs):
            t0 = time.perf_counter()
            assert a.align(p, t) == 0, "pair %d status %d" % (i, a.status)
            us[len(p) + len(t) > 4000].append((time.perf_counter() - t0) * 1e6)
            assert (a.score, a.cigar) == gold[i], "pair %d" % i
        _say("bw_align%s: %.1f us/pair over all %d utest pairs; %.1f us/pair over the %d with p+t <= 4000, "
             "%.1f over the %d ~10 kbp pairs (reused device batch; the first 10 kbp pair grows it)" % (
                 tag, np.mean(us[False] + us[True]), len(utest_pairs), np.mean(us[False]), len(us[False]),
                 np.mean(us[True]), len(us[True])))
        rng = np.random.default_rng(7)
        big_p = "".join(rng.choice(list("ACGT"), 3000))
        big_t = big_p[:1000] + "".join(rng.choice(list("ACGT"), 40)) + big_p[1000:2500] + big_p[2600:]
        r = oracle.align([(big_p, big_t)], penalties=pen)
        assert a.align(big_p, big_t) == 0
        assert a.score == int(r.alg_score[0]) and a.cigar == r.cigar(0)
        a.tuning = {"team_waves": [1, 1, 2, 0, 0, 0, 0, 0]}
        for i in (0, 150, 304):
            assert a.align(*utest_pairs[i]) == 0 and (a.score, a.cigar) == gold[i]
        assert a.align(big_p, big_t) == 0 and a.score == int(r.alg_score[0]) and a.cigar == r.cigar(0)


def test_packed2bits_api(aligners, utest_pairs):
    gold = golden_alg("")
    a = aligners((0, 4, 6, 2))
    code = {"A": 0, "C": 1, "G": 2, "T": 3}

    def pack(s):
        out = bytearray((len(s) + 3) // 4)
        for j, ch in enumerate(s):
            out[j >> 2] |= code[ch] << (2 * (j & 3))
        return bytes(out)
    for i in (3, 77, 200):
        p, t = utest_pairs[i]
        assert a.align_packed2bits(pack(p), len(p), pack(t), len(t)) == 0
        assert (a.score, a.cigar) == gold[i]


def test_align_pair_entry(utest_pairs):
    gold = golden_alg("")
    p, t = utest_pairs[10]
    score, ops = bw.align_pair(p, t, 4, 6, 2)
    assert score == gold[10][0]
    from oracle.oracle import cigar_to_runs
    assert ops == bw.runs_to_ops(cigar_to_runs(gold[10][1]))


EDGE_PAIRS = [
    ("", ""), ("A", ""), ("", "ACGT"), ("A", "A"), ("A", "C"), ("ACGT", "ACGT"),
    ("ACGTACGT" * 20, "ACGTACGT" * 20),                    # identical, > 100 bp
    ("A" * 300, "A" * 300), ("A" * 300, "A" * 250),        # homopolymers
    ("A" * 130, "C" * 130),                                # all mismatches
    ("ACGT" * 50, "TGCA" * 60),                            # unrelated
    ("ACGTTGCA" * 30, "ACGTTGCA" * 10 + "G" * 200 + "ACGTTGCA" * 20),   # long insertion
    ("ACGTTGCA" * 40, "ACGTTGCA" * 12 + "ACGTTGCA" * 8),                # long deletion
    ("C" * 101, "C" * 99 + "G"),
    ("GATTACA" * 40, ""),
]


def test_edge_cases(aligners, oracle):
    _say("edge cases")
    pen = (0, 4, 6, 2)
    ref = oracle.align(EDGE_PAIRS, pen)
    res = aligners(pen).align_batch(bw.pack_pairs(EDGE_PAIRS))
    _compare(res, ref.alg_score, [ref.cigar(i) for i in range(len(EDGE_PAIRS))], "edge")


def test_random_small_many_penalties(aligners, oracle):
    """Random short pairs (mixed similar / unrelated) under every utest penalty set."""
    rng = np.random.default_rng(11)
    pairs = []
    for _ in range(400):
        n = int(rng.integers(0, 400))
        p = "".join(rng.choice(list("ACGT"), n))
        if rng.random() < 0.3:
            t = "".join(rng.choice(list("ACGT"), int(rng.integers(0, 400))))
        else:
            t = list(p)
            for _ in range(int(rng.integers(0, max(1, n // 5) + 1))):
                if not t:
                    break
                j = int(rng.integers(0, len(t)))
                r = rng.random()
                if r < 0.34:
                    t[j] = str(rng.choice(list("ACGT")))
                elif r < 0.67:
                    t.insert(j, str(rng.choice(list("ACGT"))) * int(rng.integers(1, 20)))
                else:
                    del t[j:j + int(rng.integers(1, 20))]
            t = "".join(t)
        pairs.append((p, t))
    for pen in PENALTY_SETS.values():
        _say("random small penalties=%s" % (pen,))
        ref = oracle.align(pairs, pen, threads=16)
        res = aligners(pen).align_batch(bw.pack_pairs(pairs))
        _compare(res, ref.alg_score, [ref.cigar(i) for i in range(len(pairs))], "random%s" % (pen,))


def test_schedules_and_determinism(aligners, oracle):
    """Input order, length-sorted and an arbitrary learned order give identical results."""
    _say("schedules")
    ds = bw.generate_dataset(bw.CONFIGS["C4"], num_pairs=200)
    a = aligners((0, 4, 6, 2))
    db = a.upload(ds.pack())
    outs = []
    for sched, cost in ((bw.SCHEDULE_INPUT_ORDER, None), (bw.SCHEDULE_LENGTH_SORTED, None),
                        (bw.SCHEDULE_LEARNED, 400.0 * np.random.default_rng(0).random(200))):
        db.set_schedule(sched, cost)
        db.align()
        outs.append(db.download())
    for r in outs[1:]:
        assert np.array_equal(r.score, outs[0].score)
        assert np.array_equal(r.cigar_offset, outs[0].cigar_offset)
        assert np.array_equal(r.cigar_runs, outs[0].cigar_runs)
    db.free()


def test_learned_schedule_merges_an_underfilled_tier(aligners, oracle):
    """Predictions that straddle the 256/1024 boundary on like pairs would split them
    over two sequential, under-filled single-wave launches; the learned schedule
    hands the narrower tier's pairs to the wider one, which runs them all."""
    _say("learned tier merge")
    ds = bw.generate_dataset(bw.CONFIGS["C1"], num_pairs=300)
    ref = oracle.align(ds.pairs(), threads=16)
    a = aligners((0, 4, 6, 2))
    db = a.upload(ds.pack())
    # 100 -> first tier at least 1.25*100/2+8 = 70.5 wide (256); 500 -> 320.5 (1024)
    db.set_schedule(bw.SCHEDULE_LEARNED, np.where(np.arange(300) % 2 == 0, 100.0, 500.0))
    db.align()
    res = db.download()
    shapes = a.tier_shapes()
    assert [s["ring_width"] for s in shapes[1:3]] == [256, 1024]
    assert a.tier_pairs()[1:3] == [0, 300]
    _compare(res, ref.alg_score, [ref.cigar(i) for i in range(300)], "learned-merge")
    db.free()


def test_spill_chain_under_wrong_predictions(aligners, oracle):
    """Long divergent pairs predicted at score 0 start in the smallest workspace tier
    and must climb the spill chain tier by tier; results stay bit-exact."""
    import dataclasses
    _say("spill chain")
    c3 = bw.generate_dataset(bw.CONFIGS["C3"], num_pairs=6).pairs()
    c5 = bw.generate_dataset(dataclasses.replace(bw.CONFIGS["C5"], len_min=12_000, len_max=12_000,
                                                 num_edits=1_800), num_pairs=3).pairs()
    c1 = bw.generate_dataset(bw.CONFIGS["C1"], num_pairs=40).pairs()
    pairs = c3 + c5 + c1 + EDGE_PAIRS
    ref = oracle.align(pairs, (0, 4, 6, 2), threads=16)
    a = aligners((0, 4, 6, 2))
    db = a.upload(bw.pack_pairs(pairs))
    db.set_schedule(bw.SCHEDULE_LEARNED, np.zeros(len(pairs)))
    db.align()
    res = db.download()
    _compare(res, ref.alg_score, [ref.cigar(i) for i in range(len(pairs))], "spill-chain")
    db.free()


def test_max_steps_status(aligners, oracle):
    """A tiny step cap is reported like WFA2 (MAX_STEPS from the breakpoint search,
    UNATTAINABLE when the base case hits it)."""
    ds = bw.generate_dataset(bw.CONFIGS["C1"], num_pairs=20)
    pairs = ds.pairs() + [("ACGT" * 10, "ACGA" * 10)]
    ref = oracle.align(pairs, (0, 4, 6, 2), max_steps=60)
    res = aligners((0, 4, 6, 2), max_steps=60).align_batch(bw.pack_pairs(pairs))
    assert list(res.status) == list(ref.status)


def test_tuning_changes_speed_not_results(aligners, oracle):
    """Every team size and LDS ring width (the runtime-width LDS kernels 16, 64 and
    128 included) matches the oracle bit for bit on a mixed batch and on the utest
    pairs; the tuning round-trips through the ABI and bad values are rejected."""
    _say("tunings")
    from oracle.oracle import read_seq_file
    import os
    ds = bw.generate_dataset(bw.CONFIGS["C4"], num_pairs=120)
    short = bw.generate_dataset(bw.CONFIGS["C2"], num_pairs=300)
    utest = read_seq_file(os.path.join(os.path.dirname(__file__), "golden", "wfa2_utest", "wfa.utest.seq"))
    pairs = ds.pairs() + short.pairs() + utest[:120]
    packed = bw.pack_pairs(pairs)
    ref = oracle.align(pairs, threads=16)
    ref_cigars = [ref.cigar(i) for i in range(len(pairs))]
    for nw, lw in ((1, 16), (2, 32), (4, 64), (8, 32), (16, 128)):
        a = bw.WavefrontAligner(penalties=(0, 4, 6, 2))
        t = {"lds_ring_width": lw, "lds_seq_cap": 0, "team_waves": [nw] * bw.MAX_TIERS,
             "queue_chunk": [3] * bw.MAX_TIERS}
        a.tuning = t
        assert a.tuning["team_waves"] == [nw] * bw.MAX_TIERS and a.tuning["lds_ring_width"] == lw
        res = a.align_batch(packed)
        _compare(res, ref.alg_score, ref_cigars, "tuning nw=%d lds_ring_width=%d" % (nw, lw))
        assert np.array_equal(res.wf_score, ref.wf_score)
        a.close()
    a = bw.WavefrontAligner(penalties=(0, 4, 6, 2))
    with pytest.raises(bw.BiwfaError):
        a.tuning = {"team_waves": [3]}
    with pytest.raises(bw.BiwfaError):
        a.tuning = {"lds_ring_width": 48}
    a.close()
    # accepted values whose LDS workspaces do not fit a CU together (penalties with a
    # ring depth of 31 scores at ring width 64: 190 KiB per workgroup of 4 waves) are
    # refused at upload, before any launch, with the sizes in the message
    a = bw.WavefrontAligner(penalties=(0, 4, 20, 10))
    a.tuning = {"lds_ring_width": 64}
    with pytest.raises(bw.BiwfaError, match="more than a CU's 160 KiB"):
        a.align_batch(bw.pack_pairs(pairs[:4]))
    a.close()


def test_autotune_reports_and_keeps_results(aligners, oracle):
    """bw_autotune times candidates (progress lines), keeps a valid tuning, and the
    tuned aligner still matches the oracle bit for bit; per-tier pairs add up to
    the batch plus the re-runs."""
    _say("autotune")
    ds = bw.generate_dataset(bw.CONFIGS["C4"], num_pairs=150)
    packed = ds.pack()
    oref = oracle.align(ds.pairs(), threads=16)
    a = bw.WavefrontAligner(penalties=(0, 4, 6, 2))
    lines = []
    t = a.autotune(packed, repeats=1, progress=lines.append)
    assert len(lines) >= 5 and all(line.startswith("autotune") for line in lines)
    assert all(v in (1, 2, 4, 8, 16) for v in t["team_waves"][:6])
    db = a.upload(packed)
    db.align()
    res = db.download()
    pairs = a.tier_pairs()
    assert sum(pairs) >= len(packed) and pairs[0] + sum(pairs[1:]) == sum(pairs)
    _compare(res, oref.alg_score, [oref.cigar(i) for i in range(len(packed))], "autotuned")
    db.free()


def test_autotune_long_pairs_keeps_lds_settings(oracle):
    """A sample with no pair in the LDS tier (10 kbp pairs) tunes the global tiers
    only: one progress line per timed run, the LDS-tier settings left as they were,
    and the tuned aligner still bit-exact against the oracle."""
    _say("autotune long pairs")
    ds = bw.generate_dataset(bw.CONFIGS["C3"], num_pairs=12)
    packed = ds.pack()
    oref = oracle.align(ds.pairs(), threads=16)
    a = bw.WavefrontAligner(penalties=(0, 4, 6, 2))
    lines = []
    t = a.autotune(packed, repeats=1, progress=lines.append)
    assert any("no sample pair starts there" in line for line in lines)
    assert not any("lds queue_chunk" in line or "lds_ring_width" in line for line in lines)
    # five team candidates, each a warm-up and one timed run
    assert sum(1 for line in lines if line.startswith("autotune run")) == 10
    assert t["lds_ring_width"] == 0 and t["queue_chunk"][0] == 0
    res = a.align_batch(packed)
    assert a.tier_pairs()[0] == 0
    _compare(res, oref.alg_score, [oref.cigar(i) for i in range(len(packed))], "autotuned-long")
    a.close()
    a.close()


def _mutated_pairs(rng, n, lo, hi, rate):
    pairs = []
    for _ in range(n):
        m = int(rng.integers(lo, hi))
        p = "".join(rng.choice(list("ACGT"), m))
        t = []
        for ch in p:
            r = rng.random()
            if r < rate / 3:
                t.append(str(rng.choice([c for c in "ACGT" if c != ch])))
            elif r < 2 * rate / 3:
                t.append(ch + "".join(rng.choice(list("ACGT"), int(rng.integers(1, 6)))))
            elif r < rate:
                continue
            else:
                t.append(ch)
        pairs.append((p, "".join(t)))
    return pairs


@pytest.mark.parametrize("tag", list(PENALTY_SETS))
def test_global_tiers_all_penalties(aligners, oracle, tag):
    """Pairs too long for the LDS tier (1-6 kbp, so both the depth-first and the
    shared-task paths of the global tiers run) under every utest penalty set,
    including odd penalties (no null-score shortcut) and match < 0."""
    pen = PENALTY_SETS[tag]
    _say("global tiers penalties=%s" % (pen,))
    rng = np.random.default_rng(17)
    pairs = _mutated_pairs(rng, 24, 1000, 6000, 0.06)
    ref = oracle.align(pairs, pen, threads=16)
    res = aligners(pen).align_batch(bw.pack_pairs(pairs))
    _compare(res, ref.alg_score, [ref.cigar(i) for i in range(len(pairs))], "global%s" % tag)
    assert np.array_equal(res.wf_score, ref.wf_score)


def test_global_tiers_step_cap_and_score_only(aligners, oracle):
    """Step caps and score-only mode on pairs that run in the global tiers."""
    _say("global tiers caps")
    rng = np.random.default_rng(23)
    pairs = _mutated_pairs(rng, 16, 2000, 7000, 0.08)
    for cap in (300, 2000):
        ref = oracle.align(pairs, (0, 4, 6, 2), max_steps=cap, threads=16)
        res = aligners((0, 4, 6, 2), max_steps=cap).align_batch(bw.pack_pairs(pairs))
        assert list(res.status) == list(ref.status), cap
    ref = oracle.align(pairs, (0, 4, 6, 2), score_only=True, threads=16)
    res = aligners((0, 4, 6, 2), score_only=True).align_batch(bw.pack_pairs(pairs))
    assert set(int(s) for s in res.status) == {0}
    assert np.array_equal(res.wf_score, ref.wf_score)
    assert np.array_equal(res.score, ref.alg_score)


def test_graph_replay_matches_eager(oracle):
    """Graph mode: the captured launch sequence replayed several times, with other
    eager work and synchronisations in between, gives the oracle's results every
    time (scores, CIGARs, statuses), on a batch that spans the LDS tier and the
    global tiers' shared recursion tasks."""
    _say("graph replay")
    ds = bw.generate_dataset(bw.CONFIGS["C4"], num_pairs=160)
    short = bw.generate_dataset(bw.CONFIGS["C2"], num_pairs=200)
    pairs = ds.pairs() + short.pairs()
    packed = bw.pack_pairs(pairs)
    other = bw.generate_dataset(bw.CONFIGS["C1"], num_pairs=300).pack()
    eager = bw.WavefrontAligner(penalties=(0, 4, 6, 2))
    ref = eager.align_batch(packed)
    oref = oracle.align(pairs, threads=16)
    _compare(ref, oref.alg_score, [oref.cigar(i) for i in range(len(pairs))], "graph-eager")
    ref_other = eager.align_batch(other)
    g = bw.WavefrontAligner(penalties=(0, 4, 6, 2))
    g.graph = True
    assert g.graph
    db = g.upload(packed)
    for rep in range(4):
        db.align()
        res = db.download()
        assert np.array_equal(res.status, ref.status), rep
        assert np.array_equal(res.score, ref.score), rep
        assert np.array_equal(res.cigar_offset, ref.cigar_offset), rep
        assert np.array_equal(res.cigar_runs, ref.cigar_runs), rep
        assert sum(g.tier_pairs()) >= len(packed)
        # eager work of the program between replays, on the same device
        assert np.array_equal(eager.align_batch(other).score, ref_other.score)
        _say("replay %d identical" % rep)
    # switching graph mode off returns to eager launches with per-tier timings
    g.graph = False
    db.align()
    res = db.download()
    assert np.array_equal(res.score, ref.score) and np.array_equal(res.cigar_runs, ref.cigar_runs)
    assert sum(g.tier_timing()) > 0
    db.free()
    g.close()
    eager.close()


def test_device_batch_belongs_to_its_uploader():
    """A device batch's tiers are planned for the aligner that uploaded it; another
    aligner (other penalties) is refused instead of running kernels whose
    compile-time layout does not match the batch."""
    _say("owner check")
    pairs = bw.generate_dataset(bw.CONFIGS["C2"], num_pairs=50).pairs()
    a = bw.WavefrontAligner(penalties=(0, 2, 3, 1))
    b = bw.WavefrontAligner(penalties=(0, 4, 6, 2))
    db = a.upload(bw.pack_pairs(pairs))
    assert b._L.bw_dbatch_align(b._h, db._h) == bw.STATUS_INVALID_ARG
    assert b._L.bw_set_schedule(b._h, db._h, bw.SCHEDULE_LENGTH_SORTED, None) == bw.STATUS_INVALID_ARG
    db.align()
    res = db.download()
    assert set(int(s) for s in res.status) == {0}
    db.free()
    a.close()
    b.close()


def test_ready_guard_gives_up_and_launch_ends(oracle):
    """Failure path of the shared-task recursion, forced from the host: with the
    ready flags of published tasks withheld and the wait bounded at 1 ms, the
    aligner that claims a task gives up, every aligner of the tier leaves its work
    loop, the launch ends and the call reports BW_STATUS_INTERNAL.  The same
    aligner then aligns the pairs bit-exactly again."""
    _say("ready guard")
    pairs = bw.generate_dataset(bw.CONFIGS["C3"], num_pairs=2).pairs()
    packed = bw.pack_pairs(pairs)
    a = bw.WavefrontAligner(penalties=(0, 4, 6, 2))
    a.set_ready_guard(ticks=100_000, withhold=True)
    with pytest.raises(bw.BiwfaError, match="no progress"):
        a.align_batch(packed)
    a.set_ready_guard(0, False)
    res = a.align_batch(packed)
    ref = oracle.align(pairs, threads=16)
    _compare(res, ref.alg_score, [ref.cigar(i) for i in range(len(pairs))], "after-guard")
    a.close()


def test_c_stub_prints_the_golden_alg_lines(tmp_path, utest_pairs):
    """The integration guide's C stub (tests/c_stub/alg_line.c), compiled against include/biwfa_mi355x.h and run as
    its own process: its line for a utest pair equals WFA2's golden .alg line."""
    import subprocess
    from test_abi import build_c_stub
    exe = build_c_stub(tmp_path)
    gold = golden_alg("")
    for i in (0, 1, 152, 304):
        p, t = utest_pairs[i]
        _say("c stub pair %d" % i)
        r = subprocess.run([exe, p, t], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert r.stdout == "%d\t%s\n" % gold[i]


@pytest.mark.parametrize("tag", ["", ".p1", ".p4"])
def test_sam_cigar_without_mismatches_matches_wfa2(aligners, utest_pairs, wfa2ref, tag):
    """bw_get_sam_cigar(show_mismatches=0) against WFA2's own
    cigar_get_CIGAR(cigar, false) (alignment/cigar.c:181-240), through oracle/_ref,
    on every utest pair and on random pairs."""
    pen = PENALTY_SETS[tag]
    rng = np.random.default_rng(11)
    pairs = list(utest_pairs)
    for _ in range(40):
        p = "".join(rng.choice(list("ACGT"), int(rng.integers(1, 400))))
        t = list(p)
        for _ in range(int(rng.integers(0, 30))):
            j = int(rng.integers(0, len(t) + 1))
            k = int(rng.integers(0, 3))
            if k == 0 and j < len(t):
                t[j] = str(rng.choice(list("ACGT")))
            elif k == 1:
                t.insert(j, str(rng.choice(list("ACGT"))))
            elif j < len(t):
                del t[j]
        pairs.append((p, "".join(t) or "A"))
    ref = wfa2ref.align(pairs, penalties=pen, show_mismatches=False)
    ref_eq = wfa2ref.align(pairs, penalties=pen)
    a = aligners(pen)
    for i, (p, t) in enumerate(pairs):
        assert a.align(p, t) == 0
        got = a.sam_cigar(False).tolist()
        assert got == [int(x) for x in ref.cigar_runs(i)], "pair %d" % i
        assert a.sam_cigar(True).tolist() == [int(x) for x in ref_eq.cigar_runs(i)], "pair %d" % i
