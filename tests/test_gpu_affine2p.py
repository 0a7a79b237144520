"""GPU parity under gap-affine-2p: the single-wave 2p kernel (through the C ABI)
against WFA2's 2p goldens and the CPU oracle, bit-exact on score and CIGAR.

Run on an MI355X:  python -u -m pytest tests/test_gpu_affine2p.py -m gpu -v
"""
import numpy as np
import pytest

from conftest import AFFINE2P, golden_alg2p

pytestmark = pytest.mark.gpu

bw = pytest.importorskip("biwfa_gpu_pii2025_amd")

# penalty sets: the utest default, o2 < o1, e2 > e1, match < 0 (Eizenga transform)
SETS_2P = [AFFINE2P, (0, 4, 8, 2, 4, 3), (0, 3, 24, 1, 6, 2), (-2, 4, 6, 2, 24, 1), (0, 5, 2, 3, 10, 1)]


def _say(msg):
    print("[gpu-affine2p] " + msg, flush=True)


@pytest.fixture(scope="module")
def aligners2p():
    cache = {}

    def get(pen, score_only=False, **kw):
        key = (pen, score_only, tuple(sorted(kw.items())))
        if key not in cache:
            cache[key] = bw.WavefrontAligner(penalties=pen, score_only=score_only, **kw)
        return cache[key]
    yield get
    for a in cache.values():
        a.close()


def _random_pairs(seed, n, max_len):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m = int(rng.integers(0, max_len + 1))
        p = "".join(rng.choice(list("ACGT"), m))
        t = list(p)
        for _ in range(int(rng.integers(0, max(1, m // 5) + 1))):
            if not t:
                break
            j = int(rng.integers(0, len(t)))
            r = rng.random()
            if r < 0.4:
                t[j] = str(rng.choice(list("ACGT")))
            elif r < 0.7:
                t.insert(j, "".join(rng.choice(list("ACGT"), int(rng.integers(1, 40)))))
            else:
                del t[j:j + int(rng.integers(1, 40))]
        out.append((p, "".join(t)))
    return out


def _long_indel_pairs(seed, n_each):
    rng = np.random.default_rng(seed)
    out = []
    for n in (300, 1200, 3000):
        for _ in range(n_each):
            p = "".join(rng.choice(list("ACGT"), n))
            t, i = [], 0
            while i < n:
                r = rng.random()
                if r < 0.01:
                    t.append(str(rng.choice(list("ACGT")))); i += 1
                elif r < 0.013:
                    i += int(rng.integers(20, 200))
                elif r < 0.016:
                    t.extend(rng.choice(list("ACGT"), int(rng.integers(20, 200))))
                else:
                    t.append(p[i]); i += 1
            out.append((p, "".join(t)))
    return out


def _compare(res, ref, pairs, label, check_wf=True):
    n = len(pairs)
    bad = [i for i in range(n) if int(res.status[i]) != int(ref.status[i]) or
           (ref.status[i] == 0 and (int(res.score[i]) != int(ref.alg_score[i]) or res.cigar(i) != ref.cigar(i)))]
    assert not bad, "%s: %d/%d pairs differ, first %d (p=%d t=%d): gpu (%d, %d, %s) vs oracle (%d, %d, %s)" % (
        label, len(bad), n, bad[0], len(pairs[bad[0]][0]), len(pairs[bad[0]][1]), res.status[bad[0]],
        res.score[bad[0]], res.cigar(bad[0])[:60], ref.status[bad[0]], ref.alg_score[bad[0]],
        ref.cigar(bad[0])[:60])
    if check_wf:
        ok = ref.status == 0
        assert np.array_equal(res.wf_score[ok], ref.wf_score[ok]), label


def test_affine2p_utest_goldens(aligners2p, utest_pairs):
    """305 WFA2 utest pairs vs test.biwfa.affine2p.alg and test.biwfa.score.affine2p.alg."""
    _say("utest 2p full")
    gold = golden_alg2p()
    res = aligners2p(AFFINE2P).align_batch(bw.pack_pairs(utest_pairs))
    bad = [i for i, (s, c) in enumerate(gold) if int(res.score[i]) != s or res.cigar(i) != c]
    assert not bad, "first %d: gpu (%d, %s) vs golden %s" % (bad[0], res.score[bad[0]], res.cigar(bad[0])[:80],
                                                           gold[bad[0]][1][:80])
    assert set(int(s) for s in res.status) == {0}
    _say("utest 2p score-only")
    gold = golden_alg2p(score_only=True)
    res = aligners2p(AFFINE2P, score_only=True).align_batch(bw.pack_pairs(utest_pairs))
    assert [int(s) for s in res.score] == [g[0] for g in gold]


@pytest.mark.parametrize("pen", SETS_2P)
def test_affine2p_random_vs_oracle(aligners2p, oracle, pen):
    """Random pairs (gap runs up to 40) and long-indel pairs (runs of 20-200, where
    the second gap model wins) under five 2p penalty sets, full and score-only."""
    pairs = _random_pairs(31, 200, 600) + _long_indel_pairs(32, 4) + [("", ""), ("A", ""), ("", "ACG"),
                                                                      ("ACGT" * 30, "ACGT" * 30)]
    _say("random 2p pen=%s n=%d" % (pen, len(pairs)))
    pk = bw.pack_pairs(pairs)
    ref = oracle.align(pairs, pen, threads=16)
    _compare(aligners2p(pen).align_batch(pk), ref, pairs, "2p %s" % (pen,))
    ref = oracle.align(pairs, pen, score_only=True, threads=16)
    res = aligners2p(pen, score_only=True).align_batch(pk)
    assert np.array_equal(res.status, ref.status)
    ok = ref.status == 0
    assert np.array_equal(res.score[ok], ref.alg_score[ok])
    assert np.array_equal(res.wf_score[ok], ref.wf_score[ok])


def test_affine2p_workloads_vs_oracle(aligners2p, oracle):
    """Seeded 1 kbp (C1) and 10 kbp (C3) samples under the default 2p penalties."""
    for name, n in (("C1", 300), ("C3", 6)):
        ds = bw.generate_dataset(bw.CONFIGS[name], num_pairs=n)
        pairs = ds.pairs()
        _say("workload %s n=%d" % (name, n))
        ref = oracle.align(pairs, AFFINE2P, threads=16)
        _compare(aligners2p(AFFINE2P).align_batch(ds.pack()), ref, pairs, name)


def test_affine2p_max_steps_statuses(aligners2p, oracle):
    """A step cap gives the oracle's status per pair, and its results where it completes."""
    pairs = _random_pairs(33, 150, 500)
    pk = bw.pack_pairs(pairs)
    for cap in (5, 50, 200):
        _say("2p max_steps=%d" % cap)
        ref = oracle.align(pairs, AFFINE2P, max_steps=cap, threads=16)
        res = aligners2p(AFFINE2P, max_steps=cap).align_batch(pk)
        _compare(res, ref, pairs, "cap %d" % cap, check_wf=False)


def test_affine2p_single_pair_api(aligners2p, utest_pairs):
    """bw_align (the wavefront_align surface) on individual utest pairs, one reused aligner."""
    gold = golden_alg2p()
    a = aligners2p(AFFINE2P)
    for i in (0, 1, 2, 50, 150, 304, 3):
        p, t = utest_pairs[i]
        assert a.align(p, t) == 0
        assert (a.score, a.cigar) == (gold[i][0], gold[i][1]), i


def test_affine2p_input_order_schedule_same_results(utest_pairs):
    pk = bw.pack_pairs(utest_pairs[:120])
    with bw.WavefrontAligner(penalties=AFFINE2P, schedule=bw.SCHEDULE_INPUT_ORDER) as a, \
            bw.WavefrontAligner(penalties=AFFINE2P) as b:
        r1, r2 = a.align_batch(pk), b.align_batch(pk)
    assert np.array_equal(r1.score, r2.score)
    assert np.array_equal(r1.cigar_offset, r2.cigar_offset) and np.array_equal(r1.cigar_runs, r2.cigar_runs)


def test_affine2p_rejects_gap_affine_only_paths(aligners2p, utest_pairs):
    """The learned schedule, autotuning and the no-allocation entry point are for
    gap-affine: a 2p aligner gets BW_STATUS_INVALID_ARG with a message."""
    import ctypes
    a = aligners2p(AFFINE2P)
    pk = bw.pack_pairs(utest_pairs[:8])
    L = bw.lib()
    b = pk._struct()
    t = bw._Tuning()
    assert L.bw_autotune(a._h, ctypes.byref(b), None, 1, ctypes.byref(t)) == bw.STATUS_INVALID_ARG
    assert "2p" in bw.last_error()
    db = a.upload(pk)
    try:
        pred = np.ones(len(pk), np.float32)
        assert L.bw_set_schedule(a._h, db._h, bw.SCHEDULE_LEARNED, pred.ctypes.data_as(ctypes.c_void_p)) == \
            bw.STATUS_INVALID_ARG
        res = bw._DevResult()
        assert L.bw_align_device_noalloc(a._h, ctypes.byref(b), ctypes.byref(res)) == bw.STATUS_INVALID_ARG
    finally:
        db.free()
