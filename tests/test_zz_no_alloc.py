"""bw_align_device_noalloc (the entry point that allocates no device memory) against
bw_align_batch on the same pairs, through the C ABI: scores, internal scores,
statuses and CIGAR runs bit for bit, on two calls in a row (the second smaller,
so it must not see the first call's scratch contents).  Inputs and outputs are
device buffers made with torch (device memory only: the alignment runs in the
library's kernels).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

bw = pytest.importorskip("biwfa_gpu_pii2025_amd")
torch = pytest.importorskip("torch")

MAX_PT = 2040   # longest pair (p + t) the entry point aligns; longer ones get DEVICE_CAPACITY


def _say(msg):
    print("[no-alloc] " + msg, flush=True)


def _noalloc(aligner, packed, shrink_capacity=()):
    """Runs bw_align_device_noalloc on device copies of `packed`; returns host arrays.
    Pairs listed in shrink_capacity get one run slot less than p + t + 1."""
    n = len(packed)
    dev = torch.device("cuda:0")
    cap = packed.p_len.astype(np.int64) + packed.t_len + 1
    for i in shrink_capacity:
        cap[i] -= 1
    run_off = np.zeros(n + 1, np.int64)
    np.cumsum(cap, out=run_off[1:])
    d = {
        "words": torch.from_numpy(packed.words.view(np.int64).copy()).to(dev),
        "p_woff": torch.from_numpy(packed.p_woff).to(dev),
        "p_len": torch.from_numpy(packed.p_len).to(dev),
        "t_woff": torch.from_numpy(packed.t_woff).to(dev),
        "t_len": torch.from_numpy(packed.t_len).to(dev),
        "run_off": torch.from_numpy(run_off).to(dev),
        # outputs start as garbage: every field must be written by the call
        "score": torch.full((n,), 777, dtype=torch.int32, device=dev),
        "wf": torch.full((n,), 777, dtype=torch.int32, device=dev),
        "status": torch.full((n,), 777, dtype=torch.int32, device=dev),
        "nruns": torch.full((n,), 777, dtype=torch.int32, device=dev),
        "runs": torch.full((max(1, int(run_off[-1])),), 0x7777, dtype=torch.int32, device=dev),
    }
    b = bw._Batch()
    b.num_pairs = n
    b.num_words = len(packed.words)
    b.words = d["words"].data_ptr()
    b.pattern_word_offset = d["p_woff"].data_ptr()
    b.pattern_length = d["p_len"].data_ptr()
    b.text_word_offset = d["t_woff"].data_ptr()
    b.text_length = d["t_len"].data_ptr()
    r = bw._DevResult()
    r.score, r.wf_score = d["score"].data_ptr(), d["wf"].data_ptr()
    r.status, r.num_runs = d["status"].data_ptr(), d["nruns"].data_ptr()
    r.runs, r.run_offset = d["runs"].data_ptr(), d["run_off"].data_ptr()
    torch.cuda.synchronize()
    L = aligner._L
    st = L.bw_align_device_noalloc(aligner._h, ctypes.byref(b), ctypes.byref(r))
    assert st == 0, "bw_align_device_noalloc: %d (%s)" % (st, bw.last_error())
    st = L.bw_synchronize(aligner._h)
    assert st == 0, "bw_synchronize: %d (%s)" % (st, bw.last_error())
    out = {k: d[k].cpu().numpy() for k in ("score", "wf", "status", "nruns")}
    out["runs"] = d["runs"].cpu().numpy().view(np.uint32)
    out["run_off"] = run_off
    return out


def _check(label, pairs, out, ref, bad_capacity=()):
    n = len(pairs)
    checked = 0
    for i in range(n):
        p, t = pairs[i]
        if i in bad_capacity:
            assert out["status"][i] == bw.STATUS_INVALID_ARG and out["nruns"][i] == 0, (label, i)
            continue
        if len(p) + len(t) > MAX_PT:
            assert out["status"][i] == bw.STATUS_DEVICE_CAPACITY and out["nruns"][i] == 0, (label, i)
            continue
        got = out["runs"][out["run_off"][i]:out["run_off"][i] + out["nruns"][i]]
        assert (out["status"][i], out["score"][i], out["wf"][i]) == \
            (ref.status[i], ref.score[i], ref.wf_score[i]), (label, i)
        assert np.array_equal(got, ref.runs(i)), (label, i, bw.runs_to_cigar(got)[:60], ref.cigar(i)[:60])
        checked += 1
    return checked


def _mix(seed, n_short, n_1k, n_div):
    rng = np.random.default_rng(seed)
    pairs = bw.generate_dataset(bw.CONFIGS["C2"], num_pairs=n_short, seed=100 + seed).pairs()
    pairs += bw.generate_dataset(bw.CONFIGS["C1"], num_pairs=n_1k, seed=200 + seed).pairs()
    for _ in range(n_div):   # unrelated pairs: they outgrow the LDS tier and spill
        m = int(rng.integers(100, 500))
        pairs.append(("".join(rng.choice(list("ACGT"), m)),
                      "".join(rng.choice(list("ACGT"), int(rng.integers(100, 500))))))
    pairs += [("", ""), ("A", ""), ("", "ACGT"), ("ACGT" * 30, "ACGT" * 30),
              ("ACGT" * 400, "ACGA" * 400),   # p + t = 3200: beyond the limit
              ("A" * 1020, "A" * 1020)]       # p + t = 2040: at the limit
    return pairs


def test_noalloc_matches_batch_entry_two_calls():
    a = bw.WavefrontAligner(penalties=(0, 4, 6, 2))
    for call, (seed, n_short, n_1k, n_div) in enumerate(((1, 4000, 300, 120), (2, 700, 40, 30))):
        pairs = _mix(seed, n_short, n_1k, n_div)
        packed = bw.pack_pairs(pairs)
        _say("call %d: %d pairs" % (call, len(pairs)))
        ref = a.align_batch(packed)
        bad = (5, len(pairs) - 3) if call == 1 else ()
        out = _noalloc(a, packed, shrink_capacity=bad)
        checked = _check("call %d" % call, pairs, out, ref, bad_capacity=bad)
        assert checked >= len(pairs) - 3
        # the divergent pairs really took the spill tier
        if call == 0:
            assert a.tier_pairs()[1] > 0
        _say("call %d: %d pairs identical" % (call, checked))
    a.close()


def test_noalloc_other_penalties_and_limits():
    a = bw.WavefrontAligner(penalties=(-3, 5, 3, 2))
    pairs = _mix(3, 500, 30, 40)
    packed = bw.pack_pairs(pairs)
    ref = a.align_batch(packed)
    out = _noalloc(a, packed)
    assert _check("p5", pairs, out, ref) >= len(pairs) - 1
    # more pairs than the scratch holds: refused before anything is enqueued
    b = bw._Batch()
    b.num_pairs = (1 << 20) + 1
    r = bw._DevResult()
    assert a._L.bw_align_device_noalloc(a._h, ctypes.byref(b), ctypes.byref(r)) == bw.STATUS_INVALID_ARG
    assert "num_pairs" in bw.last_error()
    a.close()
