#!/usr/bin/env python3
"""bench.py -- throughput of the gfx950 BiWFA (score + CIGAR, bit-exact) on one MI355X.

A step is one pass of the hot path over one batch: every pair of the workload is
aligned (score + CIGAR) by the persistent kernels, inputs already resident in HBM
and outputs left there.  Default workload: the uneven mixed batch (200,000 pairs,
lengths log-uniform 100-50,000 bp, 1-15 % edits, x=4 o=6 e=2) under the learned
schedule, with the plain length-sorted schedule timed on the same batch in the
same run (fewer steps, stated in the line) and both outputs compared.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload C4]
    python bench.py --workload C1 --affine2p      # gap-affine-2p (0,4,6,2,24,1) on the 2p kernel

Prints one JSON line.  Besides the contract fields it carries
  schedules     -- plain vs learned: ms/step of each, learned_vs_plain, identical_results,
                   cost-model training and prediction figures,
  roofline      -- algorithmic int32 work (11 ops per wavefront cell, 16 under gap-affine-2p,
                   + 4 per 64-bit word compared, counted by the CPU restatement) per step over the total
                   kernel time, against the int32 VALU peak; names the dominant kernel,
  hbm           -- compulsory HBM bytes and the dominant kernel's measured HBM traffic
                   (committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes) per kernel second,
  cpu_baseline  -- WFA2-lib BiWFA (built from the reference sources) with OpenMP on all
                   of this host's cores, on a deterministic subsample of the same workload,
  parity        -- GPU vs CPU reference on that subsample (scores and CIGARs).
"""
import argparse
import glob
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0                  # MI355X HBM3E
AFFINE2P = (0, 4, 6, 2, 24, 1)         # WFA2's gap-affine-2p utest penalties
CLOCK_GHZ = 2.4                        # MI355X peak engine clock
NUM_CUS = 256
# int32 VALU peak: 256 CUs x 64 lanes per clock
INT32_PEAK_TOPS = NUM_CUS * 64 * CLOCK_GHZ * 1e9 / 1e12
# LDS: 256 CUs x 128 bytes per clock
LDS_PEAK_GBS = NUM_CUS * 128 * CLOCK_GHZ


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def compulsory_bytes(p_len, t_len, nruns):
    """Compulsory HBM bytes of a step: packed pattern+text words, per-pair offsets,
    lengths and order (24+4 B) read; score, wf_score, status, run count (16 B) and
    4 B per CIGAR run written."""
    words = (p_len.astype(np.int64) + 31) // 32 + (t_len.astype(np.int64) + 31) // 32
    return int(np.sum(8 * words) + len(p_len) * (24 + 4 + 16) + 4 * int(np.sum(nruns)))


def profile_json(kind, workload, n, penalties=(0, 4, 6, 2)):
    """A committed profiles/*_<kind>.json summary taken on this same workload (and
    penalties), or None."""
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_%s.json" % kind))):
        try:
            with open(path) as f:
                t = json.load(f)
        except (OSError, ValueError):
            continue
        if t.get("workload") == workload and int(t.get("pairs", -1)) == n and \
                list(t.get("penalties", [0, 4, 6, 2])) == list(penalties):
            t["_path"] = os.path.relpath(path, ROOT)
            return t
    return None


def host_cpu():
    """Logical cores, the cores this process may run on, physical cores and the model."""
    model, cores = None, set()
    phys = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    cores.add((phys, v))
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"logical": os.cpu_count(), "usable": usable, "physical": len(cores) or None, "model": model}


def _mismatched_pairs(off_a, runs_a, off_b, runs_b):
    """Number of pairs whose run lists differ (same pair order on both sides)."""
    if np.array_equal(off_a, off_b) and np.array_equal(runs_a[:off_a[-1]], runs_b[:off_b[-1]]):
        return 0
    return sum(1 for i in range(len(off_a) - 1)
               if not np.array_equal(runs_a[off_a[i]:off_a[i + 1]], runs_b[off_b[i]:off_b[i + 1]]))


def gather_runs(res, idx):
    """CIGAR run offsets and runs of the pairs idx of a device result, compacted."""
    lens = (res.cigar_offset[idx + 1] - res.cigar_offset[idx]).astype(np.int64)
    off = np.zeros(len(idx) + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    runs = np.concatenate([res.cigar_runs[res.cigar_offset[i]:res.cigar_offset[i + 1]] for i in idx]) \
        if len(idx) else np.zeros(0, np.uint32)
    return off, runs


def cpu_baseline(ds, res, threads, target_wall_s=8.0, every=None, penalties=(0, 4, 6, 2)):
    """Times the reference BiWFA on all `threads` host cores over a deterministic
    subsample of the workload (every k-th pair, k sized for about target_wall_s of
    wall time; the sample is repeated when one pass is shorter) and checks the GPU
    output on the subsample against it."""
    from oracle.oracle import Oracle, Wfa2Ref
    kind = "reference" if Wfa2Ref.available() else "port"
    eng = Wfa2Ref() if kind == "reference" else Oracle()
    n = len(ds)

    def view(idx):
        return (ds.seq, ds.p_off[idx], ds.p_len[idx], ds.t_off[idx], ds.t_len[idx])
    if every is None:
        # calibrate on every 1000th pair (at least 64 pairs), then size the stride
        probe = np.arange(0, n, max(1, n // max(64, n // 1000)), dtype=np.int64)
        r = eng.align(view(probe), penalties, threads=threads)
        per_pair_wall = max(r.elapsed / len(probe), 1e-9)
        m = int(min(n, max(64, target_wall_s / per_pair_wall)))
        every = max(1, n // m)
    idx = np.arange(0, n, every, dtype=np.int64)
    if len(idx) >= 100:   # 1 % warm-up (a sample of very long pairs is too dear to warm up on)
        eng.align(view(idx[:len(idx) // 100]), penalties, threads=threads)
    r = eng.align(view(idx), penalties, threads=threads)
    wall, reps = r.elapsed, 1
    while wall < target_wall_s * 0.5 and reps < 64:
        wall += eng.align(view(idx), penalties, threads=threads).elapsed
        reps += 1
    cpu_pairs_s = len(idx) * reps / wall
    # the drivers' OpenMP chunk: 16 pairs, fewer when a thread would get under 64 chunks
    per = len(idx) // (64 * threads)
    chunk = 16 if per >= 16 else max(1, per)
    s_bad = int(np.sum(res.score[idx] != r.alg_score))
    ref_off, ref_runs = r.compact()
    g_off, g_runs = gather_runs(res, idx)
    c_bad = _mismatched_pairs(g_off, g_runs, ref_off, ref_runs)
    return {
        "value": round(cpu_pairs_s, 2), "unit": "pairs/s", "cores": threads, "kind": kind,
        "sample": "%d pairs (every %d-th pair of the workload), %d pass(es), %.1f s wall on %d threads; %s" % (
            len(idx), every, reps, wall, threads, "WFA2-lib v2.3 BiWFA built from the reference sources, "
            "ultralow memory, OpenMP dynamic with chunks of %d pairs, one aligner per thread" % chunk
            if kind == "reference" else "C restatement, OpenMP dynamic with chunks of %d pairs" % chunk),
    }, {"checked": int(len(idx)), "every": int(every), "score_mismatches": s_bad, "cigar_mismatches": c_bad}, idx


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="C4")
    ap.add_argument("--pairs", type=int, default=None, help="first N pairs of the workload")
    ap.add_argument("--schedule", default=None, choices=["learned", "length", "input"],
                    help="default: learned (with plain length-sorted timed beside it) for C4, else length")
    ap.add_argument("--plain-steps", type=int, default=1,
                    help="timed steps of the plain length-sorted schedule beside the learned one")
    ap.add_argument("--plain-warmup", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="wall seconds of the CPU baseline")
    ap.add_argument("--cpu-every", type=int, default=None, metavar="K",
                    help="CPU baseline on every K-th pair (default: sized for --cpu-seconds)")
    ap.add_argument("--autotune", type=int, default=0, metavar="N",
                    help="tune launch shapes on an N-pair sample of the workload before timing")
    ap.add_argument("--autotune-repeats", type=int, default=2, metavar="R",
                    help="timed runs per autotune candidate (after one warm-up run)")
    ap.add_argument("--tuning", default=None, help="JSON launch tuning (bw_tuning_t fields)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the captured launch sequence (bw_set_graph); per-tier times read 0")
    ap.add_argument("--train-pairs", type=int, default=None,
                    help="training pairs for the cost model (default: the full 50,000-pair set)")
    ap.add_argument("--affine2p", action="store_true",
                    help="gap-affine-2p penalties (0,4,6,2,24,1) on the 2p kernel, length-sorted schedule")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.empty(args.calibrate_copy << 18, dtype=torch.int32, device="cuda").fill_(1)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copied %d MiB three times" % args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(pairs would shard across GPUs with no exchange but a final gather); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    t_start_all = time.time()
    import biwfa_gpu_pii2025_amd as bw
    cfg = bw.CONFIGS[args.workload]
    n = args.pairs or cfg.num_pairs
    schedule = args.schedule or ("learned" if args.workload == "C4" and not args.affine2p else "length")
    if args.affine2p and schedule == "learned":
        ap.error("the learned schedule is for gap-affine batches")
    penalties = AFFINE2P if args.affine2p else bw.configs.PENALTIES
    compare = schedule == "learned"
    t0 = time.time()
    ds = bw.generate_dataset(cfg, num_pairs=n)
    packed = ds.pack()
    log("generated %d pairs of %s in %.1f s" % (n, args.workload, time.time() - t0))

    base_sched = bw.SCHEDULE_INPUT_ORDER if schedule == "input" else bw.SCHEDULE_LENGTH_SORTED
    aligner = bw.WavefrontAligner(penalties=penalties, schedule=base_sched)
    if args.tuning:
        aligner.tuning = json.loads(args.tuning)
    if args.graph:
        aligner.graph = True
    sample = None
    if args.autotune > 0:
        t0 = time.time()
        sample = bw.generate_dataset(cfg, num_pairs=min(args.autotune, n), seed=cfg.seed + 1000).pack()
        tuned = aligner.autotune(sample, repeats=args.autotune_repeats, progress=log)
        log("autotuned on %d pairs in %.1f s: %s" % (len(sample), time.time() - t0, json.dumps(tuned)))
    db = aligner.upload(packed)

    def timed(label, steps, warmup):
        for w in range(warmup):
            db.align()
            log("%swarmup %d done (%.2f ms)" % (label, w, db.timing()[0]))
        km, kt, ktiers = [], [], []
        t_a = time.perf_counter()
        for k in range(steps):
            db.align()
            tot, main_ms, _ = db.timing()
            kt.append(tot); km.append(main_ms)
            ktiers.append(aligner.tier_timing())
            if steps > 2:
                log("%sstep %d: %.2f ms" % (label, k, tot))
        t_b = time.perf_counter()
        ms = (t_b - t_a) * 1000.0 / steps
        tiers = [round(float(v), 4) for v in np.mean(np.array(ktiers), axis=0)]
        log("%stimed %d steps: %.3f ms/step (kernels %.3f ms; tiers %s; tier pairs %s)" % (
            label, steps, ms, np.mean(kt), tiers, aligner.tier_pairs()))
        return ms, float(np.mean(kt)), tiers

    sched_info = None
    plain_res = None
    if compare:
        from biwfa_gpu_pii2025_amd import costmodel as cm
        plain_ms, plain_kms, plain_tiers = timed("plain: ", args.plain_steps, args.plain_warmup)
        plain_res = db.download()
        plain_pairs = aligner.tier_pairs()
        db.free()
        t0 = time.time()
        model, tinfo = cm.train_cost_model(num_pairs=args.train_pairs, epochs=30, seed=0)
        train_s = time.time() - t0
        log("cost model trained on %d aligner-labelled pairs in %.1f s (mse %.5f -> %.5f)" % (
            tinfo["pairs"], train_s, tinfo["initial_mse"], tinfo["final_mse"]))
        if sample is not None:
            spred = model.predict_score(cm.batch_features(sample)).astype(np.float32)
            aligner.autotune(sample, repeats=args.autotune_repeats, predicted_score=spred, progress=log)
        db = aligner.upload(packed)
        t0 = time.time()
        pred = cm.learned_schedule(db, model, packed)
        sched_s = time.time() - t0
        true_score = plain_res.wf_score.astype(np.float64)
        err = np.abs(np.log2(1 + pred) - np.log2(1 + np.maximum(true_score, 0)))
        rank_p = np.argsort(np.argsort(pred))
        rank_t = np.argsort(np.argsort(true_score))
        sched_info = {
            "plain": {"schedule": "length-sorted", "steps": args.plain_steps, "warmup": args.plain_warmup,
                      "ms_per_step": round(plain_ms, 3), "pairs_s": round(n / plain_ms * 1e3, 2),
                      "kernel_ms": round(plain_kms, 3), "tier_ms": plain_tiers, "tier_pairs": plain_pairs},
            "cost_model": {"train_pairs": tinfo["pairs"], "train_s": round(train_s, 2),
                           "train_mse": [round(tinfo["initial_mse"], 6), round(tinfo["final_mse"], 6)],
                           "predict_and_schedule_s": round(sched_s, 3),
                           "median_abs_log2_err": round(float(np.median(err)), 4),
                           "spearman": round(float(np.corrcoef(rank_p, rank_t)[0, 1]), 4)}}
        model.close()
    ms_per_step, kern_ms, tier_ms = timed("learned: " if compare else "", args.steps, args.warmup)
    tier_pairs = aligner.tier_pairs()
    shapes = aligner.tier_shapes()
    t0 = time.perf_counter()
    res = db.download()
    download_ms = (time.perf_counter() - t0) * 1000.0
    if sched_info is not None:
        same = (np.array_equal(res.score, plain_res.score) and np.array_equal(res.wf_score, plain_res.wf_score)
                and np.array_equal(res.cigar_offset, plain_res.cigar_offset)
                and np.array_equal(res.cigar_runs, plain_res.cigar_runs))
        sched_info["learned"] = {"ms_per_step": round(ms_per_step, 3), "pairs_s": round(n / ms_per_step * 1e3, 2),
                                 "tier_pairs": tier_pairs}
        sched_info["learned_vs_plain"] = round(sched_info["plain"]["ms_per_step"] / ms_per_step, 4)
        sched_info["identical_results"] = bool(same)
    bad_status = int(np.sum(res.status != 0))
    nruns = np.diff(res.cigar_offset)
    value = n / (ms_per_step / 1000.0)
    kern_s = kern_ms / 1000.0

    # dominant kernel: the tier with the largest share of the step's kernel time
    dom = int(np.argmax(tier_ms)) if tier_ms else 0
    dom_shape = shapes[dom] if dom < len(shapes) else {}
    dom_share = float(tier_ms[dom]) / max(1e-9, float(sum(tier_ms))) if tier_ms else None

    out = {
        "metric": "pairs/s (score+CIGAR, bit-exact) on 1 MI355X; learned vs plain schedule; vs CPU",
        "value": round(value, 2),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded generator, seed %d)" % cfg.seed,
        "config": {
            "workload": "%s: %s; %d pairs; %s; score+CIGAR; %s" % (
                cfg.name, cfg.description, n,
                "gap-affine-2p x=4 o1=6 e1=2 o2=24 e2=1" if args.affine2p else "gap-affine x=4 o=6 e=2",
                "learned (cost-model) schedule; plain length-sorted timed beside it on %d step(s)" % args.plain_steps
                if compare else "length-sorted schedule" if schedule == "length" else "input-order schedule"),
            "pairs": n,
            "schedule": schedule,
            "graph_replay": bool(args.graph),
            "pattern_length": [int(packed.p_len.min()), int(packed.p_len.max())],
            "penalties": list(penalties),
        },
        "gcups": round(float(np.sum(packed.p_len.astype(np.float64) * packed.t_len)) / (ms_per_step / 1000.0) / 1e9, 3),
        "kernel_ms": {"total": round(kern_ms, 4), "tiers": tier_ms, "tier_pairs": tier_pairs},
        "step": "reset + every tier's persistent kernel; results (score, wf_score, status, CIGAR runs "
                "in per-pair slots) stay in HBM.  Not timed: the run-compaction kernel and the copy "
                "to the host (download_ms, once after the timed steps)",
        "download_ms": round(download_ms, 3),
        "tiers": shapes,
        "tuning": aligner.tuning,
        "schedules": sched_info,
        "status_errors": bad_status,
    }

    # algorithmic int32 work of the step (counted by the CPU restatement)
    counts = profile_json("work_counts", args.workload, n, penalties)
    s2_all = float(np.sum(res.score.astype(np.float64) ** 2))
    if counts is not None:
        ops = s2_all * counts["ops_per_score_sq"]
        cells = s2_all * counts["sum_cells"] / counts["sum_score_sq"]
        words = s2_all * counts["sum_words"] / counts["sum_score_sq"]
        achieved = ops / kern_s / 1e12
        out["roofline"] = {
            "bound": "valu-int32", "achieved": round(achieved, 4), "peak": round(INT32_PEAK_TOPS, 2),
            "unit": "Tops/s", "frac": round(achieved / INT32_PEAK_TOPS, 5),
            "traffic": None,
            "kernel": dom_shape.get("kernel"), "kernel_share_of_step": round(dom_share, 4) if dom_share else None,
            "time": "total kernel time of the step (all tiers), HIP events",
            "ops_per_step": ops, "cells_per_step": cells, "words_per_step": words,
            "ops_model": counts["ops_model"],
            "counts_from": "%s (CPU restatement on %d sampled pairs; whole batch by sum(score^2) x ops/score^2, "
                           "half-sample spread %.2g)" % (counts["_path"], counts["sample_pairs"],
                                                          counts["half_sample_ratio_spread"]),
            "peak_clock_ghz": CLOCK_GHZ,
        }
        out["cells_per_s"] = round(cells / kern_s, 1)
        lds_bytes = 32.0 * cells + 16.0 * words
        if not args.affine2p:   # the 2p kernel keeps its rings in HBM
            out["lds_model"] = {"achieved": round(lds_bytes / kern_s / 1e9, 1), "peak": LDS_PEAK_GBS, "unit": "GB/s",
                                "frac": round(lds_bytes / kern_s / 1e9 / LDS_PEAK_GBS, 5),
                                "model": "32 B per cell + 16 B per word"}
    comp = compulsory_bytes(packed.p_len, packed.t_len, nruns)
    hbm = {"compulsory_bytes": comp, "compulsory_gbs": round(comp / kern_s / 1e9, 3), "peak": HBM_PEAK_GBS}
    traffic = profile_json("traffic", args.workload, n, penalties)
    if traffic is not None and traffic.get("kernel") == dom_shape.get("kernel"):
        tb = int(traffic["traffic_bytes"])
        dom_s = float(tier_ms[dom]) / 1000.0
        hbm.update({"kernel": traffic["kernel"], "traffic_bytes": tb,
                    "measured_gbs": round(tb / dom_s / 1e9, 1),
                    "measured_frac": round(tb / dom_s / 1e9 / HBM_PEAK_GBS, 4),
                    "traffic_per_cell": round(tb / out["roofline"]["cells_per_step"], 2) if "roofline" in out else None,
                    "from": traffic["_path"]})
        if "roofline" in out:
            out["roofline"]["traffic"] = tb
    out["hbm"] = hbm
    # VALU issue of the dominant kernel: the committed SQ profile's wave
    # instructions per cell (or per pair) at this run's rate.  A wave64 VALU
    # instruction holds one of a CU's 4 SIMDs for 4 clocks, so VALU instructions
    # per CU per clock is the SIMDs' VALU busy fraction.
    sq = profile_json("sq", args.workload, n, penalties)
    if sq is not None and "roofline" in out and tier_ms:
        if "valu_per_cell" in sq:
            valu = sq["valu_per_cell"] * out["roofline"]["cells_per_step"]
            salu = sq["salu_per_cell"] * out["roofline"]["cells_per_step"]
        else:
            valu, salu = sq["valu_per_pair"] * n, sq["salu_per_pair"] * n
        per = NUM_CUS * CLOCK_GHZ * 1e9 * float(tier_ms[dom]) / 1000.0
        out["valu_issue"] = {"kernel": dom_shape.get("kernel"), "valu_per_cu_clock": round(valu / per, 4),
                             "salu_per_cu_clock": round(salu / per, 4), "frac": round(valu / per, 4),
                             "from": sq["_path"]}

    if not args.no_cpu:
        t0 = time.time()
        cpu = host_cpu()
        threads = cpu["usable"]
        every = args.cpu_every or (100 if (args.workload == "C4" and n == cfg.num_pairs) else None)
        base, parity, idx = cpu_baseline(ds, res, threads, target_wall_s=args.cpu_seconds, every=every,
                                         penalties=penalties)
        base.update({"logical_cores": cpu["logical"], "physical_cores": cpu["physical"], "cpu_model": cpu["model"]})
        # the subsample's mean work against the whole batch's (work ~ score^2)
        s2 = res.score.astype(np.float64) ** 2
        ratio = float(np.mean(s2[idx]) / max(1e-9, np.mean(s2)))
        base["sample_work_vs_batch"] = round(ratio, 4)
        base["value_batch_equivalent"] = round(base["value"] * ratio, 2)
        log("cpu baseline %.1f pairs/s on %d threads (%.1f s); sample work / batch work %.3f" % (
            base["value"], threads, time.time() - t0, ratio))
        out["cpu_baseline"] = base
        out["gpu_vs_cpu"] = round(value / base["value"], 2)
        out["gpu_vs_cpu_batch_equivalent"] = round(value / base["value_batch_equivalent"], 2)
        out["parity"] = parity
    log("bench done in %.1f s" % (time.time() - t_start_all))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
