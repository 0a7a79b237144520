#!/usr/bin/env python3
"""bench.py -- throughput of the gfx950 BiWFA (score + CIGAR, bit-exact) on one MI355X.

A step is one pass of the hot path over one batch: every pair of the workload is
aligned (score + CIGAR) by the persistent kernels, inputs already resident in HBM
and outputs left there.  Default workload: the short-read batch (1,000,000 pairs
x 150 bp, 2% edits, x=4 o=6 e=2, length-sorted schedule).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload C2]

Prints one JSON line.  Besides the contract fields it carries
  roofline      -- HBM roofline of the dominant kernel (algorithmic bytes / kernel time),
  issue_roofline -- instruction issue rate of that kernel (the bound it runs against:
                   one SALU and one VALU wave instruction per CU per clock; the
                   busier of the two is reported), from the kernel's per-pair SQ
                   counts in profiles/ and the live kernel time,
  int_roofline  -- the algorithmic integer work (cells and compares) against the
                   int32 VALU peak,
  cpu_baseline  -- WFA2-lib BiWFA (built from the reference sources) with OpenMP
                   on this host's cores, on a bounded sample of the same workload,
  parity        -- GPU vs CPU reference on that sample (scores and CIGARs).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0                  # MI355X HBM3E
# instruction issue: one scalar (SALU) and one vector (VALU) wave instruction per CU
# per clock, 256 CUs at 2.4 GHz
ISSUE_PEAK_G = 256 * 2.4
# int32 VALU peak: 256 CUs x 4 SIMD x 16 lanes per clock at 2.4 GHz
INT32_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(p_len, t_len, nruns):
    """Compulsory HBM bytes: packed pattern+text words, per-pair offsets/lengths/order
    (24+4 B) read; score, wf_score, status, run count (16 B) and 4 B per CIGAR run written."""
    words = (p_len.astype(np.int64) + 31) // 32 + (t_len.astype(np.int64) + 31) // 32
    return int(np.sum(8 * words) + len(p_len) * (24 + 4 + 16) + 4 * int(np.sum(nruns)))


def traffic_bytes(workload, n):
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (scaled by a plain copy's readings), when they
    were taken on this same workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "c2_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or int(t.get("pairs", -1)) != n:
        return None
    return int(t["traffic_bytes"])


def sq_counts(workload, n):
    """Per-pair SALU / VALU wave-instruction counts of the dominant kernel from the
    committed rocprofv3 SQ counter pass, when it was taken on this same workload."""
    path = os.path.join(ROOT, "profiles", "c2_sq.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or int(t.get("pairs", -1)) != n:
        return None
    return t


def _per_pair_mismatches(off_a, runs_a, off_b, runs_b, m):
    """Number of pairs among the first m whose run lists differ."""
    if np.array_equal(off_a[:m + 1], off_b[:m + 1]) and \
            np.array_equal(runs_a[:off_a[m]], runs_b[:off_b[m]]):
        return 0
    bad = 0
    for i in range(m):
        if not np.array_equal(runs_a[off_a[i]:off_a[i + 1]], runs_b[off_b[i]:off_b[i + 1]]):
            bad += 1
    return bad


def cpu_baseline(ds, res, min_core_s=12.0, threads=None):
    """Times the reference BiWFA on a bounded sample of the workload (repeating the
    sample until about min_core_s core-seconds of CPU work are timed) and checks the
    GPU output against it; also runs the oracle restatement for the work counters."""
    from oracle.oracle import Oracle, Wfa2Ref
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count())
    kind = "reference" if Wfa2Ref.available() else "port"
    eng = Wfa2Ref() if kind == "reference" else Oracle()
    n = len(ds)

    def view(m):
        return (ds.seq, ds.p_off[:m], ds.p_len[:m], ds.t_off[:m], ds.t_len[:m])
    # calibrate on a growing slice (8, 16, 32, ... pairs until it takes about half a
    # second), then size the sample to ~min_core_s core-seconds
    probe = min(n, 8)
    while True:
        r = eng.align(view(probe), threads=threads)
        if r.elapsed >= 0.5 or probe >= n or probe >= 20000:
            break
        probe = min(n, 2 * probe)
    per_pair_core_s = max(r.elapsed * threads / probe, 1e-8)
    m = int(min(n, max(probe, min_core_s / per_pair_core_s)))
    if m > probe:
        r = eng.align(view(m), threads=threads)
    wall, reps = r.elapsed, 1
    while wall * threads < min_core_s:
        wall += eng.align(view(m), threads=threads).elapsed
        reps += 1
    cpu_pairs_s = m * reps / wall
    # parity against the GPU output on the sample
    s_bad = int(np.sum(res.score[:m] != r.alg_score[:m]))
    ref_off, ref_runs = r.compact()
    c_bad = _per_pair_mismatches(res.cigar_offset, res.cigar_runs, ref_off, ref_runs, m)
    # algorithmic work counters (cells, 64-bit words) from the restatement
    cnt_m = min(m, 100000, max(8, int(4.0 / per_pair_core_s)))   # a few core-seconds of counting
    oc = Oracle().align(view(cnt_m), threads=threads)
    o_off, o_runs = oc.compact()
    o_bad = _per_pair_mismatches(o_off, o_runs, ref_off, ref_runs, cnt_m) + \
        int(np.sum(oc.alg_score != r.alg_score[:cnt_m]))
    return {
        "value": round(cpu_pairs_s, 1), "unit": "pairs/s", "cores": threads, "kind": kind,
        "sample": "first %d pairs of the same workload, %d pass(es), %.1f s wall on %d threads (%s)" % (
            m, reps, wall, threads, "WFA2-lib v2.3 BiWFA built from the reference sources, "
            "OpenMP dynamic,16, one aligner per thread" if kind == "reference"
            else "C restatement, OpenMP dynamic,16"),
    }, {
        "checked": m, "score_mismatches": s_bad, "cigar_mismatches": c_bad,
        "oracle_vs_reference_mismatches": o_bad,
    }, {
        "cells_per_pair": float(np.mean(oc.cells)), "words_per_pair": float(np.mean(oc.words)),
        "pairs_counted": cnt_m,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="C2")
    ap.add_argument("--pairs", type=int, default=None, help="first N pairs of the workload")
    ap.add_argument("--schedule", default="length", choices=["length", "input"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", type=int, default=0, metavar="N",
                    help="tune launch shapes on an N-pair sample of the workload before timing")
    ap.add_argument("--tuning", default=None, help="JSON launch tuning (bw_tuning_t fields)")
    ap.add_argument("--graph", action="store_true",
                    help="replay the captured launch sequence (bw_set_graph); per-tier times read 0")
    ap.add_argument("--compare-schedules", action="store_true",
                    help="train the cost model, then time the plain length-sorted and the learned "
                         "schedule on the same batch (value = learned)")
    ap.add_argument("--train-pairs", type=int, default=None,
                    help="training pairs for the cost model (default: the full 50,000-pair set)")
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

    import biwfa_gpu_pii2025_amd as bw
    cfg = bw.CONFIGS[args.workload]
    n = args.pairs or cfg.num_pairs
    t0 = time.time()
    ds = bw.generate_dataset(cfg, num_pairs=n)
    packed = ds.pack()
    log("generated %d pairs of %s in %.1f s" % (n, args.workload, time.time() - t0))

    sched = bw.SCHEDULE_LENGTH_SORTED if args.schedule == "length" else bw.SCHEDULE_INPUT_ORDER
    aligner = bw.WavefrontAligner(penalties=bw.configs.PENALTIES, schedule=sched)
    if args.tuning:
        aligner.tuning = json.loads(args.tuning)
    if args.graph:
        aligner.graph = True
    sample = None
    if args.autotune > 0:
        t0 = time.time()
        sample = bw.generate_dataset(cfg, num_pairs=min(args.autotune, n), seed=cfg.seed + 1000).pack()
        tuned = aligner.autotune(sample, repeats=2, progress=log)
        log("autotuned on %d pairs in %.1f s: %s" % (len(sample), time.time() - t0, json.dumps(tuned)))
    db = aligner.upload(packed)

    def timed(label):
        for w in range(args.warmup):
            db.align()
            log("%swarmup %d done (%.2f ms)" % (label, w, db.timing()[0]))
        km, ks, kt, ktiers = [], [], [], []
        t_start = time.perf_counter()
        for k in range(args.steps):
            db.align()
            tot, main_ms, spill_ms = db.timing()
            kt.append(tot); km.append(main_ms); ks.append(spill_ms)
            ktiers.append(aligner.tier_timing())
        t_end = time.perf_counter()
        ms = (t_end - t_start) * 1000.0 / args.steps
        tiers = [round(float(v), 4) for v in np.mean(np.array(ktiers), axis=0)]
        log("%stimed %d steps: %.3f ms/step (kernels %.3f ms: main %.3f + spill %.3f; tiers %s; "
            "tier pairs %s)" % (label, args.steps, ms, np.mean(kt), np.mean(km), np.mean(ks), tiers,
                                 aligner.tier_pairs()))
        return ms, km, ks, kt, tiers

    sched_info = None
    if args.compare_schedules:
        from biwfa_gpu_pii2025_amd import costmodel as cm
        plain_ms = timed("plain: ")[0]
        plain_res = db.download()
        plain_tuning = aligner.tuning
        db.free()
        t0 = time.time()
        model, tinfo = cm.train_cost_model(num_pairs=args.train_pairs, epochs=30, seed=0)
        train_s = time.time() - t0
        log("cost model trained on %d aligner-labelled pairs in %.1f s (mse %.5f -> %.5f)" % (
            tinfo["pairs"], train_s, tinfo["initial_mse"], tinfo["final_mse"]))
        if sample is not None:
            t0 = time.time()
            spred = model.predict_score(cm.batch_features(sample)).astype(np.float32)
            tuned = aligner.autotune(sample, repeats=2, predicted_score=spred, progress=log)
            log("learned schedule autotuned in %.1f s: %s" % (time.time() - t0, json.dumps(tuned)))
        db = aligner.upload(packed)
        t0 = time.time()
        pred = cm.learned_schedule(db, model, packed)
        sched_s = time.time() - t0
        true_score = plain_res.wf_score.astype(np.float64)
        err = np.abs(np.log2(1 + pred) - np.log2(1 + np.maximum(true_score, 0)))
        rank_p = np.argsort(np.argsort(pred))
        rank_t = np.argsort(np.argsort(true_score))
        spearman = float(np.corrcoef(rank_p, rank_t)[0, 1])
        sched_info = {"plain_ms_per_step": round(plain_ms, 4), "plain_pairs_s": round(n / plain_ms * 1e3, 1),
                      "plain_tuning": plain_tuning,
                      "train_pairs": tinfo["pairs"], "train_s": round(train_s, 2),
                      "train_mse": [round(tinfo["initial_mse"], 6), round(tinfo["final_mse"], 6)],
                      "predict_and_schedule_s": round(sched_s, 3),
                      "score_prediction": {"median_abs_log2_err": round(float(np.median(err)), 4),
                                           "spearman": round(spearman, 4)}}
        model.close()
    ms_per_step, kern_main, kern_spill, kern_total, tier_ms = timed("learned: " if args.compare_schedules else "")
    res = db.download()
    if sched_info is not None:
        same = (np.array_equal(res.score, plain_res.score) and np.array_equal(res.cigar_offset, plain_res.cigar_offset)
                and np.array_equal(res.cigar_runs, plain_res.cigar_runs))
        sched_info["learned_pairs_s"] = round(n / ms_per_step * 1e3, 1)
        sched_info["learned_vs_plain"] = round(sched_info["plain_ms_per_step"] / ms_per_step, 3)
        sched_info["identical_results"] = bool(same)
    bad_status = int(np.sum(res.status != 0))
    nruns = np.diff(res.cigar_offset)

    value = n / (ms_per_step / 1000.0)
    main_s = float(np.mean(kern_main)) / 1000.0
    alg_bytes = algorithmic_bytes(packed.p_len, packed.t_len, nruns)
    achieved_gbs = alg_bytes / main_s / 1e9
    out = {
        "metric": "pairs/s (score+CIGAR, bit-exact) on 1 MI355X; learned vs plain schedule; vs CPU",
        "value": round(value, 1),
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
            "workload": "%s: %s; %d pairs; gap-affine x=4 o=6 e=2; score+CIGAR; %s schedule" % (
                cfg.name, cfg.description, n,
                "learned (cost model)" if args.compare_schedules else
                "length-sorted" if args.schedule == "length" else "input-order"),
            "pairs": n,
            "graph_replay": bool(args.graph),
            "pattern_length": [int(packed.p_len.min()), int(packed.p_len.max())],
            "penalties": list(bw.configs.PENALTIES),
        },
        "gcups": round(float(np.sum(packed.p_len.astype(np.float64) * packed.t_len)) / (ms_per_step / 1000.0) / 1e9, 3),
        "kernel_ms": {"total": round(float(np.mean(kern_total)), 4), "main": round(float(np.mean(kern_main)), 4),
                      "spill": round(float(np.mean(kern_spill)), 4), "tiers": tier_ms},
        "tuning": aligner.tuning,
        "schedules": sched_info,
        "status_errors": bad_status,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
            "traffic": traffic_bytes(args.workload, n),
            "kernel": "biwfa_persistent<Cfg<true,1,32,true>> (LDS tier, one wave per pair, ring width 32, default penalties)",
            "algorithmic_bytes_per_launch": alg_bytes,
        },
    }
    sq = sq_counts(args.workload, n)
    if sq is not None:
        salu_g = sq["salu_per_pair"] * n / main_s / 1e9
        valu_g = sq["valu_per_pair"] * n / main_s / 1e9
        top = max(salu_g, valu_g)
        out["issue_roofline"] = {
            "bound": "valu-issue" if valu_g >= salu_g else "salu-issue", "achieved": round(top, 1),
            "peak": ISSUE_PEAK_G, "unit": "G instr/s", "frac": round(top / ISSUE_PEAK_G, 4),
            "salu_frac": round(salu_g / ISSUE_PEAK_G, 4), "valu_frac": round(valu_g / ISSUE_PEAK_G, 4),
            "salu_per_pair": round(sq["salu_per_pair"], 1), "valu_per_pair": round(sq["valu_per_pair"], 1),
            "counts_from": "profiles/biwfa_persistent_lds_C2_1M_sq_counters.md",
        }
    if not args.no_cpu:
        t0 = time.time()
        cpu, parity, counters = cpu_baseline(ds, res)
        log("cpu baseline %.0f pairs/s on %d threads (%.1f s)" % (cpu["value"], cpu["cores"], time.time() - t0))
        out["cpu_baseline"] = cpu
        out["gpu_vs_cpu"] = round(value / cpu["value"], 2)
        out["parity"] = parity
        ops_per_pair = 11.0 * counters["cells_per_pair"] + 4.0 * counters["words_per_pair"]
        achieved_tops = ops_per_pair * n / main_s / 1e12
        out["int_roofline"] = {
            "bound": "valu-int32", "achieved": round(achieved_tops, 4), "peak": round(INT32_PEAK_TOPS, 2),
            "unit": "Tops/s", "frac": round(achieved_tops / INT32_PEAK_TOPS, 5),
            "ops_per_pair": round(ops_per_pair, 1), "counters": counters,
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
