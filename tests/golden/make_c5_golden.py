"""Writes the C5 golden vectors: the first 4 pairs of the divergent-long workload
(100 kbp patterns, 15,000 edit events with 10% indel runs of U{10..200} bases,
seed 5, x=4 o=6 e=2) aligned by WFA2-lib v2.3 compiled from the reference sources
(oracle/_ref/libwfa2ref.so, exact BiWFA: memory ultralow, heuristic none).

    python tests/golden/make_c5_golden.py      (about 25 s on 4 cores)

Outputs, under tests/golden/c5/:
  c5_first4.alg.gz  -- align_benchmark's .alg lines: "<score>\\t<CIGAR>"
                       (tools/align_benchmark/benchmark/benchmark_utils.c:143-179)
  c5_first4.json    -- per pair: lengths, SHA-256 of pattern and text (so the test
                       can confirm the generator still yields the same inputs),
                       WFA2's cigar->score and status.
Test infrastructure only: the product never reads these files.
"""
import gzip
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

N = 4


def main():
    import biwfa_gpu_pii2025_amd as bw
    from oracle.oracle import Wfa2Ref
    ds = bw.generate_dataset(bw.CONFIGS["C5"], num_pairs=N)
    ref = Wfa2Ref().align((ds.seq, ds.p_off, ds.p_len, ds.t_off, ds.t_len), threads=N)
    out = os.path.join(HERE, "c5")
    os.makedirs(out, exist_ok=True)
    meta = []
    with gzip.open(os.path.join(out, "c5_first4.alg.gz"), "wt") as fh:
        for i in range(N):
            p, t = ds.pair(i)
            fh.write("%d\t%s\n" % (ref.alg_score[i], ref.cigar(i)))
            meta.append({"pattern_length": len(p), "text_length": len(t),
                         "pattern_sha256": hashlib.sha256(p.encode()).hexdigest(),
                         "text_sha256": hashlib.sha256(t.encode()).hexdigest(),
                         "wfa2_cigar_score": int(ref.raw_score[i]), "status": int(ref.status[i]),
                         "runs": int(ref.nruns[i])})
    with open(os.path.join(out, "c5_first4.json"), "w") as fh:
        json.dump({"workload": "C5", "seed": 5, "pairs": meta,
                   "aligner": "WFA2-lib v2.3 from the reference sources, exact BiWFA",
                   "wall_s": round(ref.elapsed, 1)}, fh, indent=1)
    print("wrote %d pairs (%.1f s)" % (N, ref.elapsed))


if __name__ == "__main__":
    main()
