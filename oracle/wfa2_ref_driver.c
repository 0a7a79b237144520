/*
 * wfa2_ref_driver.c -- thin batch driver around the vendored WFA2-lib v2.3.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
 * WFA2-lib sources where they lie under the reference checkout (see
 * oracle/Makefile) into oracle/_ref/libwfa2ref*.so.  It is used by the tests
 * to pin the CPU restatement (oracle/biwfa_oracle.c) and by bench.py's
 * cpu_baseline leg ("kind": "reference").  Nothing in the product library
 * links or loads it.
 *
 * Aligner settings are the exact-BiWFA ones (the library defaults are the
 * WF-adaptive heuristic in high-memory mode, wavefront_attributes.c:71-78):
 *   distance_metric = gap_affine (or gap_affine_2p), heuristic = none,
 *   memory_mode = ultralow (BiWFA, wavefront_bialign.c:703),
 *   alignment_scope = compute_alignment (or compute_score).
 * One aligner per OpenMP thread, reused across pairs, as in
 * tools/align_benchmark/align_benchmark.c:504-586.
 *
 * Outputs per pair:
 *   alg_score  : the score column align_benchmark writes to .alg files
 *                (tools/align_benchmark/benchmark/benchmark_utils.c:143-179): cigar_score_gap_affine
 *                (cigar_score_gap_affine2p) of the CIGAR for full alignments, cigar->score for score-only,
 *                -1 when the CIGAR is empty.
 *   raw_score  : aligner->cigar->score as left by wavefront_align.
 *   status     : wavefront_align's return code (wfa.h:46-55).
 *   runs       : cigar_get_CIGAR(cigar, show_mismatches, ...) (cigar.c:181-240):
 *                uint32 (len<<4)|op with op '='=7, X=8, I=1, D=2 (show_mismatches
 *                true, the default), or M=0, I=1, D=2 with X folded into M
 *                (false; wfa2ref_set_show_mismatches).
 */
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

#include "wavefront/wfa.h"
#include "wavefront/wavefront_align.h"
#include "alignment/cigar.h"

static int g_show_mismatches = 1;
void wfa2ref_set_show_mismatches(int show) { g_show_mismatches = show ? 1 : 0; }

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static wavefront_aligner_t* new_aligner(const affine2p_penalties_t* p, int two_piece,
                                        int score_only, int max_steps) {
  wavefront_aligner_attr_t attr = wavefront_aligner_attr_default;
  if (two_piece) {
    attr.distance_metric = gap_affine_2p;
    attr.affine2p_penalties = *p;
  } else {
    attr.distance_metric = gap_affine;
    attr.affine_penalties.match = p->match;
    attr.affine_penalties.mismatch = p->mismatch;
    attr.affine_penalties.gap_opening = p->gap_opening1;
    attr.affine_penalties.gap_extension = p->gap_extension1;
  }
  attr.heuristic.strategy = wf_heuristic_none;
  attr.memory_mode = wavefront_memory_ultralow;
  attr.alignment_scope = score_only ? compute_score : compute_alignment;
  if (max_steps > 0) attr.system.max_alignment_steps = max_steps;
  return wavefront_aligner_new(&attr);
}

/*
 * Align pairs [0,n).  Sequences are ASCII in seqbuf at the given offsets.
 * runs_off[i]..runs_off[i+1] is the run capacity of pair i; nruns[i] is set to
 * the number of runs (or -1 when the capacity was too small).
 * Returns the wall time of the parallel region in *elapsed (seconds).
 */
static void align_batch(
    int n, const char* seqbuf,
    const int64_t* p_off, const int32_t* p_len,
    const int64_t* t_off, const int32_t* t_len,
    const affine2p_penalties_t* pen2, int two_piece, int score_only, int max_steps, int nthreads,
    int32_t* alg_score, int32_t* raw_score, int32_t* status,
    uint32_t* runs, const int64_t* runs_off, int32_t* nruns,
    double* elapsed) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const double t0 = now_seconds();
  /* chunks of 16 pairs (dynamic), fewer when that would leave a thread with
   * under 64 chunks: a small sample of long pairs must still spread over
   * every thread */
  const int chunk = n / (64 * nthreads) >= 16 ? 16 : (n / (64 * nthreads) >= 1 ? n / (64 * nthreads) : 1);
  #pragma omp parallel num_threads(nthreads)
  {
    wavefront_aligner_t* const wf = new_aligner(pen2, two_piece, score_only, max_steps);
    affine_penalties_t pen = {.match = pen2->match, .mismatch = pen2->mismatch,
                              .gap_opening = pen2->gap_opening1, .gap_extension = pen2->gap_extension1};
    #pragma omp for schedule(dynamic, chunk)
    for (int i = 0; i < n; ++i) {
      const int st = wavefront_align(wf, seqbuf + p_off[i], p_len[i],
                                     seqbuf + t_off[i], t_len[i]);
      cigar_t* const cg = wf->cigar;
      status[i] = st;
      raw_score[i] = cg->score;
      const int empty = (cg->begin_offset >= cg->end_offset);
      if (score_only) {
        alg_score[i] = cg->score;
      } else {
        alg_score[i] = empty ? -1 : two_piece ? cigar_score_gap_affine2p(cg, pen2)
                                              : cigar_score_gap_affine(cg, &pen);
      }
      if (runs != NULL) {
        int len = 0;
        uint32_t* buf = NULL;
        if (!score_only && !empty) cigar_get_CIGAR(cg, g_show_mismatches != 0, &buf, &len);
        const int64_t cap = runs_off[i + 1] - runs_off[i];
        if (len > cap) {
          nruns[i] = -1;
        } else {
          if (len > 0) memcpy(runs + runs_off[i], buf, sizeof(uint32_t) * (size_t)len);
          nruns[i] = len;
        }
      }
    }
    wavefront_aligner_delete(wf);
  }
  if (elapsed) *elapsed = now_seconds() - t0;
}

int wfa2ref_align_batch(
    int n, const char* seqbuf,
    const int64_t* p_off, const int32_t* p_len,
    const int64_t* t_off, const int32_t* t_len,
    int match, int x, int o, int e, int score_only, int max_steps, int nthreads,
    int32_t* alg_score, int32_t* raw_score, int32_t* status,
    uint32_t* runs, const int64_t* runs_off, int32_t* nruns,
    double* elapsed) {
  const affine2p_penalties_t p = {.match = match, .mismatch = x, .gap_opening1 = o,
                                  .gap_extension1 = e, .gap_opening2 = 0, .gap_extension2 = 0};
  align_batch(n, seqbuf, p_off, p_len, t_off, t_len, &p, 0, score_only, max_steps, nthreads,
              alg_score, raw_score, status, runs, runs_off, nruns, elapsed);
  return 0;
}

/* The same under gap_affine_2p (penalties match, x, o1, e1, o2, e2). */
int wfa2ref_align_batch_2p(
    int n, const char* seqbuf,
    const int64_t* p_off, const int32_t* p_len,
    const int64_t* t_off, const int32_t* t_len,
    int match, int x, int o1, int e1, int o2, int e2, int score_only, int max_steps,
    int nthreads, int32_t* alg_score, int32_t* raw_score, int32_t* status,
    uint32_t* runs, const int64_t* runs_off, int32_t* nruns, double* elapsed) {
  const affine2p_penalties_t p = {.match = match, .mismatch = x, .gap_opening1 = o1,
                                  .gap_extension1 = e1, .gap_opening2 = o2, .gap_extension2 = e2};
  align_batch(n, seqbuf, p_off, p_len, t_off, t_len, &p, 1, score_only, max_steps, nthreads,
              alg_score, raw_score, status, runs, runs_off, nruns, elapsed);
  return 0;
}

int wfa2ref_max_threads(void) { return omp_get_max_threads(); }
