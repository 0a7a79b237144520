/*
 * biwfa_oracle.h -- CPU restatement of WFA2-lib v2.3's bidirectional WFA
 * (memory_mode = ultralow, heuristic = none, gap-affine and gap-affine-2p),
 * score + CIGAR.
 *
 * TEST INFRASTRUCTURE ONLY.  Used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker; the product library never links it.
 */
#ifndef BIWFA_ORACLE_H_
#define BIWFA_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (WFA2-lib wavefront/wfa.h:46-55). */
#define BWO_STATUS_COMPLETED     0
#define BWO_STATUS_MAX_STEPS  -100
#define BWO_STATUS_OOM        -200
#define BWO_STATUS_UNATTAIN   -300
#define BWO_STATUS_CAPACITY   -400   /* run buffer too small (oracle-only) */

/* Algorithmic work counters (per call, accumulated). */
typedef struct {
  int64_t cells;        /* sum over compute steps of (hi-lo+1) from the input limits */
  int64_t words;        /* sum over extended non-null diagonals of (run/32 + 1)       */
  int64_t compute_steps;
  int64_t breakpoints;  /* breakpoint searches run                                    */
  int64_t base_cases;   /* unidirectional base cases run                              */
  int64_t max_level;    /* deepest recursion level reached                            */
} bwo_counters_t;

/*
 * Align one pair.  Penalties are the user penalties (match <= 0, x > 0, o >= 0, e > 0).
 * runs: uint32 (len<<4)|op with op '='=7, X=8, I=1, D=2 (WFA2 cigar_get_CIGAR
 * with show_mismatches=true).  *alg_score receives what align_benchmark prints
 * in the .alg score column; *wf_score the internal WFA score (penalty units).
 * Returns the status code.
 */
int bwo_align(const char* pattern, int plen, const char* text, int tlen,
              int match, int x, int o, int e, int score_only, int max_steps,
              uint32_t* runs, int runs_cap, int* nruns,
              int* alg_score, int* wf_score, bwo_counters_t* counters);

/* Batch over pairs with OpenMP (dynamic,16); returns wall seconds in *elapsed. */
int bwo_align_batch(int n, const char* seqbuf,
                    const int64_t* p_off, const int32_t* p_len,
                    const int64_t* t_off, const int32_t* t_len,
                    int match, int x, int o, int e, int score_only, int max_steps,
                    int nthreads,
                    int32_t* alg_score, int32_t* wf_score, int32_t* status,
                    uint32_t* runs, const int64_t* runs_off, int32_t* nruns,
                    int64_t* cells, int64_t* words, double* elapsed);

/* The same under gap-affine-2p penalties (match, x, o1, e1, o2, e2): gap runs
 * cost min(o1 + e1*len, o2 + e2*len) (WFA2 gap_affine_2p). */
int bwo_align_batch_2p(int n, const char* seqbuf,
                       const int64_t* p_off, const int32_t* p_len,
                       const int64_t* t_off, const int32_t* t_len,
                       int match, int x, int o1, int e1, int o2, int e2,
                       int score_only, int max_steps, int nthreads,
                       int32_t* alg_score, int32_t* wf_score, int32_t* status,
                       uint32_t* runs, const int64_t* runs_off, int32_t* nruns,
                       int64_t* cells, int64_t* words, double* elapsed);

#ifdef __cplusplus
}
#endif
#endif
