/*
 * biwfa_mi355x.h -- C ABI of libbiwfa_mi355x.so, the MI355X (gfx950) bidirectional
 * wavefront aligner.  Plain C types only; every pointer is host memory unless
 * the function says otherwise.
 *
 * Results are bit-exact with WFA2-lib v2.3 `wavefront_align` configured for
 * exact BiWFA: gap_affine (or gap_affine_2p) penalties, memory_mode = ultralow,
 * heuristic = none
 * (score from the CIGAR as align_benchmark writes it, CIGAR operations byte for
 * byte).  Paths below are relative to the reference checkout.
 *
 * Three layers:
 *   (i)   WFA2-style aligner object   -- bw_aligner_new / bw_align / bw_align_packed2bits /
 *                                        bw_aligner_delete + result accessors;
 *   (ii)  repo-style align-a-pair     -- bw_align_pair;
 *   (iii) batch                       -- bw_align_batch (host buffers in/out) and the
 *                                        device-resident bw_dbatch_* calls (inputs and
 *                                        outputs stay in HBM; what bench.py times).
 * Plus the seeded generator / 2-bit packer that feed the batch path.
 */
#ifndef BIWFA_MI355X_H_
#define BIWFA_MI355X_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: WFA2-lib wavefront/wfa.h:46-55, plus device-side codes ---- */
#define BW_STATUS_ALG_COMPLETED      0
#define BW_STATUS_ALG_PARTIAL        1
#define BW_STATUS_MAX_STEPS_REACHED  -100
#define BW_STATUS_OOM                -200
#define BW_STATUS_UNATTAINABLE       -300
#define BW_STATUS_DEVICE_CAPACITY    -400  /* a per-pair device limit was exceeded      */
#define BW_STATUS_HIP_ERROR          -401  /* a HIP runtime call failed                 */
#define BW_STATUS_INVALID_ARG        -402  /* bad penalties / lengths / pointers        */
#define BW_STATUS_NO_DEVICE          -403  /* no usable gfx950 device                   */
#define BW_STATUS_BAD_SEQUENCE       -404  /* byte outside ACGT in a 2-bit packed path  */
#define BW_STATUS_INTERNAL           -405  /* device termination guard tripped          */

/* CIGAR runs use WFA2's SAM encoding with mismatches shown
 * (cigar_get_CIGAR(show_mismatches=true), alignment/cigar.c:181-240):
 * (length << 4) | op, op = 7 '=' (match, printed 'M'), 8 'X', 1 'I', 2 'D'. */
#define BW_OP_MATCH 7
#define BW_OP_MISMATCH 8
#define BW_OP_INSERTION 1
#define BW_OP_DELETION 2

typedef struct {
  int match;          /* <= 0 (match < 0 uses WFA2's Eizenga transform) */
  int mismatch;       /* > 0  */
  int gap_opening;    /* >= 0 */
  int gap_extension;  /* > 0  */
} bw_affine_penalties_t;                 /* alignment/affine_penalties.h */

typedef struct {
  int match;           /* <= 0 (match < 0 uses WFA2's Eizenga transform) */
  int mismatch;        /* > 0  */
  int gap_opening1;    /* >= 0 */
  int gap_extension1;  /* > 0  */
  int gap_opening2;    /* >= 0 */
  int gap_extension2;  /* > 0  */
} bw_affine2p_penalties_t;               /* alignment/affine2p_penalties.h */

/* Distance metric (wavefront/wavefront_penalties.h:41-48 distance_metric_t: gap_affine,
 * gap_affine_2p).  Gap-affine-2p scores a gap run of length n as
 * min(o1 + e1*n, o2 + e2*n) and runs on its own single-wave kernel; the tiers,
 * the learned schedule, autotuning and the no-allocation entry point are for
 * gap-affine and return BW_STATUS_INVALID_ARG for a gap-affine-2p aligner. */
typedef enum {
  BW_GAP_AFFINE = 0,
  BW_GAP_AFFINE_2P = 1
} bw_distance_metric_t;

typedef enum {
  BW_SCOPE_ALIGNMENT = 0,                /* score + CIGAR (compute_alignment) */
  BW_SCOPE_SCORE = 1                     /* score only    (compute_score)     */
} bw_scope_t;

typedef enum {
  BW_SCHEDULE_INPUT_ORDER = 0,           /* pairs taken in input order                     */
  BW_SCHEDULE_LENGTH_SORTED = 1,         /* plain: longest (p+t) first                     */
  BW_SCHEDULE_LEARNED = 2                /* tiers and order from the cost model's predicted work */
} bw_schedule_t;

typedef struct {
  bw_affine_penalties_t penalties;       /* default {0,4,6,2}                          */
  int scope;                             /* bw_scope_t                                 */
  int max_alignment_steps;               /* <= 0: unlimited (wavefront_attributes.c:89) */
  int device;                            /* HIP device ordinal                         */
  int schedule;                          /* bw_schedule_t                              */
  int waves_per_cu;                      /* LDS tier only: cap on resident waves per CU
                                            (0: as many as its occupancy allows)       */
  int queue_chunk;                       /* LDS tier pairs per queue grab when the
                                            tuning's queue_chunk[0] is 0 (0: 2)        */
  int distance_metric;                   /* bw_distance_metric_t (default BW_GAP_AFFINE) */
  bw_affine2p_penalties_t affine2p_penalties;   /* BW_GAP_AFFINE_2P: default {0,4,6,2,24,1} */
} bw_attr_t;                             /* wavefront_attributes.h wavefront_aligner_attr_t */

typedef struct bw_aligner bw_aligner_t;

/* Fills *attr with the defaults above. */
void bw_attr_default(bw_attr_t* attr);

/* ---------------- (i) WFA2-style aligner ---------------- */

/* Replaces wavefront_aligner_new (wavefront/wavefront_aligner.c:421-463).
 * Returns NULL on invalid attributes or when no gfx950 device is usable
 * (bw_last_error() tells which). */
bw_aligner_t* bw_aligner_new(const bw_attr_t* attr);

/* Replaces wavefront_aligner_delete (wavefront/wavefront_aligner.c:476-501). */
void bw_aligner_delete(bw_aligner_t* aligner);

/* Replaces wavefront_align (wavefront/wavefront_align.c:212-241).  Sequences are
 * ASCII ACGT (either case).  Returns the status (BW_STATUS_*). */
int bw_align(bw_aligner_t* aligner, const char* pattern, int pattern_length,
             const char* text, int text_length);

/* Replaces wavefront_align_packed2bits (wavefront/wavefront_align.c:182-211):
 * 4 bases per byte, LSB first, A=0 C=1 G=2 T=3 (wavefront_sequences.c:102-140).
 * All ceil(len/4) bytes are decoded. */
int bw_align_packed2bits(bw_aligner_t* aligner, const uint8_t* pattern, int pattern_length,
                         const uint8_t* text, int text_length);

/* Score of the last alignment, as align_benchmark reports it
 * (tools/align_benchmark/benchmark/benchmark_utils.c:143-179): the gap-affine
 * score of the CIGAR (<= 0), or the score-only result in BW_SCOPE_SCORE. */
int bw_get_score(const bw_aligner_t* aligner);

/* Internal WFA score (penalty units, >= 0) of the last alignment. */
int bw_get_wf_score(const bw_aligner_t* aligner);

/* Status of the last alignment. */
int bw_get_status(const bw_aligner_t* aligner);

/* CIGAR operations of the last alignment, one char per column ('M','X','I','D'),
 * like cigar->operations[begin_offset,end_offset).  Valid until the next call. */
int bw_get_cigar_ops(const bw_aligner_t* aligner, const char** ops, int* length);

/* SAM-packed CIGAR, cigar_get_CIGAR (alignment/cigar.c:181-240).  show_mismatches=1
 * gives the run encoding above; 0 folds X into M (op 0) from the second operation
 * on, as cigar_compute_CIGAR does (a leading mismatch stays a 1X run).  Valid until
 * the next call. */
int bw_get_sam_cigar(bw_aligner_t* aligner, int show_mismatches,
                     const uint32_t** cigar_buffer, int* cigar_length);

/* Run-length text, cigar_sprint(print_matches=true) (alignment/cigar.c:706-740),
 * e.g. "14M1I86M".  Returns the length written (excluding NUL) or -1 if cap is too small. */
int bw_cigar_sprint(const bw_aligner_t* aligner, char* buffer, int capacity);

/* Message for the last error on this thread. */
const char* bw_last_error(void);

/* Text of a status code (wavefront_align_strerror, wavefront_aligner.c:66-83). */
const char* bw_strerror(int status);

/* ---------------- (ii) repo-style align-a-pair ---------------- */

/* The repo's per-pair entry (src/biWFA.cu:607 kernel + :756 driver, penalties
 * x/o/e from src/headers/biWFA.h:10-12, match 0), extended to return the CIGAR.
 * *score receives the (<= 0) score; ops receives 'M','X','I','D' (no NUL);
 * *ops_length its length.  Uses a per-thread default aligner on device 0. */
int bw_align_pair(const char* pattern, int pattern_length, const char* text, int text_length,
                  int x, int o, int e, int32_t* score, char* ops, int ops_capacity,
                  int* ops_length);

/* ---------------- (iii) batch ---------------- */

/* A batch of 2-bit packed pairs.  Each sequence starts on a 64-bit word boundary:
 * base j of a sequence at bits 2*(j%32) of word (word_offset + j/32).  The word
 * array needs one readable word before the first sequence and two after the last
 * (bw_batch_words_needed accounts for them; the library never reads them as bases). */
typedef struct {
  int64_t num_pairs;
  int64_t num_words;
  const uint64_t* words;
  const int64_t* pattern_word_offset;
  const int32_t* pattern_length;
  const int64_t* text_word_offset;
  const int32_t* text_length;
} bw_batch_t;

/* Per-pair results.  cigar_offset[i]..cigar_offset[i+1] delimits pair i's runs
 * in cigar_runs (compacted, input order).  Any pointer may be NULL to skip it. */
typedef struct {
  int32_t* score;           /* [num_pairs] .alg score              */
  int32_t* wf_score;        /* [num_pairs] internal WFA score      */
  int32_t* status;          /* [num_pairs]                         */
  int64_t* cigar_offset;    /* [num_pairs + 1]                     */
  uint32_t* cigar_runs;     /* [cigar_capacity]                    */
  int64_t cigar_capacity;
  int64_t cigar_total;      /* out: runs written                   */
} bw_result_t;

/* Aligns a host batch: upload, one persistent launch (plus a spill launch when
 * needed), compaction, download.  Returns 0 or a BW_STATUS_* error. */
int bw_align_batch(bw_aligner_t* aligner, const bw_batch_t* batch, bw_result_t* result);

/* Device-resident batch: inputs uploaded once, results kept in HBM. */
typedef struct bw_dbatch bw_dbatch_t;
bw_dbatch_t* bw_dbatch_upload(bw_aligner_t* aligner, const bw_batch_t* batch);
/* Enqueues the alignment of every pair on the aligner's stream (no host sync).
 * Scheduling order comes from the aligner's schedule (set by bw_set_schedule). */
int bw_dbatch_align(bw_aligner_t* aligner, bw_dbatch_t* dbatch);
/* Waits for the aligner's stream. */
int bw_synchronize(bw_aligner_t* aligner);
/* Copies results to host buffers (compacting the runs). */
int bw_dbatch_download(bw_aligner_t* aligner, bw_dbatch_t* dbatch, bw_result_t* result);
void bw_dbatch_free(bw_dbatch_t* dbatch);

/* Milliseconds of the last bw_dbatch_align, measured with HIP events on the
 * aligner's stream (valid after bw_synchronize): *total_ms the whole launch
 * sequence, *main_ms the LDS tier (every pair that fits it), *spill_ms the
 * global tiers, run in order of ring width over their own pairs and the pairs
 * the tier below passed on. */
int bw_last_timing(const bw_aligner_t* aligner, float* total_ms, float* main_ms, float* spill_ms);

/* ---------------- (iv) scheduling ---------------- */

/* Selects the pair order and starting tier of later bw_dbatch_align calls
 * (bw_schedule_t).  For BW_SCHEDULE_LEARNED, `predicted_score` holds one
 * predicted WFA score (penalty units, >= 0) per pair (host): pairs start in the
 * first tier whose ring is wide enough for it; the single-wave tiers take them
 * by predicted work, largest first, and the team tiers longest first.  The
 * schedule never changes any pair's result. */
int bw_set_schedule(bw_aligner_t* aligner, bw_dbatch_t* dbatch, int schedule,
                    const float* predicted_score);

/* ---------------- launch tuning ---------------- */

/* No counterpart in the reference (its kernel runs one pair per 32-thread block).
 * A batch runs through tiers of growing wavefront-ring width: tier 0 keeps the
 * whole workspace in LDS with one wave per pair; global tier i (i = 0, 1, ...)
 * has a ring 256 * 4^i diagonals wide in HBM (the last one as wide as the batch
 * needs) and aligns each pair with a team of team_waves[i] waves (64 lanes each)
 * of one workgroup.  A pair that outgrows a tier moves to the next one on the
 * device.  Zero fields take the built-in rule.  Tuning changes speed only. */
#define BW_MAX_TIERS 8
typedef struct {
  int lds_ring_width;                   /* tier 0 ring width: 16, 32, 64 or 128 (0: 32)      */
  int lds_seq_cap;                      /* tier 0 longest staged sequence, <= 4096 (0: 512) */
  int team_waves[BW_MAX_TIERS];         /* global tier i: 1, 2, 4, 8 or 16 waves per pair    */
  int queue_chunk[BW_MAX_TIERS];        /* tier i (0 = LDS tier): pairs per queue grab       */
} bw_tuning_t;

/* Tuning used by later uploads (bw_dbatch_upload / bw_align_batch plan their tiers
 * with it).  bw_get_tuning returns the current one (zeros = built-in rules). */
int bw_set_tuning(bw_aligner_t* aligner, const bw_tuning_t* tuning);
int bw_get_tuning(const bw_aligner_t* aligner, bw_tuning_t* tuning);

/* Per-tier milliseconds of the last bw_dbatch_align (valid after bw_synchronize):
 * ms[0] the LDS tier, ms[1 + i] global tier i.  Returns the number of tiers. */
int bw_last_tier_timing(const bw_aligner_t* aligner, float* ms, int capacity);

/* Pairs each tier took on in the last bw_dbatch_align (its initial share plus the
 * pairs the previous tier passed on); the sum minus the batch size counts re-runs.
 * Valid after bw_synchronize.  Returns the number of tiers. */
int bw_last_tier_pairs(const bw_aligner_t* aligner, int32_t* pairs, int capacity);

/* Launch shape of one tier of the last bw_dbatch_align (no counterpart in the
 * reference; for reports that name the kernel a tier ran).  `kernel` is the
 * instantiation's name as profilers print it. */
typedef struct {
  int lds;                 /* 1: the LDS tier (workspace in LDS), 0: a global tier */
  int waves_per_aligner;   /* team size: waves that align one pair together */
  int ring_width;          /* diagonals per ring slot */
  int blocks;              /* workgroups launched */
  int waves_per_block;
  int64_t aligners;        /* aligners resident at once */
  char kernel[96];
} bw_tier_shape_t;
int bw_last_tier_shape(const bw_aligner_t* aligner, int tier, bw_tier_shape_t* shape);

/* Times the sample batch under candidate tunings and keeps the fastest: team
 * waves per global tier (each tier timed on its own), then the LDS tier's queue
 * chunk and ring width (whole batch).  The sample runs under the aligner's
 * schedule, or the learned one when `predicted_score` (one per sample pair) is
 * given.  `repeats` timed runs per candidate (after one warm-up).  The result is
 * set on the aligner and copied to *best. */
int bw_autotune(bw_aligner_t* aligner, const bw_batch_t* sample, const float* predicted_score,
                int repeats, bw_tuning_t* best);

/* Optional progress reports of long calls (bw_autotune: one line per timed run
 * and one per candidate).  fn may be NULL to switch them off. */
typedef void (*bw_progress_fn)(const char* message, void* user);
int bw_set_progress(bw_aligner_t* aligner, bw_progress_fn fn, void* user);

/* Graph mode (enable != 0).  The first bw_dbatch_align of a device batch on this
 * aligner then captures the call's launch sequence -- the counter reset and one
 * persistent launch per tier; the recursion itself runs on the device -- into a
 * HIP graph, and later calls on the same batch replay it as one launch.  Results
 * are identical.  bw_last_tier_timing reports zeros after a replay (no events
 * between the tiers); bw_last_timing reports the whole sequence as the main part.
 * Replaces the reference's host loop of one launch per 100-pair chunk
 * (src/biWFA.cu:935-1004). */
int bw_set_graph(bw_aligner_t* aligner, int enable);

/* ---------------- failure-path testing ---------------- */

/* No counterpart in the reference.  For tests of the termination guard only:
 * `ready_ticks` (> 0; 0 restores the default of 60 s) bounds how long an
 * aligner of a global tier waits, in ticks of the 100 MHz device clock, for the
 * ready flag of a shared recursion task it has claimed; `withhold_ready` != 0
 * publishes such tasks without their ready flag, so every wait for one gives
 * up.  A wait that gives up raises the tier's error flag, every aligner of the
 * tier then leaves its work loop, and bw_synchronize returns
 * BW_STATUS_INTERNAL.  Affects later bw_dbatch_align / bw_align_batch calls. */
int bw_test_set_ready_guard(bw_aligner_t* aligner, int64_t ready_ticks, int withhold_ready);

/* ---------------- no device allocation ---------------- */

/* Per-pair results of bw_align_device_noalloc, all in device memory and owned by
 * the caller.  Pair i's runs (encoding above, in order) are written to
 * runs[run_offset[i] ...]; run_offset[i+1] - run_offset[i] must be at least
 * p_i + t_i + 1.  num_runs[i] receives their count. */
typedef struct {
  int32_t* score;
  int32_t* wf_score;
  int32_t* status;
  int32_t* num_runs;
  uint32_t* runs;
  const int64_t* run_offset;   /* [num_pairs + 1] */
} bw_device_result_t;

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the variant of bw_dbatch_align that calls no hipMalloc,
 * hipFree or allocating library routine.  `batch` holds DEVICE pointers (same
 * layout as for bw_dbatch_upload); the results go straight to the caller's
 * device buffers.  Everything else lives in a 64 MiB per-GPU scratch array that
 * the library holds until the process ends.  Limits: at most 1,048,576 pairs
 * per call (more: BW_STATUS_INVALID_ARG, nothing enqueued); pairs with
 * p + t <= 2040 are aligned exactly as bw_dbatch_align aligns them (LDS tier,
 * then a single-wave global tier), longer ones get status
 * BW_STATUS_DEVICE_CAPACITY and malformed ones BW_STATUS_INVALID_ARG.
 * Enqueued on the aligner's stream; call bw_synchronize before reading the
 * results.  Every call on one GPU shares its scratch: calls that overlap on one
 * GPU (other streams or threads) race. */
int bw_align_device_noalloc(bw_aligner_t* aligner, const bw_batch_t* batch, const bw_device_result_t* result);

/* ---------------- (v) learned cost model ---------------- */

/* No counterpart in the reference: a 4 -> 64 -> 64 -> 1 ReLU MLP (BF16 weights,
 * FP32 accumulation) that predicts a pair's alignment score from
 *   log2(p+1)/16, log2(t+1)/16, log2(|p-t|+1)/16, fraction of 16 sampled 12-mers
 * of the pattern found near their expected text position;
 * target log2(1 + score)/16.  Trained and evaluated by the library's own HIP
 * kernels; its predictions feed bw_set_schedule(BW_SCHEDULE_LEARNED). */
typedef struct bw_mlp bw_mlp_t;
int bw_mlp_num_params(void);
bw_mlp_t* bw_mlp_new(int device, uint64_t seed);
void bw_mlp_free(bw_mlp_t* mlp);
const char* bw_mlp_last_error(void);
/* FP32 master parameters and the BF16 weights (stored as 16-bit on the device;
 * returned widened to FP32, exactly).  Either pointer may be NULL. */
int bw_mlp_get_params(bw_mlp_t* mlp, float* master, float* bf16_weights);
int bw_mlp_set_params(bw_mlp_t* mlp, const float* master);
/* X: n x 4 features (host), y: n outputs (host). */
int bw_mlp_predict(bw_mlp_t* mlp, const float* X, int64_t n, float* y);
/* Gradient of the mean squared error w.r.t. the BF16 weights (host output). */
int bw_mlp_gradient(bw_mlp_t* mlp, const float* X, const float* y, int64_t n, float* grad);
/* Mini-batch Adam on the device; *final_loss = full-set MSE afterwards. */
int bw_mlp_train(bw_mlp_t* mlp, const float* X, const float* y, int64_t n, int epochs, int batch,
                 float lr, uint64_t seed, float* final_loss);
/* The four features of every pair of a packed batch, computed on the device. */
int bw_batch_features(int device, const bw_batch_t* batch, float* X);

/* ---------------- generator / packing ---------------- */

typedef struct {
  int64_t num_pairs;
  uint64_t seed;
  int32_t len_min, len_max;      /* pattern length; log-uniform when len_max > len_min */
  int32_t num_edits;             /* >= 0: fixed number of edit events per pair          */
  double rate_min, rate_max;     /* else round(U(rate_min, rate_max) * plen) events     */
  int32_t geometric_runs;        /* single-base I/D events get Geometric(0.5) lengths   */
  double run_frac;               /* probability that an event is an I/D run             */
  int32_t run_min, run_max;      /* run length U{run_min..run_max}                      */
} bw_dataset_params_t;

/* Pattern/text lengths of every pair (deterministic in seed and pair index). */
int bw_dataset_lengths(const bw_dataset_params_t* params, int32_t* pattern_length,
                       int32_t* text_length);
/* ASCII sequences written at the given offsets of `seq`. */
int bw_dataset_generate(const bw_dataset_params_t* params, const int64_t* pattern_offset,
                        const int64_t* text_offset, char* seq);
/* Packs ASCII sequences into 2-bit words at word offsets (see bw_batch_t). */
int bw_pack_2bit(int64_t num_seqs, const char* ascii, const int64_t* offset,
                 const int32_t* length, const int64_t* word_offset, uint64_t* words);

/* ---------------- sequence files and .alg output ---------------- */

#define BW_SEQ_FORMAT_AUTO 0   /* WFA2 when the file starts with '>' or '<', else the repo's */
#define BW_SEQ_FORMAT_WFA2 1   /* .seq: pattern line, text line, each behind a one-char marker */
#define BW_SEQ_FORMAT_REPO 2   /* header "N P T", then N pattern / text line pairs            */

typedef struct bw_seqs bw_seqs_t;
/* Reads every pair of a sequence file.  Replaces align_benchmark's input loop
 * (align_benchmark_read_input, tools/align_benchmark/align_benchmark.c:313-338;
 * the files generate_dataset writes, generate_dataset.c:387-393) and the repo
 * driver's reader (src/biWFA.cu:776-835).  NULL on error (bw_last_error). */
bw_seqs_t* bw_seqs_read(const char* path, int format);
int64_t bw_seqs_count(const bw_seqs_t* seqs);
/* Pair i's ASCII sequences (owned by seqs). */
int bw_seqs_get(const bw_seqs_t* seqs, int64_t i, const char** pattern, int* pattern_length,
                const char** text, int* text_length);
/* The pairs as a 2-bit packed batch whose arrays belong to seqs (valid until
 * bw_seqs_free or the next bw_seqs_batch); BW_STATUS_BAD_SEQUENCE for a byte
 * outside ACGT. */
int bw_seqs_batch(bw_seqs_t* seqs, bw_batch_t* batch);
void bw_seqs_free(bw_seqs_t* seqs);
/* Writes align_benchmark's lite .alg file, one "score<TAB>CIGAR" line per pair
 * (benchmark_print_output_lite / benchmark_print_output,
 * tools/align_benchmark/benchmark/benchmark_utils.c:102-117, 143-179): "-" for an
 * empty CIGAR, and -1 as the score of a full alignment whose CIGAR is empty.
 * With score_only, cigar_offset and cigar_runs are not read. */
int bw_alg_write(const char* path, int64_t num_pairs, const int32_t* score, const int64_t* cigar_offset,
                 const uint32_t* cigar_runs, int score_only);

#ifdef __cplusplus
}
#endif
#endif /* BIWFA_MI355X_H_ */
