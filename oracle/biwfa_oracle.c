/*
 * biwfa_oracle.c -- CPU restatement of WFA2-lib v2.3's bidirectional WFA
 * (BiWFA: memory_mode = ultralow, heuristic = none, gap-affine, end-to-end),
 * score + CIGAR, plain C, one pair at a time.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this file's library, and only as the
 * checker / the timed CPU baseline.  The product path never calls it.
 *
 * Every function cites the reference code it restates (paths relative to the
 * WFA2-lib directory of the reference checkout).  Semantics that matter for
 * bit-exactness and are easy to get wrong are spelled out in comments:
 *   - an offset read outside a wavefront's [lo,hi] is OFFSET_NULL, and a
 *     missing or empty input wavefront counts as the shared null wavefront
 *     whose limits are lo=1, hi=-1 (wavefront.c:110-138, wavefront_compute.c:266-305);
 *   - M offsets are either OFFSET_NULL or in bounds; I1/D1 offsets inside
 *     [lo,hi] keep whatever the recurrence produced, including out-of-bounds
 *     values (wavefront_compute_affine.c:65-85, wavefront_compute.c:579-613);
 *   - the breakpoint scans accept only strictly better scores, check D1 then
 *     I1 then M, and scan k in ascending order (wavefront_bialign.c:189-392);
 *   - backtrace ties resolve M > D-ext > D-open > I-ext > I-open through the
 *     packed (offset<<4)|type maximum (wavefront_backtrace.c:38-59,369-378).
 *
 * Pinned by tests/test_oracle.py against: the 305-pair WFA2 golden .alg files
 * (7 penalty sets, full and score-only), the vendored WFA2-lib built from
 * source on seeded random pairs, a textbook Gotoh DP on small pairs, and
 * CIGAR validity / rescoring invariants.
 */
#include "biwfa_oracle.h"

#include <limits.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

#define OFFSET_NULL (INT32_MIN / 2)          /* wavefront_offset.h:44 */
#define FALLBACK_MIN_SCORE 250               /* wavefront_bialign.c:48 */
#define FALLBACK_MIN_LENGTH 100              /* wavefront_bialign.c:49 */
#define RECOVERY_MIN_SCORE 500               /* wavefront_bialign.c:50 */

/* Internal status codes (wfa.h:52-55). */
#define ST_OK -1
#define ST_END_REACHED -2
#define ST_END_UNREACHABLE -3

/* components M, I1, D1, and for gap-affine-2p also I2, D2 (affine2p_matrix_type,
 * affine2p_penalties.h) */
enum { CM = 0, CI = 1, CD = 2, CI2 = 3, CD2 = 4 };
#define NCOMP 5

typedef struct {
  int present;     /* WFA2 keeps a non-NULL pointer for this component */
  int lo, hi;      /* trimmed limits (lo > hi: null wavefront)          */
  int clo;         /* first diagonal stored                             */
  int32_t* off;    /* off[k - clo]                                      */
  int cap;
} wf_t;

typedef struct { wf_t c[NCOMP]; } wfset_t;

typedef struct {
  int x, o, e;                  /* internal penalties */
  int o2, e2, a2p;              /* gap-affine-2p: second gap model (a2p = 1) */
  int ncomp;                    /* 3, or 5 for gap-affine-2p */
  const char* P; const char* T; /* full sequences */
  int pb, plen, tb, tlen;       /* current sub-range */
  int reverse;                  /* aligner walks the sub-range backwards */
  int modular, scope;           /* ring of `scope` scores, or all scores */
  wfset_t* wfs; int nwfs;
  int comp_begin, comp_end;
  int null_steps;
  int status, status_score;
  int end_k, end_off;
  bwo_counters_t* cnt;
} aligner_t;

typedef struct {
  int score, score_forward, score_reverse;
  int k_forward, k_reverse, offset_forward, offset_reverse;
  int component;
} breakpoint_t;                 /* wavefront_bialigner.h:43-54 */

typedef struct {
  aligner_t fwd, rev, base;
  char* ops; int nops, ops_cap;      /* master CIGAR operations, forward order */
  char* rops; int rops_cap;          /* base-case backtrace, reverse order      */
  int match, max_steps;
  int level0_wf_score;
  int err;
  bwo_counters_t cnt;
} ctx_t;

/* ------------------------------------------------------------------------- */
/* storage                                                                   */

static void wf_reserve(wf_t* w, int lo, int hi) {
  const int n = hi - lo + 1;
  if (n > w->cap) {
    int cap = w->cap ? w->cap : 64;
    while (cap < n) cap *= 2;
    free(w->off);
    w->off = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    w->cap = cap;
  }
  w->clo = lo;
  w->lo = lo;
  w->hi = hi;
  w->present = 1;
}

static wfset_t* slot_of(aligner_t* a, int score) {
  const int idx = a->modular ? score % a->scope : score;
  if (idx >= a->nwfs) {
    int n = a->nwfs ? a->nwfs : 64;
    while (n <= idx) n *= 2;
    a->wfs = (wfset_t*)realloc(a->wfs, sizeof(wfset_t) * (size_t)n);
    memset(a->wfs + a->nwfs, 0, sizeof(wfset_t) * (size_t)(n - a->nwfs));
    a->nwfs = n;
  }
  return &a->wfs[idx];
}

/* Offset of a wavefront at diagonal k; OFFSET_NULL outside [lo,hi]. */
static inline int32_t wf_get(const wf_t* w, int k) {
  return (w != NULL && w->present && w->lo <= k && k <= w->hi) ? w->off[k - w->clo] : OFFSET_NULL;
}

/* Input wavefront at `score` (wavefront_compute.c:266-305,306-352): NULL stands
 * for the shared null wavefront (missing slot, negative score, or lo > hi). */
static const wf_t* fetch_input(aligner_t* a, int score, int comp) {
  if (score < 0) return NULL;
  const int idx = a->modular ? score % a->scope : score;
  if (idx >= a->nwfs) return NULL;
  const wf_t* w = &a->wfs[idx].c[comp];
  if (!w->present || w->lo > w->hi) return NULL;
  return w;
}

/* wavefront_aligner_init + wavefront_aligner_init_wf (wavefront_aligner.c:311-417):
 * clear every wavefront and seed score 0 in the begin component at k=0. */
static void aligner_init(aligner_t* a, int comp_begin, int comp_end) {
  for (int i = 0; i < a->nwfs; ++i)
    for (int c = 0; c < NCOMP; ++c) a->wfs[i].c[c].present = 0;
  a->comp_begin = comp_begin;
  a->comp_end = comp_end;
  a->null_steps = 0;
  a->status = ST_OK;
  a->status_score = 0;
  a->end_k = INT_MAX;
  a->end_off = OFFSET_NULL;
  wfset_t* s0 = slot_of(a, 0);
  wf_t* w = &s0->c[comp_begin];
  wf_reserve(w, 0, 0);
  w->off[0] = 0;
}

/* wavefront_sequences_set_bounds (wavefront_sequences.c:274-309): the forward
 * and base aligners read P[pb+v], T[tb+h]; the reverse aligner reads the
 * reversed sub-range, P[pe-1-v], T[te-1-h]. */
static void aligner_bounds(aligner_t* a, int pb, int pe, int tb, int te) {
  a->pb = pb; a->plen = pe - pb;
  a->tb = tb; a->tlen = te - tb;
}

/* ------------------------------------------------------------------------- */
/* compute (wavefront_compute_affine.c:44-86,229-260; wavefront_compute.c)   */

static void trim_ends(aligner_t* a, wf_t* w) {       /* wavefront_compute.c:579-613 */
  const uint32_t tl = (uint32_t)a->tlen, pl = (uint32_t)a->plen;
  int k;
  for (k = w->hi; k >= w->lo; --k) {
    const int32_t o = w->off[k - w->clo];
    if ((uint32_t)o <= tl && (uint32_t)(o - k) <= pl) break;
  }
  w->hi = k;
  for (k = w->lo; k <= w->hi; ++k) {
    const int32_t o = w->off[k - w->clo];
    if ((uint32_t)o <= tl && (uint32_t)(o - k) <= pl) break;
  }
  w->lo = k;
}

/* wavefront_compute_affine2p (wavefront_compute_affine2p.c:45-108,334-367):
 * seven inputs M[s-x], M[s-o1-e1], M[s-o2-e2], I1/D1[s-e1], I2/D2[s-e2]; the
 * limits of wavefront_compute_limits_input (wavefront_compute.c:40-86) over all
 * seven; I2/D2 are allocated when one of their inputs exists
 * (wavefront_compute.c:474-493); M = max(del1, del2, misms, ins1, ins2), nulled
 * when out of bounds; every output trimmed (wavefront_compute.c:615-632). */
static void compute_affine2p(aligner_t* a, int s, wfset_t* out) {
  const wf_t* mm = fetch_input(a, s - a->x, CM);
  const wf_t* mo = fetch_input(a, s - a->o - a->e, CM);
  const wf_t* mo2 = fetch_input(a, s - a->o2 - a->e2, CM);
  const wf_t* ie = fetch_input(a, s - a->e, CI);
  const wf_t* de = fetch_input(a, s - a->e, CD);
  const wf_t* ie2 = fetch_input(a, s - a->e2, CI2);
  const wf_t* de2 = fetch_input(a, s - a->e2, CD2);
  if (!mm && !mo && !mo2 && !ie && !de && !ie2 && !de2) {   /* null step (:346-357) */
    a->null_steps++;
    return;
  }
  a->null_steps = 0;
  int lo = mm ? mm->lo : 1, hi = mm ? mm->hi : -1, t;
  t = (mo ? mo->lo : 1) - 1;   if (lo > t) lo = t;
  t = (mo ? mo->hi : -1) + 1;  if (hi < t) hi = t;
  t = (ie ? ie->lo : 1) + 1;   if (lo > t) lo = t;
  t = (ie ? ie->hi : -1) + 1;  if (hi < t) hi = t;
  t = (de ? de->lo : 1) - 1;   if (lo > t) lo = t;
  t = (de ? de->hi : -1) - 1;  if (hi < t) hi = t;
  t = (mo2 ? mo2->lo : 1) - 1; if (lo > t) lo = t;
  t = (mo2 ? mo2->hi : -1) + 1; if (hi < t) hi = t;
  t = (ie2 ? ie2->lo : 1) + 1; if (lo > t) lo = t;
  t = (ie2 ? ie2->hi : -1) + 1; if (hi < t) hi = t;
  t = (de2 ? de2->lo : 1) - 1; if (lo > t) lo = t;
  t = (de2 ? de2->hi : -1) - 1; if (hi < t) hi = t;
  a->cnt->cells += hi - lo + 1;
  a->cnt->compute_steps++;
  const int has_i = (mo != NULL || ie != NULL), has_d = (mo != NULL || de != NULL);
  const int has_i2 = (mo2 != NULL || ie2 != NULL), has_d2 = (mo2 != NULL || de2 != NULL);
  wf_t* om = &out->c[CM];
  wf_t* oi = &out->c[CI];
  wf_t* od = &out->c[CD];
  wf_t* oi2 = &out->c[CI2];
  wf_t* od2 = &out->c[CD2];
  wf_reserve(om, lo, hi);
  if (has_i) wf_reserve(oi, lo, hi);
  if (has_d) wf_reserve(od, lo, hi);
  if (has_i2) wf_reserve(oi2, lo, hi);
  if (has_d2) wf_reserve(od2, lo, hi);
  const uint32_t tl = (uint32_t)a->tlen, pl = (uint32_t)a->plen;
  for (int k = lo; k <= hi; ++k) {
    int32_t a1 = wf_get(mo, k - 1), b1 = wf_get(ie, k - 1);
    const int32_t ins1 = (a1 >= b1 ? a1 : b1) + 1;
    a1 = wf_get(mo2, k - 1); b1 = wf_get(ie2, k - 1);
    const int32_t ins2 = (a1 >= b1 ? a1 : b1) + 1;
    const int32_t ins = ins1 >= ins2 ? ins1 : ins2;
    a1 = wf_get(mo, k + 1); b1 = wf_get(de, k + 1);
    const int32_t del1 = (a1 >= b1 ? a1 : b1);
    a1 = wf_get(mo2, k + 1); b1 = wf_get(de2, k + 1);
    const int32_t del2 = (a1 >= b1 ? a1 : b1);
    const int32_t del = del1 >= del2 ? del1 : del2;
    const int32_t misms = wf_get(mm, k) + 1;
    int32_t mx = (misms >= ins ? misms : ins);
    mx = (del >= mx ? del : mx);
    if ((uint32_t)mx > tl || (uint32_t)(mx - k) > pl) mx = OFFSET_NULL;
    om->off[k - lo] = mx;
    if (has_i) oi->off[k - lo] = ins1;
    if (has_d) od->off[k - lo] = del1;
    if (has_i2) oi2->off[k - lo] = ins2;
    if (has_d2) od2->off[k - lo] = del2;
  }
  trim_ends(a, om);
  if (has_i) trim_ends(a, oi);
  if (has_d) trim_ends(a, od);
  if (has_i2) trim_ends(a, oi2);
  if (has_d2) trim_ends(a, od2);
}

static void compute_affine(aligner_t* a, int s) {
  /* Reserve the output slot first: growing the slot table may move it.  The
   * inputs (scores s-x, s-o-e, s-e) never share the ring slot of s. */
  wfset_t* out = slot_of(a, s);
  for (int q = 0; q < NCOMP; ++q) out->c[q].present = 0;
  if (a->a2p) { compute_affine2p(a, s, out); return; }
  const wf_t* mm = fetch_input(a, s - a->x, CM);
  const wf_t* mo = fetch_input(a, s - a->o - a->e, CM);
  const wf_t* ie = fetch_input(a, s - a->e, CI);
  const wf_t* de = fetch_input(a, s - a->e, CD);
  if (!mm && !mo && !ie && !de) {                    /* null step (:245-251) */
    a->null_steps++;
    return;
  }
  a->null_steps = 0;
  /* wavefront_compute_limits_input (wavefront_compute.c:40-86); a null input
   * contributes the null wavefront's limits lo=1, hi=-1. */
  int lo = mm ? mm->lo : 1, hi = mm ? mm->hi : -1, t;
  t = (mo ? mo->lo : 1) - 1;  if (lo > t) lo = t;
  t = (mo ? mo->hi : -1) + 1; if (hi < t) hi = t;
  t = (ie ? ie->lo : 1) + 1;  if (lo > t) lo = t;
  t = (ie ? ie->hi : -1) + 1; if (hi < t) hi = t;
  t = (de ? de->lo : 1) - 1;  if (lo > t) lo = t;
  t = (de ? de->hi : -1) - 1; if (hi < t) hi = t;
  a->cnt->cells += hi - lo + 1;
  a->cnt->compute_steps++;
  /* I1/D1 exist only if one of their inputs does (wavefront_compute.c:454-472). */
  const int has_i = (mo != NULL || ie != NULL), has_d = (mo != NULL || de != NULL);
  wf_t* om = &out->c[CM];
  wf_t* oi = &out->c[CI];
  wf_t* od = &out->c[CD];
  wf_reserve(om, lo, hi);
  if (has_i) wf_reserve(oi, lo, hi);
  if (has_d) wf_reserve(od, lo, hi);
  const uint32_t tl = (uint32_t)a->tlen, pl = (uint32_t)a->plen;
  for (int k = lo; k <= hi; ++k) {
    int32_t a1 = wf_get(mo, k - 1), b1 = wf_get(ie, k - 1);
    const int32_t ins1 = (a1 >= b1 ? a1 : b1) + 1;
    a1 = wf_get(mo, k + 1); b1 = wf_get(de, k + 1);
    const int32_t del1 = (a1 >= b1 ? a1 : b1);
    const int32_t misms = wf_get(mm, k) + 1;
    int32_t mx = (misms >= ins1 ? misms : ins1);
    mx = (del1 >= mx ? del1 : mx);
    if ((uint32_t)mx > tl || (uint32_t)(mx - k) > pl) mx = OFFSET_NULL;
    om->off[k - lo] = mx;
    if (has_i) oi->off[k - lo] = ins1;
    if (has_d) od->off[k - lo] = del1;
  }
  trim_ends(a, om);
  if (has_i) trim_ends(a, oi);
  if (has_d) trim_ends(a, od);
}

/* ------------------------------------------------------------------------- */
/* extend + termination (wavefront_extend.c:86-211, wavefront_termination.c:37-114) */

static int lcp(const aligner_t* a, int v, int h) {
  const int lp = a->plen - v, lt = a->tlen - h;
  const int lim = lp < lt ? lp : lt;
  int n = 0;
  if (!a->reverse) {
    const char* p = a->P + a->pb + v;
    const char* t = a->T + a->tb + h;
    while (n < lim && p[n] == t[n]) ++n;
  } else {
    const char* p = a->P + a->pb + a->plen - 1 - v;
    const char* t = a->T + a->tb + a->tlen - 1 - h;
    while (n < lim && p[-n] == t[-n]) ++n;
  }
  return n;
}

static int termination(aligner_t* a, int s) {
  const int k_end = a->tlen - a->plen, off_end = a->tlen;
  const wf_t* w = &slot_of(a, s)->c[a->comp_end];
  if (!w->present || w->lo > k_end || k_end > w->hi) return 0;
  if (w->off[k_end - w->clo] < off_end) return 0;
  a->end_k = k_end;
  a->end_off = off_end;
  return 1;
}

/* Returns 1 when the aligner is done (end reached / unreachable). */
static int extend_end2end(aligner_t* a, int s, int* max_antidiag) {
  if (max_antidiag) *max_antidiag = 0;
  wf_t* m = &slot_of(a, s)->c[CM];
  if (!m->present) {
    if (a->null_steps > a->scope) {
      a->status = ST_END_UNREACHABLE;
      a->status_score = s;
      return 1;
    }
    return 0;
  }
  int max_ak = 0;
  for (int k = m->lo; k <= m->hi; ++k) {
    int32_t off = m->off[k - m->clo];
    if (off == OFFSET_NULL) continue;
    const int run = lcp(a, off - k, off);
    a->cnt->words += run / 32 + 1;
    off += run;
    m->off[k - m->clo] = off;
    const int ak = 2 * off - k;
    if (max_ak < ak) max_ak = ak;
  }
  if (termination(a, s)) {
    a->status = ST_END_REACHED;
    a->status_score = s;
    return 1;
  }
  if (max_antidiag) *max_antidiag = max_ak;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* breakpoint detection (wavefront_bialign.c:189-519)                        */

static void breakpoint_indel2indel(const aligner_t* a0, int forward, int s0, int s1,
                                   const wf_t* w0, const wf_t* w1, int comp, int gap_open,
                                   breakpoint_t* bp) {
  const int tl = a0->tlen, pl = a0->plen;
  const int lo0 = w0->lo, hi0 = w0->hi;
  const int lo1 = tl - pl - w1->hi, hi1 = tl - pl - w1->lo;   /* WAVEFRONT_K_INVERSE */
  if (hi1 < lo0 || hi0 < lo1) return;
  const int min_hi = hi0 < hi1 ? hi0 : hi1, max_lo = lo0 > lo1 ? lo0 : lo1;
  for (int k0 = max_lo; k0 <= min_hi; ++k0) {
    const int k1 = tl - pl - k0;
    const int dh0 = w0->off[k0 - w0->clo];
    const int dh1 = w1->off[k1 - w1->clo];
    if (dh0 + dh1 >= tl && s0 + s1 - gap_open < bp->score) {
      if (forward) {
        if (dh0 - k0 > pl || dh0 > tl) continue;
        bp->score_forward = s0; bp->score_reverse = s1;
        bp->k_forward = k0; bp->k_reverse = k1;
        bp->offset_forward = dh0; bp->offset_reverse = dh1;
      } else {
        if (dh1 - k1 > pl || dh1 > tl) continue;
        bp->score_forward = s1; bp->score_reverse = s0;
        bp->k_forward = k1; bp->k_reverse = k0;
        bp->offset_forward = dh1; bp->offset_reverse = dh0;
      }
      bp->score = s0 + s1 - gap_open;
      bp->component = comp;
      return;
    }
  }
}

static void breakpoint_m2m(const aligner_t* a0, int forward, int s0, int s1,
                           const wf_t* w0, const wf_t* w1, breakpoint_t* bp) {
  const int tl = a0->tlen, pl = a0->plen;
  const int lo0 = w0->lo, hi0 = w0->hi;
  const int lo1 = tl - pl - w1->hi, hi1 = tl - pl - w1->lo;
  if (hi1 < lo0 || hi0 < lo1) return;
  const int min_hi = hi0 < hi1 ? hi0 : hi1, max_lo = lo0 > lo1 ? lo0 : lo1;
  for (int k0 = max_lo; k0 <= min_hi; ++k0) {
    const int k1 = tl - pl - k0;
    const int m0 = w0->off[k0 - w0->clo];
    const int m1 = w1->off[k1 - w1->clo];
    if (m0 + m1 >= tl && s0 + s1 < bp->score) {
      if (forward) {
        bp->score_forward = s0; bp->score_reverse = s1;
        bp->k_forward = k0; bp->k_reverse = k1;
        bp->offset_forward = m0; bp->offset_reverse = m1;
      } else {
        bp->score_forward = s1; bp->score_reverse = s0;
        bp->k_forward = k1; bp->k_reverse = k0;
        bp->offset_forward = m1; bp->offset_reverse = m0;
      }
      bp->score = s0 + s1;
      bp->component = CM;
      return;
    }
  }
}

static void bialign_overlap(aligner_t* a0, aligner_t* a1, int s0, int s1, int forward,
                            breakpoint_t* bp) {
  const int scope = a0->scope, gap_open = a0->o;
  wfset_t* set0 = slot_of(a0, s0);
  if (!set0->c[CM].present) return;
  for (int i = 0; i < scope; ++i) {
    const int si = s1 - i;
    if (si < 0) break;
    wfset_t* set1 = slot_of(a1, si);
    if (a0->a2p) {                   /* I2/D2 breakpoints first (:347-364) */
      if (s0 + si - a0->o2 >= bp->score) continue;
      if (set0->c[CD2].present && set1->c[CD2].present)
        breakpoint_indel2indel(a0, forward, s0, si, &set0->c[CD2], &set1->c[CD2], CD2, a0->o2, bp);
      if (set0->c[CI2].present && set1->c[CI2].present)
        breakpoint_indel2indel(a0, forward, s0, si, &set0->c[CI2], &set1->c[CI2], CI2, a0->o2, bp);
    }
    if (s0 + si - gap_open >= bp->score) continue;
    if (set0->c[CD].present && set1->c[CD].present)
      breakpoint_indel2indel(a0, forward, s0, si, &set0->c[CD], &set1->c[CD], CD, gap_open, bp);
    if (set0->c[CI].present && set1->c[CI].present)
      breakpoint_indel2indel(a0, forward, s0, si, &set0->c[CI], &set1->c[CI], CI, gap_open, bp);
    if (s0 + si >= bp->score) continue;
    if (set1->c[CM].present)
      breakpoint_m2m(a0, forward, s0, si, &set0->c[CM], &set1->c[CM], bp);
  }
}

static int find_breakpoint(ctx_t* c, int comp_begin, int comp_end, breakpoint_t* bp) {
  aligner_t* f = &c->fwd;
  aligner_t* r = &c->rev;
  c->cnt.breakpoints++;
  aligner_init(f, comp_begin, comp_end);
  aligner_init(r, comp_end, comp_begin);
  const int max_antidiagonal = f->plen + f->tlen - 1;
  int sf = 0, sr = 0, f_max = 0, r_max = 0, ak = 0;
  bp->score = INT_MAX;
  if (extend_end2end(f, sf, &f_max)) return f->status;
  if (extend_end2end(r, sr, &r_max)) return r->status;
  int last_forward = 0;
  for (;;) {                                                 /* phase 1 (:447-480) */
    if (f_max + r_max >= max_antidiagonal) break;
    ++sf;
    compute_affine(f, sf);
    const int fq = extend_end2end(f, sf, &ak);
    if (f_max < ak) f_max = ak;
    last_forward = 1;
    if (fq) return f->status;
    if (f_max + r_max >= max_antidiagonal) break;
    ++sr;
    compute_affine(r, sr);
    const int rq = extend_end2end(r, sr, &ak);
    if (r_max < ak) r_max = ak;
    last_forward = 0;
    if (rq) return r->status;
    if (sr + sf >= c->max_steps) return BWO_STATUS_MAX_STEPS;
  }
  /* phase 2 (:482-516); the opening penalty of the overlap bound is
   * wavefront_bialign_overlap_gopen_adjust (:396-410): max(o1, o2) for 2p */
  const int scope = f->scope, gap_open = (f->a2p && f->o2 > f->o) ? f->o2 : f->o;
  for (;;) {
    if (last_forward) {
      const int min_r = (sr > scope - 1) ? sr - (scope - 1) : 0;
      if (sf + min_r - gap_open >= bp->score) break;
      bialign_overlap(f, r, sf, sr, 1, bp);
      ++sr;
      compute_affine(r, sr);
      if (extend_end2end(r, sr, NULL)) return r->status;
    }
    const int min_f = (sf > scope - 1) ? sf - (scope - 1) : 0;
    if (min_f + sr - gap_open >= bp->score) break;
    bialign_overlap(r, f, sr, sf, 0, bp);
    ++sf;
    compute_affine(f, sf);
    if (extend_end2end(f, sf, NULL)) return f->status;
    if (sr + sf >= c->max_steps) return BWO_STATUS_MAX_STEPS;
    last_forward = 1;
  }
  return ST_OK;
}

/* ------------------------------------------------------------------------- */
/* base case: unidirectional WFA + backtrace                                 */
/* (wavefront_bialign.c:155-185, wavefront_unialign.c:54-273,               */
/*  wavefront_backtrace.c:320-529)                                           */

static void ops_push(ctx_t* c, char op, int n) {
  if (n <= 0) return;
  if (c->nops + n > c->ops_cap) {
    int cap = c->ops_cap ? c->ops_cap : 1024;
    while (cap < c->nops + n) cap *= 2;
    c->ops = (char*)realloc(c->ops, (size_t)cap);
    c->ops_cap = cap;
  }
  memset(c->ops + c->nops, op, (size_t)n);
  c->nops += n;
}

/* reverse-order op buffer used by the backtrace */
static void rops_push(ctx_t* c, int* n, char op, int cnt) {
  if (cnt <= 0) return;
  if (*n + cnt > c->rops_cap) {
    int cap = c->rops_cap ? c->rops_cap : 1024;
    while (cap < *n + cnt) cap *= 2;
    c->rops = (char*)realloc(c->rops, (size_t)cap);
    c->rops_cap = cap;
  }
  memset(c->rops + *n, op, (size_t)cnt);
  *n += cnt;
}

#define BT_M 9
#define BT_D2_EXT 8
#define BT_D2_OPEN 7
#define BT_D1_EXT 6
#define BT_D1_OPEN 5
#define BT_I1_EXT 2
#define BT_I1_OPEN 1
#define BT_I2_EXT 4
#define BT_I2_OPEN 3

/* wavefront_backtrace_* source lookups (wavefront_backtrace.c:64-219): the
 * candidate is the packed (offset << 4) | type, or OFFSET_NULL itself. */
static int64_t bt_src(aligner_t* a, int score, int comp, int k, int add, int type) {
  if (score < 0 || score >= a->nwfs) return OFFSET_NULL;
  const wf_t* w = &a->wfs[score].c[comp];
  if (!w->present || !(w->lo <= k && k <= w->hi)) return OFFSET_NULL;
  return (int64_t)(w->off[k - w->clo] + add) * 16 | type;
}
static inline int64_t max64(int64_t p, int64_t q) { return p >= q ? p : q; }

static int backtrace_affine(ctx_t* c, aligner_t* a, int score) {
  int matrix = a->comp_end;
  int k = a->end_k, off = a->end_off;
  int h = off, v = off - k;
  int n = 0;
  if (matrix == CM) {
    if (v < a->plen) rops_push(c, &n, 'D', a->plen - v);
    if (h < a->tlen) rops_push(c, &n, 'I', a->tlen - h);
  }
  while (v > 0 && h > 0 && score > 0) {
    const int mismatch = score - a->x;
    const int gap_open1 = score - a->o - a->e;
    const int gap_extend1 = score - a->e;
    const int gap_open2 = score - a->o2 - a->e2;
    const int gap_extend2 = score - a->e2;
    int64_t max_all;
    if (matrix == CM) {
      const int64_t misms = bt_src(a, mismatch, CM, k, 1, BT_M);
      const int64_t io = bt_src(a, gap_open1, CM, k - 1, 1, BT_I1_OPEN);
      const int64_t ix = bt_src(a, gap_extend1, CI, k - 1, 1, BT_I1_EXT);
      const int64_t dop = bt_src(a, gap_open1, CM, k + 1, 0, BT_D1_OPEN);
      const int64_t dx = bt_src(a, gap_extend1, CD, k + 1, 0, BT_D1_EXT);
      max_all = max64(misms, max64(max64(io, ix), max64(dop, dx)));
      if (a->a2p) {                  /* the I2/D2 sources (wavefront_backtrace.c:377-386) */
        const int64_t io2 = bt_src(a, gap_open2, CM, k - 1, 1, BT_I2_OPEN);
        const int64_t ix2 = bt_src(a, gap_extend2, CI2, k - 1, 1, BT_I2_EXT);
        const int64_t dop2 = bt_src(a, gap_open2, CM, k + 1, 0, BT_D2_OPEN);
        const int64_t dx2 = bt_src(a, gap_extend2, CD2, k + 1, 0, BT_D2_EXT);
        max_all = max64(max_all, max64(max64(io2, ix2), max64(dop2, dx2)));
      }
    } else if (matrix == CI) {
      max_all = max64(bt_src(a, gap_open1, CM, k - 1, 1, BT_I1_OPEN),
                      bt_src(a, gap_extend1, CI, k - 1, 1, BT_I1_EXT));
    } else if (matrix == CD) {
      max_all = max64(bt_src(a, gap_open1, CM, k + 1, 0, BT_D1_OPEN),
                      bt_src(a, gap_extend1, CD, k + 1, 0, BT_D1_EXT));
    } else if (matrix == CI2) {
      max_all = max64(bt_src(a, gap_open2, CM, k - 1, 1, BT_I2_OPEN),
                      bt_src(a, gap_extend2, CI2, k - 1, 1, BT_I2_EXT));
    } else {
      max_all = max64(bt_src(a, gap_open2, CM, k + 1, 0, BT_D2_OPEN),
                      bt_src(a, gap_extend2, CD2, k + 1, 0, BT_D2_EXT));
    }
    if (max_all < 0) break;
    if (matrix == CM) {
      const int max_off = (int)(max_all >> 4);
      rops_push(c, &n, 'M', off - max_off);
      off = max_off;
      v = off - k; h = off;
      if (v <= 0 || h <= 0) break;
    }
    switch ((int)(max_all & 15)) {
      case BT_M:       score = mismatch;    matrix = CM; rops_push(c, &n, 'X', 1); --off; break;
      case BT_I1_OPEN: score = gap_open1;   matrix = CM; rops_push(c, &n, 'I', 1); --k; --off; break;
      case BT_I1_EXT:  score = gap_extend1; matrix = CI; rops_push(c, &n, 'I', 1); --k; --off; break;
      case BT_D1_OPEN: score = gap_open1;   matrix = CM; rops_push(c, &n, 'D', 1); ++k; break;
      case BT_D1_EXT:  score = gap_extend1; matrix = CD; rops_push(c, &n, 'D', 1); ++k; break;
      case BT_I2_OPEN: score = gap_open2;   matrix = CM; rops_push(c, &n, 'I', 1); --k; --off; break;
      case BT_I2_EXT:  score = gap_extend2; matrix = CI2; rops_push(c, &n, 'I', 1); --k; --off; break;
      case BT_D2_OPEN: score = gap_open2;   matrix = CM; rops_push(c, &n, 'D', 1); ++k; break;
      case BT_D2_EXT:  score = gap_extend2; matrix = CD2; rops_push(c, &n, 'D', 1); ++k; break;
      default: return -1;
    }
    v = off - k; h = off;
  }
  if (matrix == CM) {
    if (v > 0 && h > 0) {
      const int nm = v < h ? v : h;
      rops_push(c, &n, 'M', nm);
      v -= nm; h -= nm;
    }
    rops_push(c, &n, 'D', v);
    rops_push(c, &n, 'I', h);
  } else if (v != 0 || h != 0 || (score != 0 && c->match == 0)) {
    return -1;       /* "I?/D?-Beginning backtrace error" (wavefront_backtrace.c:517-524) */
  }
  /* cigar_append_forward of the base CIGAR (cigar.c:125-136) */
  for (int i = n - 1; i >= 0; --i) ops_push(c, c->rops[i], 1);
  return 0;
}

/* Returns an internal status (ST_OK on success) and the score in *score_out. */
static int bialign_base(ctx_t* c, int pb, int pe, int tb, int te,
                        int comp_begin, int comp_end, int* score_out) {
  aligner_t* a = &c->base;
  c->cnt.base_cases++;
  aligner_bounds(a, pb, pe, tb, te);
  aligner_init(a, comp_begin, comp_end);
  int score = 0;
  for (;;) {
    if (extend_end2end(a, score, NULL)) {
      if (a->status != ST_END_REACHED) return BWO_STATUS_UNATTAIN;
      if (backtrace_affine(c, a, score) != 0) { c->err = 1; return BWO_STATUS_UNATTAIN; }
      *score_out = score;
      return ST_OK;
    }
    ++score;
    compute_affine(a, score);
    if (score >= c->max_steps) return BWO_STATUS_UNATTAIN;  /* base reports MAX_STEPS as unattainable */
  }
}

static int bialign_alignment(ctx_t* c, int pb, int pe, int tb, int te,
                             int comp_begin, int comp_end, int score_remaining, int level) {
  const int plen = pe - pb, tlen = te - tb;
  if (level > c->cnt.max_level) c->cnt.max_level = level;
  if (tlen == 0) { ops_push(c, 'D', plen); return ST_OK; }
  if (plen == 0) { ops_push(c, 'I', tlen); return ST_OK; }
  int base_score = 0;
  if (score_remaining <= FALLBACK_MIN_SCORE) {
    const int st = bialign_base(c, pb, pe, tb, te, comp_begin, comp_end, &base_score);
    if (level == 0) c->level0_wf_score = base_score;
    return st;
  }
  aligner_bounds(&c->fwd, pb, pe, tb, te);
  aligner_bounds(&c->rev, pb, pe, tb, te);
  breakpoint_t bp;
  int st = find_breakpoint(c, comp_begin, comp_end, &bp);
  if (st != ST_OK) {                                  /* :520-547 */
    if (st == ST_END_REACHED) {
      const int reached = (c->fwd.status == ST_END_REACHED) ? c->fwd.status_score
                                                            : c->rev.status_score;
      if (reached <= RECOVERY_MIN_SCORE) {
        st = bialign_base(c, pb, pe, tb, te, comp_begin, comp_end, &base_score);
        if (level == 0) c->level0_wf_score = base_score;
        return st;
      }
      return ST_END_UNREACHABLE;
    }
    return st;
  }
  const int bh = bp.offset_forward, bv = bp.offset_forward - bp.k_forward;
  st = bialign_alignment(c, pb, pb + bv, tb, tb + bh, comp_begin, bp.component,
                         bp.score_forward, level + 1);
  if (st != ST_OK) return st;
  st = bialign_alignment(c, pb + bv, pe, tb + bh, te, bp.component, comp_end,
                         bp.score_reverse, level + 1);
  if (st != ST_OK) return st;
  if (level == 0) c->level0_wf_score = bp.score;
  return ST_OK;
}

/* ------------------------------------------------------------------------- */
/* driver                                                                    */

/* User penalties: gap-affine (a2p = 0) or gap-affine-2p (a2p = 1, second gap
 * model o2 + e2 * length). */
typedef struct { int match, x, o, e, o2, e2, a2p; } pen_t;

static void ctx_setup(ctx_t* c, const char* P, int plen, const char* T, int tlen,
                      const pen_t* u, int max_steps) {
  /* wavefront_penalties_set_affine (wavefront_penalties.c:95-131) and
   * wavefront_penalties_set_affine2p (:132-173): Eizenga's transform for match < 0 */
  const int match = u->match;
  int ix = u->x, io = u->o, ie = u->e, io2 = u->o2, ie2 = u->e2;
  if (match < 0) {
    ix = 2 * u->x - 2 * match; io = 2 * u->o; ie = 2 * u->e - match;
    io2 = 2 * u->o2; ie2 = 2 * u->e2 - match;
  }
  /* wavefront_components_dimensions_affine / _affine2p (wavefront_components.c:81-112) */
  int scope_indel = io + ie;
  if (u->a2p && io2 + ie2 > scope_indel) scope_indel = io2 + ie2;
  const int scope = (scope_indel > ix ? scope_indel : ix) + 1;
  aligner_t* as[3] = {&c->fwd, &c->rev, &c->base};
  for (int i = 0; i < 3; ++i) {
    aligner_t* a = as[i];
    a->x = ix; a->o = io; a->e = ie;
    a->a2p = u->a2p;
    a->o2 = u->a2p ? io2 : 0; a->e2 = u->a2p ? ie2 : 0;
    a->ncomp = u->a2p ? 5 : 3;
    a->P = P; a->T = T;
    a->scope = scope;
    a->modular = (i < 2);          /* breakpoint aligners are score-only (modular) */
    a->reverse = (i == 1);
    a->cnt = &c->cnt;
    aligner_bounds(a, 0, plen, 0, tlen);
  }
  c->match = match;
  c->max_steps = (max_steps > 0) ? max_steps : INT_MAX;
  c->nops = 0;
  c->err = 0;
  c->level0_wf_score = INT32_MIN;
}

static void ctx_free(ctx_t* c) {
  aligner_t* as[3] = {&c->fwd, &c->rev, &c->base};
  for (int i = 0; i < 3; ++i) {
    for (int s = 0; s < as[i]->nwfs; ++s)
      for (int q = 0; q < NCOMP; ++q) free(as[i]->wfs[s].c[q].off);
    free(as[i]->wfs);
  }
  free(c->ops);
  free(c->rops);
}

/* cigar_score_gap_affine (cigar.c:277-303) over the operation string */
static int score_gap_affine(const char* ops, int n, int match, int x, int o, int e) {
  char last = 0;
  int score = 0;
  for (int i = 0; i < n; ++i) {
    switch (ops[i]) {
      case 'M': score -= match; break;
      case 'X': score -= x; break;
      case 'D': score -= e + (last == 'D' ? 0 : o); break;
      case 'I': score -= e + (last == 'I' ? 0 : o); break;
    }
    last = ops[i];
  }
  return score;
}

/* cigar_score_gap_affine2p (cigar.c:304-342): each maximal gap run costs the
 * cheaper of its two gap models */
static int score_gap_affine2p(const char* ops, int n, const pen_t* u) {
  int score = 0;
  for (int i = 0; i < n;) {
    int j = i;
    while (j < n && ops[j] == ops[i]) ++j;
    const int len = j - i;
    switch (ops[i]) {
      case 'M': score -= u->match * len; break;
      case 'X': score -= u->x * len; break;
      default: {
        const int s1 = u->o + u->e * len, s2 = u->o2 + u->e2 * len;
        score -= s1 < s2 ? s1 : s2;
      }
    }
    i = j;
  }
  return score;
}

static int ops_to_runs(const char* ops, int n, uint32_t* runs, int cap) {
  int nr = 0;
  for (int i = 0; i < n;) {
    int j = i;
    while (j < n && ops[j] == ops[i]) ++j;
    const uint32_t code = ops[i] == 'M' ? 7u : ops[i] == 'X' ? 8u : ops[i] == 'I' ? 1u : 2u;
    if (nr >= cap) return -1;
    runs[nr++] = ((uint32_t)(j - i) << 4) | code;
    i = j;
  }
  return nr;
}

static int align_ctx(ctx_t* c, const char* P, int plen, const char* T, int tlen,
                     const pen_t* u, int score_only, int max_steps,
                     uint32_t* runs, int runs_cap, int* nruns, int* alg_score, int* wf_score) {
  const int match = u->match;
  ctx_setup(c, P, plen, T, tlen, u, max_steps);
  if (nruns) *nruns = 0;
  if (score_only) {
    /* wavefront_bialign_compute_score (wavefront_bialign.c:662-699) */
    breakpoint_t bp;
    const int st = find_breakpoint(c, CM, CM, &bp);
    if (st != ST_OK && st != ST_END_REACHED) {
      *wf_score = INT32_MIN;
      *alg_score = INT32_MIN;
      return (st == BWO_STATUS_MAX_STEPS) ? st : BWO_STATUS_UNATTAIN;
    }
    int s = bp.score;
    if (st == ST_END_REACHED)
      s = (c->fwd.status == ST_END_REACHED) ? c->fwd.status_score : c->rev.status_score;
    *wf_score = s;
    /* wavefront_compute_classic_score (wavefront_compute.c:108-120) */
    *alg_score = (match == 0) ? -s : ((-match) * (plen + tlen) - s) / 2;
    return BWO_STATUS_COMPLETED;
  }
  /* wavefront_bialign (wavefront_bialign.c:703-730) */
  const int min_length = (plen > tlen ? plen : tlen) <= FALLBACK_MIN_LENGTH;
  int st = bialign_alignment(c, 0, plen, 0, tlen, CM, CM, min_length ? 0 : INT_MAX, 0);
  if (c->err) st = BWO_STATUS_UNATTAIN;
  int status;
  if (st == ST_OK) status = BWO_STATUS_COMPLETED;
  else if (st == BWO_STATUS_MAX_STEPS || st == BWO_STATUS_OOM) status = st;
  else status = BWO_STATUS_UNATTAIN;
  *wf_score = c->level0_wf_score;
  if (c->nops <= 0) *alg_score = -1;
  else if (u->a2p) *alg_score = score_gap_affine2p(c->ops, c->nops, u);
  else *alg_score = score_gap_affine(c->ops, c->nops, match, u->x, u->o, u->e);
  if (runs) {
    const int nr = ops_to_runs(c->ops, c->nops, runs, runs_cap);
    if (nr < 0) { if (nruns) *nruns = -1; return BWO_STATUS_CAPACITY; }
    if (nruns) *nruns = nr;
  }
  return status;
}

int bwo_align(const char* pattern, int plen, const char* text, int tlen,
              int match, int x, int o, int e, int score_only, int max_steps,
              uint32_t* runs, int runs_cap, int* nruns,
              int* alg_score, int* wf_score, bwo_counters_t* counters) {
  ctx_t c;
  memset(&c, 0, sizeof(c));
  const pen_t u = {match, x, o, e, 0, 0, 0};
  const int st = align_ctx(&c, pattern, plen, text, tlen, &u, score_only,
                           max_steps, runs, runs_cap, nruns, alg_score, wf_score);
  if (counters) *counters = c.cnt;
  ctx_free(&c);
  return st;
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void align_batch(int n, const char* seqbuf,
                        const int64_t* p_off, const int32_t* p_len,
                        const int64_t* t_off, const int32_t* t_len,
                        const pen_t* u, int score_only, int max_steps, int nthreads,
                        int32_t* alg_score, int32_t* wf_score, int32_t* status,
                        uint32_t* runs, const int64_t* runs_off, int32_t* nruns,
                        int64_t* cells, int64_t* words, double* elapsed) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const double t0 = now_seconds();
  /* chunks of 16 pairs (dynamic), fewer when that would leave a thread with
   * under 64 chunks: a small sample of long pairs must still spread over
   * every thread */
  const int chunk = n / (64 * nthreads) >= 16 ? 16 : (n / (64 * nthreads) >= 1 ? n / (64 * nthreads) : 1);
  #pragma omp parallel num_threads(nthreads)
  {
    ctx_t c;
    memset(&c, 0, sizeof(c));
    #pragma omp for schedule(dynamic, chunk)
    for (int i = 0; i < n; ++i) {
      memset(&c.cnt, 0, sizeof(c.cnt));
      int nr = 0, as = 0, ws = 0;
      const int64_t cap = runs ? runs_off[i + 1] - runs_off[i] : 0;
      status[i] = align_ctx(&c, seqbuf + p_off[i], p_len[i], seqbuf + t_off[i], t_len[i],
                            u, score_only, max_steps,
                            runs ? runs + runs_off[i] : NULL, (int)cap, &nr, &as, &ws);
      alg_score[i] = as;
      wf_score[i] = ws;
      if (nruns) nruns[i] = nr;
      if (cells) cells[i] = c.cnt.cells;
      if (words) words[i] = c.cnt.words;
    }
    ctx_free(&c);
  }
  if (elapsed) *elapsed = now_seconds() - t0;
}

int bwo_align_batch(int n, const char* seqbuf,
                    const int64_t* p_off, const int32_t* p_len,
                    const int64_t* t_off, const int32_t* t_len,
                    int match, int x, int o, int e, int score_only, int max_steps,
                    int nthreads,
                    int32_t* alg_score, int32_t* wf_score, int32_t* status,
                    uint32_t* runs, const int64_t* runs_off, int32_t* nruns,
                    int64_t* cells, int64_t* words, double* elapsed) {
  const pen_t u = {match, x, o, e, 0, 0, 0};
  align_batch(n, seqbuf, p_off, p_len, t_off, t_len, &u, score_only, max_steps, nthreads,
              alg_score, wf_score, status, runs, runs_off, nruns, cells, words, elapsed);
  return 0;
}

int bwo_align_batch_2p(int n, const char* seqbuf,
                       const int64_t* p_off, const int32_t* p_len,
                       const int64_t* t_off, const int32_t* t_len,
                       int match, int x, int o1, int e1, int o2, int e2,
                       int score_only, int max_steps, int nthreads,
                       int32_t* alg_score, int32_t* wf_score, int32_t* status,
                       uint32_t* runs, const int64_t* runs_off, int32_t* nruns,
                       int64_t* cells, int64_t* words, double* elapsed) {
  const pen_t u = {match, x, o1, e1, o2, e2, 1};
  align_batch(n, seqbuf, p_off, p_len, t_off, t_len, &u, score_only, max_steps, nthreads,
              alg_score, wf_score, status, runs, runs_off, nruns, cells, words, elapsed);
  return 0;
}
