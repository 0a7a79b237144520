// biwfa_kernel.hip -- gfx950 bidirectional WFA (score + CIGAR): one wave64, or a
// team of wave64s in one workgroup, per pair.
//
// Behaviour follows WFA2-lib v2.3 `wavefront_align` in exact BiWFA mode
// (memory_mode = ultralow, heuristic = none, gap-affine):
//   recursion driver      wavefront/wavefront_bialign.c:581-658, 703-730
//   breakpoint search     wavefront/wavefront_bialign.c:411-519
//   overlap detection     wavefront/wavefront_bialign.c:189-392
//   affine compute        wavefront/wavefront_compute_affine.c:44-86, 229-260
//                         wavefront/wavefront_compute.c:40-104, 306-352, 382-632
//   extend / termination  wavefront/wavefront_extend.c:86-211, wavefront_termination.c:37-114
//   base case + backtrace wavefront/wavefront_bialign.c:155-185,
//                         wavefront/wavefront_unialign.c:54-273,
//                         wavefront/wavefront_backtrace.c:38-529
//
// Device mapping:
//   * A persistent grid of workgroups; each aligner (a wave, or a team) pulls
//     pairs from a device work queue (atomic counter over a schedule order).
//   * Diagonals are spread over the 64 lanes; scores, limits and the recursion
//     state are wave-uniform (scalar registers).
//   * Wavefront limits (lo/hi per component) are kept in vector registers, one
//     lane per ring slot (or per base-case score), read with v_readlane: no
//     memory round trip per header access.
//   * Forward and reverse offsets live in modular rings (scope scores x M/I1/D1 x
//     wcap diagonals, k indexed modulo wcap) -- O(s) memory per pair.  The base
//     case keeps every wavefront of its sub-problem in a bump arena that reuses
//     the ring storage.
//   * Sequences are 2-bit packed; extension compares 32 bases per 64-bit word
//     (xor + count-trailing-zeros forward, count-leading-zeros backward).
//   * Instantiations: Cfg<true,1> keeps the workspace and both sequences in LDS,
//     one wave per pair; a pair that outgrows an LDS capacity is pushed to a
//     spill list and re-run by a global-workspace instantiation with larger
//     capacities.  Cfg<false,NW> with NW > 1 aligns one pair per workgroup of
//     NW waves (a team): each step's diagonals are interleaved over the waves in
//     64-diagonal chunks, and the per-step reductions (trimmed limits, maximum
//     antidiagonal, first overlap hit) go through LDS with one barrier each.
#include <hip/hip_runtime.h>
#include <limits.h>

#include "biwfa_device.hpp"

namespace bw {

#define DI __device__ __forceinline__

enum { CM = 0, CI = 1, CD = 2 };
constexpr int kAbsent = INT_MAX;   // header lo of a component WFA2 has not allocated
// uint32 run op codes (SAM encoding with mismatches shown, cigar.c:38-56)
constexpr int kOpM = 7, kOpX = 8, kOpI = 1, kOpD = 2;

DI int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
DI int rdl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
DI int imax(int a, int b) { return a >= b ? a : b; }
DI int imin(int a, int b) { return a <= b ? a : b; }
// v0, v1 or v2 by an index 0..2 that may differ between lanes, written as
// masking arithmetic: a select of register-array elements would be folded into
// a select of their addresses and pin the array to scratch.
DI int pick3(int v0, int v1, int v2, int w) {
  return (v0 & -(int)(w == 0)) | (v1 & -(int)(w == 1)) | (v2 & -(int)(w == 2));
}

// Orders this wave's workspace writes before later reads by other lanes.  LDS
// operations of one wave execute in issue order, so the LDS kernel only needs
// the compiler not to move memory operations across this point; the global
// workspace needs its stores performed before the loads are issued.
template <bool kLds>
DI void wave_sync() {
  if (kLds) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_wave_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
}

// Orders the aligner's workspace writes before later reads by any of its
// lanes: a wave barrier for a single-wave aligner, a workgroup barrier (with
// its release/acquire fences) for a team.
template <class K>
DI void team_sync() {
  if constexpr (K::kNW == 1) wave_sync<K::kLds>();
  else __syncthreads();
}

// Max over the wave of non-negative values (DPP row scans, then row broadcasts).
DI int wave_max_nonneg(int v) {
  v = imax(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = imax(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = imax(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = imax(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = imax(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = imax(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return rdl(v, kWave - 1);
}

// Max / min over the wave of any values (DPP row scans, then row broadcasts).
DI int wave_max_any(int v) {
  v = imax(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = imax(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = imax(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = imax(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = imax(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = imax(v, __builtin_amdgcn_update_dpp(INT_MIN, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return rdl(v, kWave - 1);
}
DI int wave_min_any(int v) {
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return rdl(v, kWave - 1);
}

// A wavefront component as seen by readers: offsets for k in [lo,hi] live at
// element base + (k & mask) of the offset store (mask = wcap-1 in a ring, -1 in
// the arena).  lo > hi is the null wavefront (WFA2's wavefront_null has lo=1, hi=-1).
struct Wf {
  int lo, hi, base, mask;
};
DI Wf wf_null() { return Wf{1, -1, 0, 0}; }

// One aligner's view of a sub-problem.
struct Dir {
  int pb, plen, tb, tlen;   // sub-range of the pair
  int rev;                  // 1: reads the sub-range backwards (reverse aligner)
  int ring;                 // 0 forward ring, 1 reverse ring, -1 base-case arena
  int cb, ce;               // begin / end component
  int null_steps;
  int status, status_score;
};

struct BP {                 // wavefront_bialigner.h:43-54
  int score, sf, sr, kf, kr, of, orr, comp;
};

template <class K>
struct Ctx {
  int32_t* ws;    // shared workspace (ring / arena, then the private regions)
  int32_t* wd;    // the ring / arena offsets (at the workspace base)
  int32_t* pv;    // this wave's private region (base headers, backtrace runs, stack)
  const uint64_t* P;
  const uint64_t* T;
  int seq_lds;    // P and T point into LDS (staged), else into global memory
  int lane;
  int wid;        // wave index within the team (0 for a single-wave aligner)
  bool leader;    // lane 0 of wave 0: the one lane that writes results
  int32_t* tred;  // team reduction scratch in LDS (team kernels)
  int tph;        // its double-buffer phase
  WsLayout L;
  int x, o, e, scope, max_steps;
  int um, ux, uo, ue;   // user penalties (global kernels)
  int upen;             // the same in lanes 0..3 (LDS kernel)
  int qm0, qm1, qm2;    // LDS kernel: lane masks (-1 / 0) of the input gather's component
  int qlo, qhi;         // LDS kernel: the gather's per-lane limit adjustments
  int qdelta, qcomp;    // LDS kernel: the gather's per-lane score distance and component
  int om0, om1, om2;    // LDS kernel: lane masks of the overlap candidates' kind (p % 3)
  int arena_top;
  // headers in registers: ring (lane = ring*scope + slot) and base case (lane = score)
  int rlo[3], rhi[3];
  int blo[3], bhi[3], bkb[3];
  // presence: a header lo other than kAbsent (ring and base-case headers alike)
  int gmask;      // scores s with (s & gmask) != 0 are null steps (even penalties)
  uint32_t* out;
  int out_cap, out_n, pend_op, pend_len, alg_score;
  int spill;      // a workspace capacity was exceeded
  int cap_err;    // the pair's output run capacity was exceeded
  int guard;      // score bound of the current sub-problem (termination guard)
  int epoch;      // ready-flag value of this launch: call count << 3 | tier (never 0)
};

// ---------------------------------------------------------------------------
// offset store: WFA2's int32 offsets, indexed from the start of the ring /
// arena region (vector accessors load or store 4 consecutive offsets).
template <class K>
DI int ld_off(const Ctx<K>& c, int i) { return c.wd[(uint32_t)i]; }
template <class K>
DI void st_off(const Ctx<K>& c, int i, int v) { c.wd[(uint32_t)i] = v; }
// 4 offsets at a multiple-of-4 index (16-byte aligned: ring slots start at
// multiples of the ring width, global-tier arena wavefronts at multiples of 4)
template <class K>
DI int4 ld_off4(const Ctx<K>& c, int i) { return *reinterpret_cast<const int4*>(c.wd + (uint32_t)i); }
template <class K>
DI void st_off4(const Ctx<K>& c, int i, const int4& v) { *reinterpret_cast<int4*>(c.wd + (uint32_t)i) = v; }
template <class K>
DI int rd(const Ctx<K>& c, const Wf& w, int k) {
  return (k >= w.lo && k <= w.hi) ? ld_off(c, w.base + (k & w.mask)) : kOffsetNull;
}
// rd without a branch: every lane loads, at k clamped into [lo, hi] (for the
// null wavefront lo=1, hi=-1, mask 0: index base+0), and lanes outside take NULL.
template <class K>
DI int rd_clamped(const Ctx<K>& c, const Wf& w, int k) {
  const int kk = imin(imax(k, w.lo), w.hi);
  const int v = ld_off(c, w.base + (kk & w.mask));
  return ((k >= w.lo) & (k <= w.hi)) ? v : kOffsetNull;
}
// rd without a branch for a ring input: base + (k & mask) stays inside the
// ring slot for every k, so no clamp is needed.
template <class K>
DI int rd_ring(const Ctx<K>& c, const Wf& w, int k) {
  const int v = ld_off(c, w.base + (k & w.mask));
  return ((k >= w.lo) & (k <= w.hi)) ? v : kOffsetNull;
}

// Upper bound of the optimal score of a plen x tlen sub-problem plus the
// slack the breakpoint search may run past it; a search that goes beyond it
// can only be a defect, and is stopped with kStGuard instead of looping.
template <class K>
DI int score_guard(const Ctx<K>& c, int plen, int tlen) {
  const int mn = imin(plen, tlen), df = imax(plen, tlen) - mn;
  return mn * c.x + c.o + df * c.e + c.o + 2 * c.scope + 16;
}

// Register-array access with a run-time component index, written as masking
// arithmetic rather than a select: a select of the three elements would be
// folded into a select of their addresses and pin the aligner state to scratch.
DI int sel3(const int (&a)[3], int comp) {
  return (a[0] & -(int)(comp == CM)) | (a[1] & -(int)(comp == CI)) | (a[2] & -(int)(comp == CD));
}

// ring slot of (score - delta) given the slot of score; 0 <= delta < scope
DI int slot_back(int slot, int delta, int scope) {
  const int r = slot - delta;
  return r < 0 ? r + scope : r;
}

// ---------------------------------------------------------------------------
// team reductions (kNW > 1).  Every wave of a team runs the same uniform
// control flow on identical copies of the aligner state; values that depend on
// the diagonals a wave owns are combined through a double-buffered LDS area
// (kNW x kRedSlots ints per phase).  Two phases suffice: a wave can only write
// a phase again after passing the barrier of the other phase, which every wave
// reaches only once its reads of the first are done.

template <class K>
DI int32_t* team_buf(Ctx<K>& c) {
  int32_t* b = c.tred + c.tph * (K::kNW * kRedSlots);
  c.tph ^= 1;
  return b;
}

// In-place max over the team of eight wave-uniform values (one barrier).
// Lane j of every wave publishes value j; after the barrier lane j folds the
// waves' copies of value j.  The lane's value is picked with masking
// arithmetic: a select chain would become a select of the arguments' addresses
// and move them to scratch.
template <class K>
DI void team_max8(Ctx<K>& c, int& v0, int& v1, int& v2, int& v3, int& v4, int& v5, int& v6, int& v7) {
  int32_t* b = team_buf(c);
  const int l = c.lane;
  const int a0 = v0, a1 = v1, a2 = v2, a3 = v3, a4 = v4, a5 = v5, a6 = v6, a7 = v7;
  const int mine = (a0 & -(int)(l == 0)) | (a1 & -(int)(l == 1)) | (a2 & -(int)(l == 2)) |
                   (a3 & -(int)(l == 3)) | (a4 & -(int)(l == 4)) | (a5 & -(int)(l == 5)) |
                   (a6 & -(int)(l == 6)) | (a7 & -(int)(l == 7));
  if (l < 8) b[c.wid * kRedSlots + l] = mine;
  __syncthreads();
  int r = INT_MIN;
  if (l < 8) {
#pragma unroll
    for (int w = 0; w < K::kNW; ++w) r = imax(r, b[w * kRedSlots + l]);
  }
  v0 = rdl(r, 0); v1 = rdl(r, 1); v2 = rdl(r, 2); v3 = rdl(r, 3);
  v4 = rdl(r, 4); v5 = rdl(r, 5); v6 = rdl(r, 6); v7 = rdl(r, 7);
}
// In-place max over the team of sixteen wave-uniform values (one barrier), as
// team_max8 with lanes 0..15 publishing.
template <class K>
DI void team_max16(Ctx<K>& c, int (&v)[16]) {
  int32_t* b = team_buf(c);
  const int l = c.lane;
  int mine = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) mine |= v[j] & -(int)(l == j);
  if (l < 16) b[c.wid * kRedSlots + l] = mine;
  __syncthreads();
  int r = INT_MIN;
  if (l < 16) {
#pragma unroll
    for (int w = 0; w < K::kNW; ++w) r = imax(r, b[w * kRedSlots + l]);
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = rdl(r, j);
}
template <class K>
DI int team_max1(Ctx<K>& c, int v) {
  int32_t* b = team_buf(c);
  if (c.lane == 0) b[c.wid * kRedSlots] = v;
  __syncthreads();
  int r = INT_MIN;
  if (c.lane < K::kNW) r = b[c.lane * kRedSlots];
  r = imax(r, __builtin_amdgcn_update_dpp(INT_MIN, r, 0x111, 0xf, 0xf, false));  // row_shr:1
  r = imax(r, __builtin_amdgcn_update_dpp(INT_MIN, r, 0x112, 0xf, 0xf, false));  // row_shr:2
  r = imax(r, __builtin_amdgcn_update_dpp(INT_MIN, r, 0x114, 0xf, 0xf, false));  // row_shr:4
  r = imax(r, __builtin_amdgcn_update_dpp(INT_MIN, r, 0x118, 0xf, 0xf, false));  // row_shr:8
  return rdl(r, 15);
}

// Min over the team of the values the waves stored at slot j of `buf`
// (lanes 0..15 load one wave's value each, then a row-wide DPP scan).
template <class K>
DI int team_min_slot(const Ctx<K>& c, const int32_t* buf, int j) {
  int v = (c.lane < K::kNW) ? buf[c.lane * kRedSlots + j] : INT_MAX;
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = imin(v, __builtin_amdgcn_update_dpp(INT_MAX, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  return rdl(v, 15);
}

// ---------------------------------------------------------------------------
// headers

// Start of the ring / base-case arena, as an index into the offset store: the
// layout puts the offsets at the workspace base (the host checks L.data == 0),
// so it is the compile-time constant 0 (every uniform value the kernels keep out
// of their scalar registers helps: they run close to the scalar register limit).
template <class K>
DI int ws_data(const Ctx<K>&) { return 0; }
// ring width (a power of two): compile-time in the default LDS configuration
template <class K>
DI int wcap(const Ctx<K>& c) { return K::kW > 0 ? K::kW : c.L.wcap; }
// base-case arena capacity: the LDS tier's arena is its two rings' space
// (2 x scope x 3 x wcap), compile-time when both the width and the penalties are
template <class K>
DI int arena_cap(const Ctx<K>& c) {
  return (K::kLds && K::kW > 0 && K::kDefPen) ? 2 * kDefScope * 3 * K::kW : c.L.arena_cap;
}

// Offsets in a wave's private region of the backtrace run buffer and the
// recursion stack (compile-time in the LDS kernel: its region holds no
// base-case headers, so the runs start at 0 and the stack after them).
template <class K>
DI int runs_off(const Ctx<K>& c) { return K::kLds ? 0 : c.L.runs_tmp; }
template <class K>
DI int stack_off(const Ctx<K>& c) { return K::kLds ? kLdsRcap : c.L.stack; }

template <class K>
DI int ring_base(const Ctx<K>& c, int ring, int slot, int comp) {
  return ws_data(c) + ((ring * c.scope + slot) * 3 + comp) * wcap(c);
}

// Presence of (ring slot | base score) for component `comp`: a header lo other
// than kAbsent in the slot's header lane (ring) or lane `score` (base case, LDS
// kernel).  No scalar presence masks: the kernels run close to their scalar
// register limit, and spilled scalars cost more than a lane read.
template <class K>
DI bool is_present(const Ctx<K>& c, const Dir& d, int score, int slot, int comp) {
  if (d.ring >= 0) return rdl(sel3(c.rlo, comp), d.ring * c.scope + slot) != kAbsent;
  return rdl(sel3(c.blo, comp), score) != kAbsent;
}

// Component `comp` at `score` (ring slot `slot`), WFA2's
// wavefront_compute_get_*wavefront (wavefront_compute.c:266-305).  *present
// reports a non-NULL WFA2 pointer (possibly with lo > hi).
template <class K>
DI Wf get_wf(const Ctx<K>& c, const Dir& d, int score, int slot, int comp, int* present) {
  *present = 0;
  if (score < 0) return wf_null();
  if (d.ring >= 0) {
    if (!is_present(c, d, score, slot, comp)) return wf_null();
    *present = 1;
    const int j = d.ring * c.scope + slot;
    const int lo = rdl(sel3(c.rlo, comp), j);
    const int hi = rdl(sel3(c.rhi, comp), j);
    if (lo > hi) return Wf{lo, hi, 0, 0};
    return Wf{lo, hi, ring_base(c, d.ring, slot, comp), wcap(c) - 1};
  }
  int lo, hi, kb;
  if (K::kLds) {
    if (!is_present(c, d, score, slot, comp)) return wf_null();
    lo = rdl(sel3(c.blo, comp), score);
    hi = rdl(sel3(c.bhi, comp), score);
    kb = rdl(sel3(c.bkb, comp), score);
  } else {
    const int h = c.L.bhdr + (score * 3 + comp) * 3;
    lo = uni(c.pv[h]);
    hi = uni(c.pv[h + 1]);
    kb = uni(c.pv[h + 2]);
  }
  if (lo == kAbsent) return wf_null();
  *present = 1;
  return Wf{lo, hi, kb, -1};
}


DI void lane_set(int (&a)[3], int comp, int lane, int j, int v) {
  const bool me = lane == j;
  a[0] = (me && comp == CM) ? v : a[0];
  a[1] = (me && comp == CI) ? v : a[1];
  a[2] = (me && comp == CD) ? v : a[2];
}


// lo == kAbsent marks an unallocated component
template <class K>
DI void set_hdr(Ctx<K>& c, const Dir& d, int score, int slot, int comp, int lo, int hi, int kb) {
  const bool on = lo != kAbsent;
  if (d.ring >= 0) {
    const int j = d.ring * c.scope + slot;
    lane_set(c.rlo, comp, c.lane, j, lo);   // kAbsent marks an absent component
    if (on) lane_set(c.rhi, comp, c.lane, j, hi);
  } else if (K::kLds) {
    lane_set(c.blo, comp, c.lane, score, lo);   // kAbsent marks an absent component
    if (on) {
      lane_set(c.bhi, comp, c.lane, score, hi);
      lane_set(c.bkb, comp, c.lane, score, kb);
    }
  } else {
    const int h = c.L.bhdr + (score * 3 + comp) * 3;
    c.pv[h] = lo;
    c.pv[h + 1] = hi;
    c.pv[h + 2] = kb;
    // register copy of the last 64 scores (lane score % 64) for the steps'
    // input gather; the backtrace reads the memory copy.  It reuses the ring
    // header registers, which no base case needs (a breakpoint search starts
    // by clearing its rings' presence bits).
    lane_set(c.rlo, comp, c.lane, score & (kWave - 1), lo);
    lane_set(c.rhi, comp, c.lane, score & (kWave - 1), hi);
    lane_set(c.bkb, comp, c.lane, score & (kWave - 1), kb);
  }
}

// ---------------------------------------------------------------------------
// sequences: 2-bit packed, base j at bits 2*(j%32) of word j/32

// The sequences are read as 32-bit halves of the packed words, 16 bases each:
// base j is half j / 16, bits 2 * (j % 16).  One funnel shift forms the 16-base
// window starting at base 16w + sh from the halves w and w + 1.
typedef const __attribute__((address_space(3))) uint32_t* lds_halves;
typedef const __attribute__((address_space(1))) uint32_t* global_halves;

DI uint32_t window16(uint32_t lo, uint32_t hi, int sh) {
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> (sh << 1));
}

// Longest common extension from (v,h) of the aligner's sub-problem, 16 bases
// per 32-bit compare, clamped to the sub-problem's end.  The four halves of an
// iteration are loaded before either window is formed, so the two sequence
// reads overlap.  Typed pointers keep the loads in the address space the
// pair's words live in (generic loads would wait on both the LDS and the
// memory counters).
template <bool kOneExit, class Ptr>
DI int lcp_words(Ptr P, Ptr T, const Dir& d, int v, int h);
template <class Ptr>
DI int lcp_words64(Ptr P, Ptr T, const Dir& d, int v, int h);
typedef const __attribute__((address_space(3))) uint64_t* lds_words;
typedef const __attribute__((address_space(1))) uint64_t* global_words;

// Team kernels compare 32 bases per 64-bit step instead (measured faster there:
// fewer loop trips on long matches, and no more registers than they spill now).
template <class K>
DI int lcp(const Ctx<K>& c, const Dir& d, int v, int h) {
  if constexpr (K::kNW > 1) {
    if (c.seq_lds) return lcp_words64((lds_words)c.P, (lds_words)c.T, d, v, h);
    return lcp_words64((global_words)c.P, (global_words)c.T, d, v, h);
  }
  if (K::kLds || c.seq_lds) return lcp_words<K::kLds>((lds_halves)c.P, (lds_halves)c.T, d, v, h);
  return lcp_words<false>((global_halves)c.P, (global_halves)c.T, d, v, h);
}

template <bool kOneExit, class Ptr>
DI int lcp_words(Ptr P, Ptr T, const Dir& d, int v, int h) {
  const int lim = imin(d.plen - v, d.tlen - h);
  // forward: windows start at P[pb+v+n], T[tb+h+n]; reverse: windows end at
  // P[pe-1-v-n], T[te-1-h-n] and the first mismatch is the highest differing base
  const bool rev = d.rev != 0;
  const int pp = rev ? d.pb + d.plen - 16 - v : d.pb + v;
  const int tp = rev ? d.tb + d.tlen - 16 - h : d.tb + h;
  // windows move 16 bases (one half) per iteration: the in-half shifts stay put
  const int sp = pp & 15, st = tp & 15, dw = rev ? -1 : 1;
  int wp = pp >> 4, wt = tp >> 4;
  int n = 0;
  for (;;) {
    const uint32_t p0 = P[wp], p1 = P[wp + 1];
    const uint32_t t0 = T[wt], t1 = T[wt + 1];
    const uint32_t x = window16(p0, p1, sp) ^ window16(t0, t1, st);
    if constexpr (kOneExit) {
      // One exit test per window: a lane leaves at its first mismatch or past
      // the limit (a mismatch-free window counts 16 bases), so the divergent
      // loop carries a single break mask instead of one per exit (fewer scalar
      // mask operations per window, in the kernel that is bound by issue).
      const int z = x == 0 ? 32 : rev ? __builtin_clz(x) : __builtin_ctz(x);
      n += z >> 1;
      if (x != 0 || n >= lim) break;
    } else {
      if (x) {
        n += (rev ? __builtin_clz(x) : __builtin_ctz(x)) >> 1;
        break;
      }
      n += 16;
      if (n >= lim) break;
    }
    wp += dw;
    wt += dw;
  }
  return imin(n, lim);
}

// The same over 32-base windows of two 64-bit words (word w, bits 2 * (j % 32)).
DI uint64_t window32(uint64_t lo, uint64_t hi, int sh) {
  const int b = sh << 1;
  return (lo >> b) | ((hi << 1) << (63 - b));
}

template <class Ptr>
DI int lcp_words64(Ptr P, Ptr T, const Dir& d, int v, int h) {
  const int lim = imin(d.plen - v, d.tlen - h);
  const bool rev = d.rev != 0;
  const int pp = rev ? d.pb + d.plen - 32 - v : d.pb + v;
  const int tp = rev ? d.tb + d.tlen - 32 - h : d.tb + h;
  const int sp = pp & 31, st = tp & 31, dw = rev ? -1 : 1;
  int wp = pp >> 5, wt = tp >> 5;
  int n = 0;
  for (;;) {
    const uint64_t p0 = P[wp], p1 = P[wp + 1];
    const uint64_t t0 = T[wt], t1 = T[wt + 1];
    const uint64_t x = window32(p0, p1, sp) ^ window32(t0, t1, st);
    if (x) {
      n += (rev ? __builtin_clzll(x) : __builtin_ctzll(x)) >> 1;
      break;
    }
    n += 32;
    if (n >= lim) break;
    wp += dw;
    wt += dw;
  }
  return imin(n, lim);
}

// Extension of two diagonals of a lane over their first 16 bases (the global
// tiers' vector steps call it for diagonals 0-1 and 2-3 of their four).  The
// sequences are read as 32-bit halves of the packed words, 16 bases each, and
// one funnel shift forms a window from two consecutive halves: base j is half
// j / 16, bits 2 * (j % 16).  All four reads are issued before either window is
// formed.  Returns the matches within the window, 16 when it matches completely
// (on random sequence almost never away from the alignment path); the caller
// continues those diagonals word by word from base 16.  (v, h) must be
// in-bounds positions (callers pass 0, 0 for diagonals not extended).

template <class Ptr>
DI void lcp_first2(Ptr P, Ptr T, const Dir& d, int v0, int h0, int v1, int h1, int& n0, int& n1) {
  const bool rev = d.rev != 0;
  const int pp0 = rev ? d.pb + d.plen - 16 - v0 : d.pb + v0;
  const int tp0 = rev ? d.tb + d.tlen - 16 - h0 : d.tb + h0;
  const int pp1 = rev ? d.pb + d.plen - 16 - v1 : d.pb + v1;
  const int tp1 = rev ? d.tb + d.tlen - 16 - h1 : d.tb + h1;
  const uint32_t a0 = P[pp0 >> 4], a1 = P[(pp0 >> 4) + 1], b0 = T[tp0 >> 4], b1 = T[(tp0 >> 4) + 1];
  const uint32_t c0 = P[pp1 >> 4], c1 = P[(pp1 >> 4) + 1], e0 = T[tp1 >> 4], e1 = T[(tp1 >> 4) + 1];
  const uint32_t x0 = window16(a0, a1, pp0 & 15) ^ window16(b0, b1, tp0 & 15);
  const uint32_t x1 = window16(c0, c1, pp1 & 15) ^ window16(e0, e1, tp1 & 15);
  n0 = x0 == 0 ? 16 : (rev ? __builtin_clz(x0) : __builtin_ctz(x0)) >> 1;
  n1 = x1 == 0 ? 16 : (rev ? __builtin_clz(x1) : __builtin_ctz(x1)) >> 1;
}

// ---------------------------------------------------------------------------
// compute: wavefront_compute_affine (wavefront_compute_affine.c:229-260)
// `slot` is score % scope for ring aligners.

template <class K>
DI void null_step(Ctx<K>& c, Dir& d, int s, int slot) {
  d.null_steps++;
  set_hdr(c, d, s, slot, CM, kAbsent, 0, 0);
  set_hdr(c, d, s, slot, CI, kAbsent, 0, 0);
  set_hdr(c, d, s, slot, CD, kAbsent, 0, 0);
}


// The four inputs of a step -- M[s-x], M[s-o-e], I1[s-e], D1[s-e] -- and the
// limits they give (wavefront_compute_limits_input, wavefront_compute.c:40-86),
// lane-parallel: lane q < 4 fetches input q's header from the header lanes
// (ds_bpermute), tests its presence, normalises an absent or empty input to the
// null wavefront (lo=1, hi=-1, which still takes part in the limits as in WFA2),
// and the limits are reduced over lanes 0..3.  Same results as reading the four
// inputs one by one (wavefront_compute_get_*wavefront, wavefront_compute.c:266-305)
// with empty inputs taken as null, with the per-input control in vector lanes
// instead of one scalar chain per input.  Ring headers live in the ring header
// lanes; base-case headers in the LDS kernel's per-score lanes, and in the
// global kernels in a register copy of the last 64 scores (see set_hdr).
// Returns false when all four inputs are null.  Ring inputs use mask wcap-1,
// base-case inputs mask -1; a null input's base keeps a clamped read inside the
// workspace.
// pick3 by the gather's per-lane component (CM, CI, CD for lanes 0-1, 2, 3)
template <class K>
DI int pick_q(const Ctx<K>& c, int v0, int v1, int v2, int cq) {
  if constexpr (K::kLds) return (v0 & c.qm0) | (v1 & c.qm1) | (v2 & c.qm2);
  return pick3(v0, v1, v2, cq);
}

template <class K>
DI bool gather_inputs(const Ctx<K>& c, const Dir& d, int s, int slot, Wf& mm, Wf& mo, Wf& ie, Wf& de,
                      int& lo, int& hi, bool& has_i, bool& has_d) {
  const int q = c.lane;
  // q = 3 shares s-e with q = 2
  const int delta = K::kLds ? c.qdelta : pick3(c.x, c.o + c.e, c.e, imin(q, 2));
  const int cq = K::kLds ? c.qcomp : q <= 1 ? CM : q == 2 ? CI : CD;
  const int sq = s - delta;
  const bool ring = d.ring >= 0;
  int hl, lq, hq, kq;
  if (ring) {
    int sl = slot - delta;
    sl += sl < 0 ? c.scope : 0;
    hl = d.ring * c.scope + sl;
    const int hb = (hl & (kWave - 1)) << 2;
    lq = pick_q(c, __builtin_amdgcn_ds_bpermute(hb, c.rlo[CM]), __builtin_amdgcn_ds_bpermute(hb, c.rlo[CI]),
               __builtin_amdgcn_ds_bpermute(hb, c.rlo[CD]), cq);
    hq = pick_q(c, __builtin_amdgcn_ds_bpermute(hb, c.rhi[CM]), __builtin_amdgcn_ds_bpermute(hb, c.rhi[CI]),
               __builtin_amdgcn_ds_bpermute(hb, c.rhi[CD]), cq);
    kq = ws_data(c) + (hl * 3 + cq) * wcap(c);
  } else if constexpr (K::kLds) {
    hl = sq;
    const int hb = (hl & (kWave - 1)) << 2;
    lq = pick_q(c, __builtin_amdgcn_ds_bpermute(hb, c.blo[CM]), __builtin_amdgcn_ds_bpermute(hb, c.blo[CI]),
               __builtin_amdgcn_ds_bpermute(hb, c.blo[CD]), cq);
    hq = pick_q(c, __builtin_amdgcn_ds_bpermute(hb, c.bhi[CM]), __builtin_amdgcn_ds_bpermute(hb, c.bhi[CI]),
               __builtin_amdgcn_ds_bpermute(hb, c.bhi[CD]), cq);
    kq = pick_q(c, __builtin_amdgcn_ds_bpermute(hb, c.bkb[CM]), __builtin_amdgcn_ds_bpermute(hb, c.bkb[CI]),
               __builtin_amdgcn_ds_bpermute(hb, c.bkb[CD]), cq);
  } else {
    // global workspace: base-case headers of the last 64 scores in register
    // lanes (lane score % 64, lo = kAbsent when absent; see set_hdr); a
    // step's inputs are at most scope-1 < 64 scores back and every score in
    // between was written
    hl = sq;
    const int hb = (hl & (kWave - 1)) << 2;
    lq = pick_q(c, __builtin_amdgcn_ds_bpermute(hb, c.rlo[CM]), __builtin_amdgcn_ds_bpermute(hb, c.rlo[CI]),
               __builtin_amdgcn_ds_bpermute(hb, c.rlo[CD]), cq);
    hq = pick_q(c, __builtin_amdgcn_ds_bpermute(hb, c.rhi[CM]), __builtin_amdgcn_ds_bpermute(hb, c.rhi[CI]),
               __builtin_amdgcn_ds_bpermute(hb, c.rhi[CD]), cq);
    kq = pick_q(c, __builtin_amdgcn_ds_bpermute(hb, c.bkb[CM]), __builtin_amdgcn_ds_bpermute(hb, c.bkb[CI]),
               __builtin_amdgcn_ds_bpermute(hb, c.bkb[CD]), cq);
  }
  // Only lanes 0..3 are read below (their readlanes, the ballot's low 4 bits
  // and the limits reduced into lane 3), so lanes 4.. need no masking.
  const bool present = sq >= 0 && lq != kAbsent;   // a header lo other than kAbsent
  const bool nonempty = present && lq <= hq;
  if (!nonempty) {
    lq = 1;
    hq = -1;
    kq = ring ? 0 : 1;
  }
  // limits: M[s-x] [lo, hi], M[s-o-e] [lo-1, hi+1], I1[s-e] [lo+1, hi+1], D1[s-e] [lo-1, hi-1]
  int lc, hc;
  if constexpr (K::kLds) {
    lc = lq + c.qlo;
    hc = hq + c.qhi;
  } else {
    lc = lq + (q == 0 ? 0 : q == 2 ? 1 : -1);
    hc = hq + (q == 0 ? 0 : q == 3 ? -1 : 1);
  }
  lc = imin(lc, __builtin_amdgcn_update_dpp(INT_MAX, lc, 0x111, 0xf, 0xf, false));  // row_shr:1
  hc = imax(hc, __builtin_amdgcn_update_dpp(INT_MIN, hc, 0x111, 0xf, 0xf, false));
  lc = imin(lc, __builtin_amdgcn_update_dpp(INT_MAX, lc, 0x112, 0xf, 0xf, false));  // row_shr:2
  hc = imax(hc, __builtin_amdgcn_update_dpp(INT_MIN, hc, 0x112, 0xf, 0xf, false));
  const uint32_t ne = (uint32_t)__ballot(nonempty) & 0xfu;
  const int mk = ring ? wcap(c) - 1 : -1;
  mm = Wf{rdl(lq, 0), rdl(hq, 0), rdl(kq, 0), mk};
  mo = Wf{rdl(lq, 1), rdl(hq, 1), rdl(kq, 1), mk};
  ie = Wf{rdl(lq, 2), rdl(hq, 2), rdl(kq, 2), mk};
  de = Wf{rdl(lq, 3), rdl(hq, 3), rdl(kq, 3), mk};
  lo = rdl(lc, 3);
  hi = rdl(hc, 3);
  has_i = (ne & 0x6u) != 0;
  has_d = (ne & 0xau) != 0;
  return ne != 0;
}

// Lane i receives lane i-1's value, lane 0 `first` (DPP wave_shr:1).
DI int from_prev_lane(int v, int first) { return __builtin_amdgcn_update_dpp(first, v, 0x138, 0xf, 0xf, false); }
// Lane i receives lane i+1's value, lane 63 `last` (DPP wave_shl:1).
DI int from_next_lane(int v, int last) { return __builtin_amdgcn_update_dpp(last, v, 0x130, 0xf, 0xf, false); }
DI int v4get(const int4& v, int j) { return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w; }
// a[j] for a lane-varying j, as masking arithmetic (a select of register-array
// elements would become a select of their addresses and pin the array to scratch)
DI int pick4(const int (&a)[4], int j) {
  return (a[0] & -(int)(j == 0)) | (a[1] & -(int)(j == 1)) | (a[2] & -(int)(j == 2)) | (a[3] & -(int)(j == 3));
}

// Per-lane results of a step's diagonals: first/last in-bounds diagonal of
// M, I1, D1 (trims), maximum antidiagonal, and the end component's offset at
// the end diagonal (INT_MIN unless this lane holds it).
struct StepAcc {
  int fm, lm, fi, li, fd, ld, lmax, endv;
};

// Index of the 4 offsets of input w at diagonals k0 .. k0+3 (k0 a multiple of
// 4): in a ring they stay in the slot for any k0; in the arena k0 is clamped
// into the input's allocation (whole 4-offset groups from lo & ~3 to hi | 3),
// and a null input reads element 0.  Values outside [lo, hi] are masked later.
DI int in4_index(const Wf& w, int k0, bool ring) {
  if (ring) return w.base + (k0 & w.mask);
  if (w.lo > w.hi) return 0;
  return w.base + imin(imax(k0, w.lo & ~3), w.hi & ~3);
}
DI int4 mask4(int4 v, const Wf& w, int k0) {
  v.x = (k0 >= w.lo && k0 <= w.hi) ? v.x : kOffsetNull;
  v.y = (k0 + 1 >= w.lo && k0 + 1 <= w.hi) ? v.y : kOffsetNull;
  v.z = (k0 + 2 >= w.lo && k0 + 2 <= w.hi) ? v.z : kOffsetNull;
  v.w = (k0 + 3 >= w.lo && k0 + 3 <= w.hi) ? v.w : kOffsetNull;
  return v;
}

// A global-tier step over [lo, hi] in tiles of 256 diagonals: lane l of the
// wave that owns a tile computes diagonals kt + 4l .. kt + 4l + 3
// (wavefront_compute_affine_idm, wavefront_compute_affine.c:65-85, then the
// extension) from one 16-byte read per input, with the diagonal-minus-one and
// diagonal-plus-one inputs of its outer diagonals taken from the neighbouring
// lanes (DPP wave shifts) and, at the tile edges, from one read each.  Tiles
// start at lo & ~3 and go to the team's waves in turn.  Same results as the
// one-diagonal-per-lane loop: identical arithmetic per diagonal; reads outside
// an input's [lo, hi] are NULL; writes cover whole 4-offset groups, so the
// caller guarantees they alias nothing live (ring: (hi|3) - (lo&~3) < wcap;
// arena: the output allocation spans those groups).
template <class K, bool kMax>
DI void step_tiles(const Ctx<K>& c, const Dir& d, int lo, int hi, const Wf& mm, const Wf& mo, const Wf& ie,
                   const Wf& de, bool has_i, bool has_d, int bm, int bi, int bd, int mask, int wid, int nw,
                   StepAcc& acc) {
  const bool ring = d.ring >= 0;
  const uint32_t tl = (uint32_t)d.tlen, pl = (uint32_t)d.plen;
  const int k_end = d.tlen - d.plen;
  const int ce = d.ce;
  const int t0 = lo & ~3;
  constexpr int kTile = 4 * kWave;
  for (int kt = t0 + wid * kTile; kt <= hi; kt += nw * kTile) {
    const int k0 = kt + 4 * c.lane;
    // every input read issued before any is used
    int4 vm = ld_off4(c, in4_index(mm, k0, ring));
    int4 vo = ld_off4(c, in4_index(mo, k0, ring));
    int4 vi = ld_off4(c, in4_index(ie, k0, ring));
    int4 vd = ld_off4(c, in4_index(de, k0, ring));
    const int o_prev = ring ? rd_ring(c, mo, kt - 1) : rd_clamped(c, mo, kt - 1);
    const int i_prev = ring ? rd_ring(c, ie, kt - 1) : rd_clamped(c, ie, kt - 1);
    const int o_next = ring ? rd_ring(c, mo, kt + kTile) : rd_clamped(c, mo, kt + kTile);
    const int d_next = ring ? rd_ring(c, de, kt + kTile) : rd_clamped(c, de, kt + kTile);
    // a tile strictly inside every input needs no masks (a null input is never inside)
    const bool inner = mm.lo <= kt && kt + kTile - 1 <= mm.hi && mo.lo <= kt - 1 && kt + kTile <= mo.hi &&
                       ie.lo <= kt - 1 && kt + kTile - 2 <= ie.hi && de.lo <= kt + 1 && kt + kTile <= de.hi;
    if (!inner) {
      vm = mask4(vm, mm, k0);
      vo = mask4(vo, mo, k0);
      vi = mask4(vi, ie, k0);
      vd = mask4(vd, de, k0);
    }
    // neighbours across lanes (all lanes active here)
    const int o_m1 = from_prev_lane(vo.w, o_prev), i_m1 = from_prev_lane(vi.w, i_prev);
    const int o_p1 = from_next_lane(vo.x, o_next), d_p1 = from_next_lane(vd.x, d_next);
    int ins1[4], del1[4], mx[4], vv[4], hh[4], ext[4];
    bool act[4], live[4];
    ins1[0] = imax(o_m1, i_m1) + 1;
    ins1[1] = imax(vo.x, vi.x) + 1;
    ins1[2] = imax(vo.y, vi.y) + 1;
    ins1[3] = imax(vo.z, vi.z) + 1;
    del1[0] = imax(vo.y, vd.y);
    del1[1] = imax(vo.z, vd.z);
    del1[2] = imax(vo.w, vd.w);
    del1[3] = imax(o_p1, d_p1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + j;
      mx[j] = imax(del1[j], imax(v4get(vm, j) + 1, ins1[j]));
      act[j] = k >= lo && k <= hi;
      live[j] = act[j] && !((uint32_t)mx[j] > tl || (uint32_t)(mx[j] - k) > pl);
      vv[j] = live[j] ? mx[j] - k : 0;
      hh[j] = live[j] ? mx[j] : 0;
    }
    if (c.seq_lds) {
      lcp_first2((lds_halves)c.P, (lds_halves)c.T, d, vv[0], hh[0], vv[1], hh[1], ext[0], ext[1]);
      lcp_first2((lds_halves)c.P, (lds_halves)c.T, d, vv[2], hh[2], vv[3], hh[3], ext[2], ext[3]);
    } else {
      lcp_first2((global_halves)c.P, (global_halves)c.T, d, vv[0], hh[0], vv[1], hh[1], ext[0], ext[1]);
      lcp_first2((global_halves)c.P, (global_halves)c.T, d, vv[2], hh[2], vv[3], hh[3], ext[2], ext[3]);
    }
    // a diagonal whose first window matched completely continues word by word
    // (rare away from the alignment path); one diagonal at a time per lane
    int more = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      more |= (live[j] && ext[j] == 16 && imin(d.plen - vv[j], d.tlen - hh[j]) > 16) ? (1 << j) : 0;
    while (more) {
      const int j = __builtin_ctz(more);
      more &= more - 1;
      const int v = pick4(vv, j), h = pick4(hh, j);
      const int r = 16 + lcp(c, d, v + 16, h + 16);
      ext[0] = j == 0 ? r : ext[0];
      ext[1] = j == 1 ? r : ext[1];
      ext[2] = j == 2 ? r : ext[2];
      ext[3] = j == 3 ? r : ext[3];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) ext[j] = imin(ext[j], imin(d.plen - vv[j], d.tlen - hh[j]));
    // in-bounds flags for wavefront_compute_trim_ends (wavefront_compute.c:579-613)
    bool in_i[4], in_d[4];
    bool all_in = true;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + j;
      if (live[j]) {
        mx[j] += ext[j];
        if constexpr (kMax) acc.lmax = imax(acc.lmax, 2 * mx[j] - k);
      } else {
        mx[j] = kOffsetNull;
      }
      in_i[j] = act[j] && (uint32_t)ins1[j] <= tl && (uint32_t)(ins1[j] - k) <= pl;
      in_d[j] = act[j] && (uint32_t)del1[j] <= tl && (uint32_t)(del1[j] - k) <= pl;
      all_in = all_in && live[j] && in_i[j] && in_d[j];
      if (k == k_end && act[j]) acc.endv = (ce == CM) ? mx[j] : (ce == CI) ? ins1[j] : del1[j];
    }
    if (all_in) {   // every component in bounds on all four diagonals (interior tiles)
      acc.fm = imin(acc.fm, k0); acc.lm = imax(acc.lm, k0 + 3);
      acc.fi = imin(acc.fi, k0); acc.li = imax(acc.li, k0 + 3);
      acc.fd = imin(acc.fd, k0); acc.ld = imax(acc.ld, k0 + 3);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = k0 + j;
        acc.fm = imin(acc.fm, live[j] ? k : INT_MAX); acc.lm = imax(acc.lm, live[j] ? k : INT_MIN);
        acc.fi = imin(acc.fi, in_i[j] ? k : INT_MAX); acc.li = imax(acc.li, in_i[j] ? k : INT_MIN);
        acc.fd = imin(acc.fd, in_d[j] ? k : INT_MAX); acc.ld = imax(acc.ld, in_d[j] ? k : INT_MIN);
      }
    }
    if (k0 <= hi) {   // whole groups: positions past hi (up to hi | 3) are never read
      const int o = ring ? (k0 & mask) : k0;
      st_off4(c, bm + o, int4{mx[0], mx[1], mx[2], mx[3]});
      if (has_i) st_off4(c, bi + o, int4{ins1[0], ins1[1], ins1[2], ins1[3]});
      if (has_d) st_off4(c, bd + o, int4{del1[0], del1[1], del1[2], del1[3]});
    }
  }
}

// The whole step in one 64-diagonal chunk (the executing wave only):
// branch-free input reads, trims straight from the ballots.  Writes this
// wave's trims, maximum antidiagonal and end-diagonal offset.
template <class K, bool kMax>
DI void step_chunk(const Ctx<K>& c, const Dir& d, int lo, int hi, const Wf& mm, const Wf& mo, const Wf& ie,
                   const Wf& de, bool has_i, bool has_d, int bm, int bi, int bd, int mask, int& fm, int& lm,
                   int& fi, int& li, int& fd, int& ld, int& lmax, int& endu) {
  const uint32_t tl = (uint32_t)d.tlen, pl = (uint32_t)d.plen;
  const int k_end = d.tlen - d.plen;
  const int ce = d.ce;
  const int k = lo + c.lane;
  const bool act = k <= hi;
  int ins1, del1, mx;
  if (d.ring >= 0) {
    ins1 = imax(rd_ring(c, mo, k - 1), rd_ring(c, ie, k - 1)) + 1;
    del1 = imax(rd_ring(c, mo, k + 1), rd_ring(c, de, k + 1));
    mx = imax(del1, imax(rd_ring(c, mm, k) + 1, ins1));
  } else {
    ins1 = imax(rd_clamped(c, mo, k - 1), rd_clamped(c, ie, k - 1)) + 1;
    del1 = imax(rd_clamped(c, mo, k + 1), rd_clamped(c, de, k + 1));
    mx = imax(del1, imax(rd_clamped(c, mm, k) + 1, ins1));
  }
  const bool live = act && !((uint32_t)mx > tl || (uint32_t)(mx - k) > pl);
  int ak = 0;
  if (live) {                                   // extension along the diagonal
    mx += lcp(c, d, mx - k, mx);
    ak = 2 * mx - k;
  }
  mx = live ? mx : kOffsetNull;
  if (act) {
    st_off(c, bm + (k & mask), mx);
    if (has_i) st_off(c, bi + (k & mask), ins1);
    if (has_d) st_off(c, bd + (k & mask), del1);
  }
  const uint64_t b_m = __ballot(live);
  const uint64_t b_i = __ballot(act && (uint32_t)ins1 <= tl && (uint32_t)(ins1 - k) <= pl);
  const uint64_t b_d = __ballot(act && (uint32_t)del1 <= tl && (uint32_t)(del1 - k) <= pl);
  if (b_m) { fm = lo + __builtin_ctzll(b_m); lm = lo + 63 - __builtin_clzll(b_m); }
  if (b_i) { fi = lo + __builtin_ctzll(b_i); li = lo + 63 - __builtin_clzll(b_i); }
  if (b_d) { fd = lo + __builtin_ctzll(b_d); ld = lo + 63 - __builtin_clzll(b_d); }
  const int ev = (ce == CM) ? mx : (ce == CI) ? ins1 : del1;
  if (k_end >= lo && k_end <= hi) endu = rdl(ev, k_end - lo);
  if constexpr (kMax) lmax = wave_max_nonneg(ak);
}

// A global-tier step in tiles (wave `wid` of `nw` taking every nw-th tile),
// reduced over the wave: trims, maximum antidiagonal and end-diagonal offset.
template <class K, bool kMax>
DI void step_tiles_reduced(const Ctx<K>& c, const Dir& d, int lo, int hi, const Wf& mm, const Wf& mo,
                           const Wf& ie, const Wf& de, bool has_i, bool has_d, int bm, int bi, int bd, int mask,
                           int wid, int nw, int& fm, int& lm, int& fi, int& li, int& fd, int& ld, int& lmax,
                           int& endu) {
  StepAcc acc{INT_MAX, INT_MIN, INT_MAX, INT_MIN, INT_MAX, INT_MIN, 0, INT_MIN};
  step_tiles<K, kMax>(c, d, lo, hi, mm, mo, ie, de, has_i, has_d, bm, bi, bd, mask, wid, nw, acc);
  fm = wave_min_any(acc.fm); lm = wave_max_any(acc.lm);
  fi = wave_min_any(acc.fi); li = wave_max_any(acc.li);
  fd = wave_min_any(acc.fd); ld = wave_max_any(acc.ld);
  // offset at k_end: lane ((k_end - (lo & ~3)) / 4) % 64 of the wave owning its tile
  const int k_end = d.tlen - d.plen;
  if (k_end >= lo && k_end <= hi) endu = rdl(acc.endv, ((k_end - (lo & ~3)) >> 2) & (kWave - 1));
  if constexpr (kMax) lmax = wave_max_nonneg(acc.lmax);
}

// The end of a step once the team agrees on its values: trimmed limits into
// the headers (wavefront_compute_trim_ends, wavefront_compute.c:579-613) and
// the termination test (wavefront_termination_end2end, wavefront_termination.c:37-114).
// Returns 1 when the end is reached; *max_ak gets the maximum antidiagonal.
template <class K>
DI int apply_step(Ctx<K>& c, Dir& d, int s, int slot, int lo, bool has_i, bool has_d, int bm, int bi, int bd,
                  int fm, int lm, int fi, int li, int fd, int ld, int lmax, int endu, int* max_ak) {
  // trimmed limits: hi = last in-bounds (else lo-1), lo = first in-bounds (else lo)
  if (lm == INT_MIN) { fm = lo; lm = lo - 1; }
  if (li == INT_MIN) { fi = lo; li = lo - 1; }
  if (ld == INT_MIN) { fd = lo; ld = lo - 1; }
  set_hdr(c, d, s, slot, CM, fm, lm, bm);
  set_hdr(c, d, s, slot, CI, has_i ? fi : kAbsent, li, bi);
  set_hdr(c, d, s, slot, CD, has_d ? fd : kAbsent, ld, bd);
  // wavefront_termination_end2end: the end component holds k_end with offset >= tlen
  const int k_end = d.tlen - d.plen;
  const int ce = d.ce;
  const bool ce_present = ce == CM || (ce == CI ? has_i : has_d);
  const int ce_lo = ce == CM ? fm : ce == CI ? fi : fd;
  const int ce_hi = ce == CM ? lm : ce == CI ? li : ld;
  if (ce_present && ce_lo <= k_end && k_end <= ce_hi && endu >= d.tlen) {
    d.status = kStEndReached;
    d.status_score = s;
    return 1;
  }
  *max_ak = lmax;
  return 0;
}

// One WFA step at score s: wavefront_compute_affine (wavefront_compute_affine.c:229-260)
// followed by wavefront_extend_end2end(_max) (wavefront_extend.c:86-211) and the
// termination test (wavefront_termination.c:37-114), fused: the lane that
// computes M[s][k] extends it straight away, so the new wavefront is written
// once and the trims, the maximum antidiagonal and the end-diagonal offset are
// combined in a single reduction.  Extension only reads M[s][k] and the
// sequences, and the trimmed-away diagonals are NULL, so this is the same
// computation as compute-all-then-extend.  Returns 1 when the aligner is done
// (end reached, or END_UNREACHABLE after scope+1 null steps); *max_ak gets the
// maximum antidiagonal 2*offset-k over the extended cells.
template <class K, bool kMax = true>
DI int compute_extend(Ctx<K>& c, Dir& d, int s, int slot, int* max_ak) {
  *max_ak = 0;
  if (d.ring < 0 && s >= (K::kLds ? kLdsHcap : c.L.hcap)) { c.spill = 1; return 0; }
  bool null = (s & c.gmask) != 0;
  // With penalties that share a power-of-two factor g, only multiples of g are
  // reachable scores: the other steps are null without looking at the inputs.
  Wf mm, mo, ie, de;
  int lo = 0, hi = 0;
  bool has_i = false, has_d = false;
  if (!null) null = !gather_inputs(c, d, s, slot, mm, mo, ie, de, lo, hi, has_i, has_d);
  if (null) {                                   // null step (wavefront_compute_affine.c:245-251)
    if (K::kLds && (s & c.gmask) != 0) {
      // A score that is not a multiple of g is never read in a single-wave
      // aligner: step inputs are s minus multiples of g, the overlap test and
      // the backtrace skip such scores.  Its headers are left as they are.
      d.null_steps++;
    } else {
      null_step(c, d, s, slot);
    }
    if (d.null_steps > c.scope) {               // nothing left to extend (wavefront_extend.c)
      d.status = kStEndUnreachable;
      d.status_score = s;
      return 1;
    }
    return 0;
  }
  d.null_steps = 0;
  const int width = hi - lo + 1;
  int bm, bi, bd, mask;
  if (d.ring >= 0) {
    if (width > wcap(c)) { c.spill = 1; return 0; }
    bm = ring_base(c, d.ring, slot, CM);
    bi = ring_base(c, d.ring, slot, CI);
    bd = ring_base(c, d.ring, slot, CD);
    mask = wcap(c) - 1;
  } else if constexpr (K::kLds) {
    const int need = width * (1 + (int)has_i + (int)has_d);
    if (c.arena_top + need > arena_cap(c)) { c.spill = 1; return 0; }
    int top = ws_data(c) + c.arena_top;
    bm = top - lo; top += width;
    bi = top - lo; if (has_i) top += width;
    bd = top - lo; if (has_d) top += width;
    c.arena_top = top - ws_data(c);
    mask = -1;
  } else {
    // global tiers: each component spans whole 4-offset groups (lo & ~3 to
    // hi | 3) from a multiple-of-4 index, for the 16-byte reads and writes
    const int a0 = lo & ~3, span = (hi | 3) - a0 + 1;
    int top = (c.arena_top + 3) & ~3;
    const int need = span * (1 + (int)has_i + (int)has_d);
    if (top + need > arena_cap(c)) { c.spill = 1; return 0; }
    bm = top - a0; top += span;
    bi = top - a0; if (has_i) top += span;
    bd = top - a0; if (has_d) top += span;
    c.arena_top = top;
    mask = -1;
  }
  const uint32_t tl = (uint32_t)d.tlen, pl = (uint32_t)d.plen;
  const int k_end = d.tlen - d.plen;
  const int ce = d.ce;
  int fm = INT_MAX, lm = INT_MIN, fi = INT_MAX, li = INT_MIN, fd = INT_MAX, ld = INT_MIN;
  int lmax = 0, endv = INT_MIN, endu = INT_MIN;
  if (width <= kWave) {
    // One chunk holds the whole wavefront (in the LDS tier always, for ring
    // widths up to 64; elsewhere the narrow steps of deep recursion levels and
    // base cases).  In a team only wave 0 works; the others add neutral values
    // to the team reduction.
    if (K::kNW == 1 || c.wid == 0)
      step_chunk<K, kMax>(c, d, lo, hi, mm, mo, ie, de, has_i, has_d, bm, bi, bd, mask, fm, lm, fi, li, fd, ld,
                          lmax, endu);
  } else if (!K::kLds && (d.ring < 0 || (hi | 3) - (lo & ~3) + 1 <= wcap(c))) {
    step_tiles_reduced<K, kMax>(c, d, lo, hi, mm, mo, ie, de, has_i, has_d, bm, bi, bd, mask, c.wid, K::kNW, fm, lm,
                                fi, li, fd, ld, lmax, endu);
  } else {
  // A wave takes kG of its chunks at a time and issues all their loads before
  // any store: the stores of one chunk could alias the next chunk's inputs as
  // far as the compiler knows, so without the grouping every chunk would wait
  // out a full memory latency on its own.  (The LDS ring is fast enough and
  // narrow enough for one chunk at a time.)
  constexpr int kG = K::kLds ? 1 : 2;
  constexpr int kStride = K::kNW * kWave;
  for (int kb0 = lo + c.wid * kWave; kb0 <= hi; kb0 += kG * kStride) {
    int ins1[kG], del1[kG], misms[kG];
#pragma unroll
    for (int g = 0; g < kG; ++g) {
      const int k = kb0 + g * kStride + c.lane;
      // wavefront_compute_affine_idm (wavefront_compute_affine.c:65-85)
      ins1[g] = imax(rd(c, mo, k - 1), rd(c, ie, k - 1)) + 1;
      del1[g] = imax(rd(c, mo, k + 1), rd(c, de, k + 1));
      misms[g] = rd(c, mm, k) + 1;
    }
#pragma unroll
    for (int g = 0; g < kG; ++g) {
      const int kb = kb0 + g * kStride;
      if (kb > hi) break;
      const int k = kb + c.lane;
      const bool act = k <= hi;
      int mx = imax(del1[g], imax(misms[g], ins1[g]));
      const bool live = act && !((uint32_t)mx > tl || (uint32_t)(mx - k) > pl);
      if (live) {                                 // extension along the diagonal
        mx += lcp(c, d, mx - k, mx);
        lmax = imax(lmax, 2 * mx - k);
      } else {
        mx = kOffsetNull;
      }
      if (act) {
        st_off(c, bm + (k & mask), mx);
        if (has_i) st_off(c, bi + (k & mask), ins1[g]);
        if (has_d) st_off(c, bd + (k & mask), del1[g]);
        if (k == k_end) endv = (ce == CM) ? mx : (ce == CI) ? ins1[g] : del1[g];
      }
      // in-bounds flags for wavefront_compute_trim_ends (wavefront_compute.c:579-613)
      const bool in_i = act && (uint32_t)ins1[g] <= tl && (uint32_t)(ins1[g] - k) <= pl;
      const bool in_d = act && (uint32_t)del1[g] <= tl && (uint32_t)(del1[g] - k) <= pl;
      if constexpr (K::kLds) {
        const uint64_t b_m = __ballot(live), b_i = __ballot(in_i), b_d = __ballot(in_d);
        if (b_m) { if (fm == INT_MAX) fm = kb + __builtin_ctzll(b_m); lm = kb + 63 - __builtin_clzll(b_m); }
        if (b_i) { if (fi == INT_MAX) fi = kb + __builtin_ctzll(b_i); li = kb + 63 - __builtin_clzll(b_i); }
        if (b_d) { if (fd == INT_MAX) fd = kb + __builtin_ctzll(b_d); ld = kb + 63 - __builtin_clzll(b_d); }
      } else {
        // per-lane extremes, reduced once after the loop: no scalar work per chunk
        fm = imin(fm, live ? k : INT_MAX); lm = imax(lm, live ? k : INT_MIN);
        fi = imin(fi, in_i ? k : INT_MAX); li = imax(li, in_i ? k : INT_MIN);
        fd = imin(fd, in_d ? k : INT_MAX); ld = imax(ld, in_d ? k : INT_MIN);
      }
    }
  }
  if constexpr (!K::kLds) {
    fm = wave_min_any(fm); lm = wave_max_any(lm);
    fi = wave_min_any(fi); li = wave_max_any(li);
    fd = wave_min_any(fd); ld = wave_max_any(ld);
  }
  // offset at k_end: held by lane (k_end - lo) % 64 of the wave owning its chunk
  if (k_end >= lo && k_end <= hi) endu = rdl(endv, (k_end - lo) & (kWave - 1));
  if constexpr (kMax) lmax = wave_max_nonneg(lmax);
  else lmax = 0;
  }
  if constexpr (K::kNW > 1) {   // combine the team's waves
    int nfm = -fm, nfi = -fi, nfd = -fd;
    team_max8(c, nfm, lm, nfi, li, nfd, ld, lmax, endu);
    fm = -nfm; fi = -nfi; fd = -nfd;
  } else {
    wave_sync<K::kLds>();
  }
  return apply_step(c, d, s, slot, lo, has_i, has_d, bm, bi, bd, fm, lm, fi, li, fd, ld, lmax, endu, max_ak);
}

// ---------------------------------------------------------------------------
// extend + termination (wavefront_extend.c:86-211, wavefront_termination.c:37-114)

template <class K>
DI bool terminated(const Ctx<K>& c, const Dir& d, int s, int slot) {
  const int k_end = d.tlen - d.plen;
  int present;
  const Wf w = get_wf(c, d, s, slot, d.ce, &present);
  if (!present || w.lo > k_end || k_end > w.hi) return false;
  return uni(ld_off(c, w.base + (k_end & w.mask))) >= d.tlen;
}

// Returns 1 when the aligner is done; *max_ak gets the max antidiagonal.
template <class K>
DI int extend_step(Ctx<K>& c, Dir& d, int s, int slot, int* max_ak) {
  *max_ak = 0;
  int present;
  const Wf m = get_wf(c, d, s, slot, CM, &present);
  if (!present) {
    if (d.null_steps > c.scope) {
      d.status = kStEndUnreachable;
      d.status_score = s;
      return 1;
    }
    return 0;
  }
  int lmax = 0;
  for (int kb = m.lo + c.wid * kWave; kb <= m.hi; kb += K::kNW * kWave) {
    const int k = kb + c.lane;
    if (k <= m.hi) {
      const int a = m.base + (k & m.mask);
      int off = ld_off(c, a);
      if (off != kOffsetNull) {
        off += lcp(c, d, off - k, off);
        st_off(c, a, off);
        lmax = imax(lmax, 2 * off - k);
      }
    }
  }
  if constexpr (K::kNW == 1) {
    wave_sync<K::kLds>();
    lmax = wave_max_nonneg(lmax);
  } else {
    lmax = team_max1(c, wave_max_nonneg(lmax));
  }
  if (terminated(c, d, s, slot)) {
    d.status = kStEndReached;
    d.status_score = s;
    return 1;
  }
  *max_ak = lmax;
  return 0;
}

// ---------------------------------------------------------------------------
// breakpoint detection (wavefront_bialign.c:189-392)

// First k0 (ascending) of the overlapping interval that satisfies the overlap
// condition; component `comp` (CD/CI: indel-to-indel, CM: M-to-M).
template <class K>
DI void scan_breakpoint(const Ctx<K>& c, bool forward, int s0, int s1, const Wf& w0,
                        const Wf& w1, int comp, int tl, int pl, BP& bp) {
  const int lo0 = w0.lo, hi0 = w0.hi;
  const int lo1 = tl - pl - w1.hi, hi1 = tl - pl - w1.lo;
  if (hi1 < lo0 || hi0 < lo1) return;
  const int min_hi = imin(hi0, hi1), max_lo = imax(lo0, lo1);
  const int sc = (comp == CM) ? s0 + s1 : s0 + s1 - c.o;
  if (!(sc < bp.score)) return;
  for (int kb = max_lo; kb <= min_hi; kb += kWave) {
    const int k0 = kb + c.lane;
    const int k1 = tl - pl - k0;
    bool ok = false;
    int o0 = 0, o1 = 0;
    if (k0 <= min_hi) {
      o0 = ld_off(c, w0.base + (k0 & w0.mask));
      o1 = ld_off(c, w1.base + (k1 & w1.mask));
      ok = o0 + o1 >= tl;
      if (ok && comp != CM) {   // indel: forward coordinates must be in bounds (:226-240)
        ok = forward ? !(o0 - k0 > pl || o0 > tl) : !(o1 - k1 > pl || o1 > tl);
      }
    }
    const uint64_t b = __ballot(ok);
    if (b) {
      const int j = __builtin_ctzll(b);
      const int a0 = rdl(o0, j), a1 = rdl(o1, j);
      const int kk0 = kb + j, kk1 = tl - pl - kk0;
      if (forward) {
        bp.sf = s0; bp.sr = s1; bp.kf = kk0; bp.kr = kk1; bp.of = a0; bp.orr = a1;
      } else {
        bp.sf = s1; bp.sr = s0; bp.kf = kk1; bp.kr = kk0; bp.of = a1; bp.orr = a0;
      }
      bp.score = sc;
      bp.comp = comp;
      return;
    }
  }
}

// Team form of scan_breakpoint: the first hit among the chunks this wave owns
// (INT_MAX if none).  The team's answer is the minimum over its waves.
template <class K>
DI int scan_first(const Ctx<K>& c, bool forward, const Wf& w0, const Wf& w1, int comp, int tl, int pl) {
  const int lo1 = tl - pl - w1.hi, hi1 = tl - pl - w1.lo;
  if (hi1 < w0.lo || w0.hi < lo1) return INT_MAX;
  const int min_hi = imin(w0.hi, hi1), max_lo = imax(w0.lo, lo1);
  for (int kb = max_lo + c.wid * kWave; kb <= min_hi; kb += K::kNW * kWave) {
    const int k0 = kb + c.lane;
    const int k1 = tl - pl - k0;
    bool ok = false;
    if (k0 <= min_hi) {
      const int o0 = ld_off(c, w0.base + (k0 & w0.mask));
      const int o1 = ld_off(c, w1.base + (k1 & w1.mask));
      ok = o0 + o1 >= tl;
      if (ok && comp != CM) ok = forward ? !(o0 - k0 > pl || o0 > tl) : !(o1 - k1 > pl || o1 > tl);
    }
    const uint64_t b = __ballot(ok);
    if (b) return kb + __builtin_ctzll(b);
  }
  return INT_MAX;
}

// Records the team's hit (slot j of `buf`) if it improves the breakpoint.
template <class K>
DI void accept_hit(const Ctx<K>& c, const int32_t* buf, int j, bool forward, int s0, int s1,
                   const Wf& w0, const Wf& w1, int comp, int tl, int pl, BP& bp) {
  const int sc = (comp == CM) ? s0 + s1 : s0 + s1 - c.o;
  if (!(sc < bp.score)) return;
  const int kk0 = team_min_slot(c, buf, j);
  if (kk0 == INT_MAX) return;
  const int kk1 = tl - pl - kk0;
  const int a0 = uni(ld_off(c, w0.base + (kk0 & w0.mask)));
  const int a1 = uni(ld_off(c, w1.base + (kk1 & w1.mask)));
  if (forward) {
    bp.sf = s0; bp.sr = s1; bp.kf = kk0; bp.kr = kk1; bp.of = a0; bp.orr = a1;
  } else {
    bp.sf = s1; bp.sr = s0; bp.kf = kk1; bp.kr = kk0; bp.of = a1; bp.orr = a0;
  }
  bp.score = sc;
  bp.comp = comp;
}

// Lane-parallel form of the overlap test below for a single-wave aligner with
// its rings in LDS.  The sequential order visits i = 0..scope-1 and at each
// score s1-i the D, I and M components, taking a hit only when it strictly
// improves the breakpoint; the breakpoint it ends with is therefore the hit
// with the smallest score, ties going to the earliest visit, and within one
// component the smallest k.  Here lane p = 3i + {0: D, 1: I, 2: M} sets up the
// candidate of one visit (header gather, presence, diagonal intersection), the
// candidates that can improve are compacted into a table, every (candidate, k)
// is tested in fixed-width lane slots, and the key (score, p, k) is minimised
// over the wave.  Requires 3*scope <= 64, wcap <= 64, ring offsets < 4096,
// 2*3*scope <= rcap (the table borrows the backtrace run buffer, idle during a
// breakpoint search) and s0 + s1 < 2^18 (key width); the caller checks.
DI int lane_div3(int p) { return (p * 43) >> 7; }   // p / 3 for 0 <= p < 128

template <class K>
DI void overlap_lanes(Ctx<K>& c, const Dir& a0, int slot0, const Dir& a1, int s0, int s1, int slot1,
                      bool forward, int tl, int pl, BP& bp) {
  const int dk = tl - pl;
  const int p = c.lane;
  const int i = lane_div3(p);
  const int cc = p - 3 * i;                          // 0: D, 1: I, 2: M
  const int si = s1 - i;
  const int sc = s0 + si - (c.o & ~c.om2);          // an indel candidate pays the opening
  int sl = slot1 - i;
  sl += sl < 0 ? c.scope : 0;
  const int comp = 2 - cc;                           // CD, CI, CM
  const int j = a1.ring * c.scope + sl;              // header lane of a1's slot
  const int jb = (j & (kWave - 1)) << 2;
  const int lo_m = __builtin_amdgcn_ds_bpermute(jb, c.rlo[CM]);
  const int lo_i = __builtin_amdgcn_ds_bpermute(jb, c.rlo[CI]);
  const int lo_d = __builtin_amdgcn_ds_bpermute(jb, c.rlo[CD]);
  const int hi_m = __builtin_amdgcn_ds_bpermute(jb, c.rhi[CM]);
  const int hi_i = __builtin_amdgcn_ds_bpermute(jb, c.rhi[CI]);
  const int hi_d = __builtin_amdgcn_ds_bpermute(jb, c.rhi[CD]);
  const int lo1 = (lo_d & c.om0) | (lo_i & c.om1) | (lo_m & c.om2);
  const int hi1 = (hi_d & c.om0) | (hi_i & c.om1) | (hi_m & c.om2);
  // a0's component of this lane's candidate, at score s0 (header lane j0)
  const int j0 = a0.ring * c.scope + slot0;
  const int lo0 = (rdl(c.rlo[CD], j0) & c.om0) | (rdl(c.rlo[CI], j0) & c.om1) | (rdl(c.rlo[CM], j0) & c.om2);
  const int hi0 = (rdl(c.rhi[CD], j0) & c.om0) | (rdl(c.rhi[CI], j0) & c.om1) | (rdl(c.rhi[CM], j0) & c.om2);
  const int b0 = ws_data(c) + (j0 * 3 + comp) * wcap(c);
  const int b1 = ws_data(c) + (j * 3 + comp) * wcap(c);
  const bool p0 = lo0 != kAbsent, p1 = lo1 != kAbsent;   // presence: header lo set
  const int klo = imax(lo0, dk - hi1), khi = imin(hi0, dk - lo1);
  const bool ok = i < c.scope && si >= 0 && (si & c.gmask) == 0 && sc < bp.score && p0 && p1 &&
                  klo <= khi;
  const uint64_t vm = __ballot(ok);
  if (!vm) return;
  const int nr = __builtin_popcountll(vm);
  int32_t* const tab = c.pv + runs_off(c);
  if (ok) {
    const int r = __builtin_amdgcn_mbcnt_hi((uint32_t)(vm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)vm, 0));
    tab[2 * r] = (klo & 0xffff) | (khi << 16);
    tab[2 * r + 1] = b0 | (b1 << 12) | (p << 24);
  }
  wave_sync<K::kLds>();
  const int wm = wcap(c) - 1;
  const int sh = __builtin_ctz(wcap(c));
  const int off = c.lane & wm;
  int best = INT_MAX;
  for (int r0 = 0; r0 < nr; r0 += kWave >> sh) {
    const int r = r0 + (c.lane >> sh);
    if (r < nr) {
      const int e0 = tab[2 * r], e1 = tab[2 * r + 1];
      const int k0 = (int)(int16_t)(e0 & 0xffff) + off;
      if (k0 <= (e0 >> 16)) {
        const int k1 = dk - k0;
        const int pp = (e1 >> 24) & 63;
        const int o0 = ld_off(c, (e1 & 0xfff) + (k0 & wm));
        const int o1 = ld_off(c, ((e1 >> 12) & 0xfff) + (k1 & wm));
        const int ii = lane_div3(pp);
        const bool indel = pp - 3 * ii != 2;
        bool hit = o0 + o1 >= tl;
        if (hit && indel) hit = forward ? !(o0 - k0 > pl || o0 > tl) : !(o1 - k1 > pl || o1 > tl);
        if (hit) {
          const int scc = s0 + s1 - ii - (indel ? c.o : 0);
          best = imin(best, (scc << 12) | (pp << 6) | off);
        }
      }
    }
  }
  const int key = INT_MAX - wave_max_nonneg(INT_MAX - best);
  // the table is read before anything else writes the run buffer
  wave_sync<K::kLds>();
  if (key == INT_MAX) return;
  const int pp = (key >> 6) & 63;
  const int ii = lane_div3(pp);
  const int ccw = pp - 3 * ii;
  const int s1w = s1 - ii;
  const int kk0 = rdl(klo, pp) + (key & 63), kk1 = dk - kk0;
  const int a0v = uni(ld_off(c, rdl(b0, pp) + (kk0 & wm)));
  const int a1v = uni(ld_off(c, rdl(b1, pp) + (kk1 & wm)));
  if (forward) {
    bp.sf = s0; bp.sr = s1w; bp.kf = kk0; bp.kr = kk1; bp.of = a0v; bp.orr = a1v;
  } else {
    bp.sf = s1w; bp.sr = s0; bp.kf = kk1; bp.kr = kk0; bp.of = a1v; bp.orr = a0v;
  }
  bp.score = key >> 12;
  bp.comp = ccw == 0 ? CD : ccw == 1 ? CI : CM;
}

template <class K>
DI void overlap(Ctx<K>& c, const Dir& a0, const Dir& a1, int s0, int slot0, int s1,
                int slot1, bool forward, BP& bp) {
  if (s0 & c.gmask) return;     // a score that is not a multiple of g holds nothing
  const int tl = a0.tlen, pl = a0.plen;
  if constexpr (K::kLds) {
    if (3 * c.scope <= kWave && wcap(c) <= kWave && 2 * 3 * c.scope * wcap(c) <= 4096 &&
        6 * c.scope <= kLdsRcap && s0 + s1 < (1 << 18)) {
      if (rdl(c.rlo[CM], a0.ring * c.scope + slot0) == kAbsent) return;   // no M wavefront at s0
      overlap_lanes(c, a0, slot0, a1, s0, s1, slot1, forward, tl, pl, bp);
      return;
    }
  }
  int pm0, pi0, pd0;
  const Wf m0 = get_wf(c, a0, s0, slot0, CM, &pm0);
  if (!pm0) return;
  const Wf i0 = get_wf(c, a0, s0, slot0, CI, &pi0);
  const Wf d0 = get_wf(c, a0, s0, slot0, CD, &pd0);
  // ring slots of a1 holding any component (null scores hold none)
  const uint64_t live1 = 0xffffffffull;   // every slot is examined; get_wf tests presence
  if constexpr (K::kNW == 1) {
    for (int i = 0; i < c.scope; ++i) {
      const int si = s1 - i;
      if (si < 0) break;
      if (s0 + si - c.o >= bp.score || (si & c.gmask)) continue;
      const int sl = slot_back(slot1, i, c.scope);
      if (!((live1 >> sl) & 1ull)) continue;
      int p1;
      const Wf d1 = get_wf(c, a1, si, sl, CD, &p1);
      if (pd0 && p1) scan_breakpoint(c, forward, s0, si, d0, d1, CD, tl, pl, bp);
      const Wf i1 = get_wf(c, a1, si, sl, CI, &p1);
      if (pi0 && p1) scan_breakpoint(c, forward, s0, si, i0, i1, CI, tl, pl, bp);
      if (s0 + si >= bp.score) continue;
      const Wf m1 = get_wf(c, a1, si, sl, CM, &p1);
      if (p1) scan_breakpoint(c, forward, s0, si, m0, m1, CM, tl, pl, bp);
    }
  } else {
    // Every scan the sequential order could still take (judged by the score the
    // call starts with) runs at once, each wave over its chunks; one barrier
    // later the waves replay the sequential order on the team minima.  A scan's
    // result does not depend on earlier acceptances, only whether it is taken.
    int32_t* buf = team_buf(c);
    int32_t* mine = buf + c.wid * kRedSlots;
    const int entry = bp.score;
    for (int i = 0; i < c.scope; ++i) {
      const int si = s1 - i;
      int hd = INT_MAX, hi = INT_MAX, hm = INT_MAX;
      const int sl0 = slot_back(slot1, i, c.scope);
      if (si >= 0 && s0 + si - c.o < entry && ((live1 >> sl0) & 1ull)) {
        const int sl = sl0;
        int p1;
        const Wf d1 = get_wf(c, a1, si, sl, CD, &p1);
        if (pd0 && p1) hd = scan_first(c, forward, d0, d1, CD, tl, pl);
        const Wf i1 = get_wf(c, a1, si, sl, CI, &p1);
        if (pi0 && p1) hi = scan_first(c, forward, i0, i1, CI, tl, pl);
        if (s0 + si < entry) {
          const Wf m1 = get_wf(c, a1, si, sl, CM, &p1);
          if (p1) hm = scan_first(c, forward, m0, m1, CM, tl, pl);
        }
      }
      if (c.lane == 0) {
        mine[3 * i] = hd;
        mine[3 * i + 1] = hi;
        mine[3 * i + 2] = hm;
      }
    }
    __syncthreads();
    for (int i = 0; i < c.scope; ++i) {
      const int si = s1 - i;
      if (si < 0) break;
      if (s0 + si - c.o >= bp.score) continue;
      const int sl = slot_back(slot1, i, c.scope);
      if (!((live1 >> sl) & 1ull)) continue;
      int p1;
      const Wf d1 = get_wf(c, a1, si, sl, CD, &p1);
      if (pd0 && p1) accept_hit(c, buf, 3 * i, forward, s0, si, d0, d1, CD, tl, pl, bp);
      const Wf i1 = get_wf(c, a1, si, sl, CI, &p1);
      if (pi0 && p1) accept_hit(c, buf, 3 * i + 1, forward, s0, si, i0, i1, CI, tl, pl, bp);
      if (s0 + si >= bp.score) continue;
      const Wf m1 = get_wf(c, a1, si, sl, CM, &p1);
      if (p1) accept_hit(c, buf, 3 * i + 2, forward, s0, si, m0, m1, CM, tl, pl, bp);
    }
    // the winners' offsets are read before any wave moves the rings on
    __syncthreads();
  }
}

// Phase 1 of the breakpoint search on a team: forward step s and reverse step s
// read only their own direction's rings, so the two halves of the team compute
// them at once (each half in tiles, or its first wave for a step of at most 64
// diagonals) and one team reduction serves both; the headers of both are then
// set.  Returns false, having changed nothing, when either step is null or its
// ring slot is nearly full; the caller then steps one direction at a time.
// *fq / *rq: the direction reached its end; *akf / *akr: maximum antidiagonals.
template <class K>
DI bool compute_extend_pair(Ctx<K>& c, Dir& f, Dir& r, int s, int qf, int qr, int& fq, int& rq, int& akf,
                            int& akr) {
  Wf fmm, fmo, fie, fde, rmm, rmo, rie, rde;
  int flo = 0, fhi = 0, rlo = 0, rhi = 0;
  bool fhas_i = false, fhas_d = false, rhas_i = false, rhas_d = false;
  if (!gather_inputs(c, f, s, qf, fmm, fmo, fie, fde, flo, fhi, fhas_i, fhas_d)) return false;
  if (!gather_inputs(c, r, s, qr, rmm, rmo, rie, rde, rlo, rhi, rhas_i, rhas_d)) return false;
  const int w = wcap(c);
  if ((fhi | 3) - (flo & ~3) + 1 > w || (rhi | 3) - (rlo & ~3) + 1 > w) return false;
  f.null_steps = 0;
  r.null_steps = 0;
  const int mask = w - 1;
  const int fbm = ring_base(c, f.ring, qf, CM), fbi = ring_base(c, f.ring, qf, CI), fbd = ring_base(c, f.ring, qf, CD);
  const int rbm = ring_base(c, r.ring, qr, CM), rbi = ring_base(c, r.ring, qr, CI), rbd = ring_base(c, r.ring, qr, CD);
  constexpr int kHalf = K::kNW / 2;
  const bool rside = c.wid >= kHalf;
  const int hw = rside ? c.wid - kHalf : c.wid;
  int fm = INT_MAX, lm = INT_MIN, fi = INT_MAX, li = INT_MIN, fd = INT_MAX, ld = INT_MIN, lmax = 0, endu = INT_MIN;
  if (!rside) {
    if (fhi - flo + 1 <= kWave) {
      if (hw == 0)
        step_chunk<K, true>(c, f, flo, fhi, fmm, fmo, fie, fde, fhas_i, fhas_d, fbm, fbi, fbd, mask, fm, lm, fi, li,
                            fd, ld, lmax, endu);
    } else {
      step_tiles_reduced<K, true>(c, f, flo, fhi, fmm, fmo, fie, fde, fhas_i, fhas_d, fbm, fbi, fbd, mask, hw, kHalf,
                                  fm, lm, fi, li, fd, ld, lmax, endu);
    }
  } else {
    if (rhi - rlo + 1 <= kWave) {
      if (hw == 0)
        step_chunk<K, true>(c, r, rlo, rhi, rmm, rmo, rie, rde, rhas_i, rhas_d, rbm, rbi, rbd, mask, fm, lm, fi, li,
                            fd, ld, lmax, endu);
    } else {
      step_tiles_reduced<K, true>(c, r, rlo, rhi, rmm, rmo, rie, rde, rhas_i, rhas_d, rbm, rbi, rbd, mask, hw, kHalf,
                                  fm, lm, fi, li, fd, ld, lmax, endu);
    }
  }
  // slots 0..7 the forward step's values, 8..15 the reverse step's (INT_MIN is
  // neutral for every one of them under max)
  const int own[8] = {-fm, lm, -fi, li, -fd, ld, lmax, endu};
  int v[16];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    v[j] = rside ? INT_MIN : own[j];
    v[8 + j] = rside ? own[j] : INT_MIN;
  }
  team_max16(c, v);
  akf = 0;
  akr = 0;
  fq = apply_step(c, f, s, qf, flo, fhas_i, fhas_d, fbm, fbi, fbd, -v[0], v[1], -v[2], v[3], -v[4], v[5], v[6], v[7],
                  &akf);
  rq = apply_step(c, r, s, qr, rlo, rhas_i, rhas_d, rbm, rbi, rbd, -v[8], v[9], -v[10], v[11], -v[12], v[13], v[14],
                  v[15], &akr);
  return true;
}

// wavefront_aligner_init_wf (wavefront_aligner.c:311-383) for a ring aligner.
template <class K>
DI void ring_init(Ctx<K>& c, Dir& d) {
  // a team's slower waves may still read the previous sub-problem's workspace
  if constexpr (K::kNW > 1) __syncthreads();
  {   // the ring's header lanes read absent
    const bool mine = c.lane >= d.ring * c.scope && c.lane < (d.ring + 1) * c.scope;
    for (int q = 0; q < 3; ++q) c.rlo[q] = mine ? kAbsent : c.rlo[q];
  }
  set_hdr(c, d, 0, 0, d.cb, 0, 0, 0);
  st_off(c, ring_base(c, d.ring, 0, d.cb), 0);
  d.null_steps = 0;
  d.status = kStOk;
  d.status_score = 0;
  team_sync<K>();
}

DI int next_slot(int slot, int scope) { return slot + 1 == scope ? 0 : slot + 1; }

// wavefront_bialign_find_breakpoint (wavefront_bialign.c:411-519)
template <class K>
DI int find_breakpoint(Ctx<K>& c, Dir& f, Dir& r, BP& bp) {
  ring_init(c, f);
  ring_init(c, r);
  const int max_antidiagonal = f.plen + f.tlen - 1;
  int sf = 0, sr = 0, qf = 0, qr = 0, f_max = 0, r_max = 0, ak = 0;
  bp.score = INT_MAX;
  c.guard = score_guard(c, f.plen, f.tlen);
  if (extend_step(c, f, sf, qf, &f_max)) return f.status;
  if (extend_step(c, r, sr, qr, &r_max)) return r.status;
  bool last_forward = false;
  for (;;) {
    if (f_max + r_max >= max_antidiagonal) break;
    if (sf + sr > c.guard) return kStGuard;
    if constexpr (!K::kLds && K::kNW > 1) {
      // Both next steps at once when they are not null by parity.  The reverse
      // step then runs before WFA2's check between the two; when that check
      // ends phase 1, the step is undone: it wrote the slot of score s - scope,
      // a score that is not a multiple of g (scope = max(o+e, x) + 1 is 1 mod g),
      // hence null and absent, and its headers and the direction's state are
      // restored (its ring data is never read while absent).
      const int s2 = sf + 1;
      if (sf == sr && c.gmask != 0 && (s2 & c.gmask) == 0) {
        const int qf2 = next_slot(qf, c.scope), qr2 = next_slot(qr, c.scope);
        const int j = r.ring * c.scope + qr2;
        const int h0 = rdl(c.rlo[CM], j), h1 = rdl(c.rlo[CI], j), h2 = rdl(c.rlo[CD], j);
        const int u0 = rdl(c.rhi[CM], j), u1 = rdl(c.rhi[CI], j), u2 = rdl(c.rhi[CD], j);
        const Dir rsave = r;
        int fq = 0, rq = 0, akf = 0, akr = 0;
        if (compute_extend_pair(c, f, r, s2, qf2, qr2, fq, rq, akf, akr)) {
          sf = s2; qf = qf2;
          if (f_max < akf) f_max = akf;
          last_forward = true;
          if (fq) return f.status;
          if (f_max + r_max >= max_antidiagonal) {   // WFA2 stops here, before the reverse step
            r = rsave;
            lane_set(c.rlo, CM, c.lane, j, h0); lane_set(c.rlo, CI, c.lane, j, h1); lane_set(c.rlo, CD, c.lane, j, h2);
            lane_set(c.rhi, CM, c.lane, j, u0); lane_set(c.rhi, CI, c.lane, j, u1); lane_set(c.rhi, CD, c.lane, j, u2);
            break;
          }
          sr = s2; qr = qr2;
          if (r_max < akr) r_max = akr;
          last_forward = false;
          if (rq) return r.status;
          if (sr + sf >= c.max_steps) return kStMaxSteps;
          continue;
        }
      }
    }
    ++sf; qf = next_slot(qf, c.scope);
    const int fq = compute_extend(c, f, sf, qf, &ak);
    if (c.spill) return kStCapacity;
    if (f_max < ak) f_max = ak;
    last_forward = true;
    if (fq) return f.status;
    if (f_max + r_max >= max_antidiagonal) break;
    ++sr; qr = next_slot(qr, c.scope);
    const int rq = compute_extend(c, r, sr, qr, &ak);
    if (c.spill) return kStCapacity;
    if (r_max < ak) r_max = ak;
    last_forward = false;
    if (rq) return r.status;
    if (sr + sf >= c.max_steps) return kStMaxSteps;
  }
  for (;;) {
    if (last_forward) {
      const int min_r = (sr > c.scope - 1) ? sr - (c.scope - 1) : 0;
      if (sf + min_r - c.o >= bp.score) break;
      overlap(c, f, r, sf, qf, sr, qr, true, bp);
      ++sr; qr = next_slot(qr, c.scope);
      const int rq = compute_extend<K, false>(c, r, sr, qr, &ak);
      if (c.spill) return kStCapacity;
      if (rq) return r.status;
    }
    const int min_f = (sf > c.scope - 1) ? sf - (c.scope - 1) : 0;
    if (min_f + sr - c.o >= bp.score) break;
    if (sf + sr > c.guard) return kStGuard;
    overlap(c, r, f, sr, qr, sf, qf, false, bp);
    ++sf; qf = next_slot(qf, c.scope);
    const int fq = compute_extend<K, false>(c, f, sf, qf, &ak);
    if (c.spill) return kStCapacity;
    if (fq) return f.status;
    if (sr + sf >= c.max_steps) return kStMaxSteps;
    last_forward = true;
  }
  return kStOk;
}

// ---------------------------------------------------------------------------
// output runs

// User penalties for the .alg score: scalar registers in the global kernels,
// lanes 0..3 of one vector register in the LDS kernel (which runs close to its
// scalar register limit; they are read once per CIGAR run).
template <class K> DI int pen_m(const Ctx<K>& c) { return K::kLds ? rdl(c.upen, 0) : c.um; }
template <class K> DI int pen_x(const Ctx<K>& c) { return K::kLds ? rdl(c.upen, 1) : c.ux; }
template <class K> DI int pen_o(const Ctx<K>& c) { return K::kLds ? rdl(c.upen, 2) : c.uo; }
template <class K> DI int pen_e(const Ctx<K>& c) { return K::kLds ? rdl(c.upen, 3) : c.ue; }

template <class K>
DI void out_flush(Ctx<K>& c) {
  if (c.pend_len <= 0) return;
  if (c.out_n >= c.out_cap) {
    c.cap_err = 1;
  } else {
    if (c.leader) c.out[c.out_n] = ((uint32_t)c.pend_len << 4) | (uint32_t)c.pend_op;
    c.out_n++;
  }
  // cigar_score_gap_affine (cigar.c:277-303) over maximal runs
  if (c.pend_op == kOpM) c.alg_score -= pen_m(c) * c.pend_len;
  else if (c.pend_op == kOpX) c.alg_score -= pen_x(c) * c.pend_len;
  else c.alg_score -= pen_o(c) + pen_e(c) * c.pend_len;
  c.pend_len = 0;
  c.pend_op = -1;
}
template <class K>
DI void out_append(Ctx<K>& c, int op, int n) {
  if (n <= 0) return;
  if (op == c.pend_op) { c.pend_len += n; return; }
  out_flush(c);
  c.pend_op = op;
  c.pend_len = n;
}

// reversed runs of one backtrace, merged as they are emitted
struct RevRuns {
  int n, op, len;
};
template <class K>
DI void rev_emit(Ctx<K>& c, RevRuns& rr, int op, int n) {
  if (n <= 0) return;
  if (op == rr.op) { rr.len += n; return; }
  if (rr.len > 0) {
    if (rr.n >= (K::kLds ? kLdsRcap : c.L.rcap)) { c.spill = 1; return; }
    if (c.lane == 0) c.pv[runs_off(c) + rr.n] = (int)(((uint32_t)rr.len << 4) | (uint32_t)rr.op);
    rr.n++;
  }
  rr.op = op;
  rr.len = n;
}

// ---------------------------------------------------------------------------
// base case: wavefront_bialign_base (wavefront_bialign.c:155-185)

constexpr int kBtM = 9, kBtD1Ext = 6, kBtD1Open = 5, kBtI1Ext = 2, kBtI1Open = 1;

template <class K>
DI long long bt_src(const Ctx<K>& c, const Dir& b, int score, int comp, int k, int add, int type) {
  int present;
  const Wf w = get_wf(c, b, score, 0, comp, &present);
  if (!present || !(w.lo <= k && k <= w.hi)) return (long long)kOffsetNull;
  return (long long)(uni(ld_off(c, w.base + k)) + add) * 16 | type;
}
DI long long max64(long long a, long long b) { return a >= b ? a : b; }

// The packed maximum of bt_src over the sources of (matrix, score, k) when the
// base-case headers live in memory (global workspace): the five header reads
// are issued together, then the five offset reads, with no branch in between,
// so a backtrace step costs two memory round trips instead of up to ten.
template <class K>
DI long long bt_max_mem(const Ctx<K>& c, int score, int matrix, int k) {
  const int sj[5] = {score - c.x, score - c.o - c.e, score - c.e, score - c.o - c.e, score - c.e};
  const int cj[5] = {CM, CM, CI, CM, CD};
  const int dk[5] = {0, -1, -1, 1, 1};
  const int ad[5] = {1, 1, 1, 0, 0};
  const int ty[5] = {kBtM, kBtI1Open, kBtI1Ext, kBtD1Open, kBtD1Ext};
  const bool en[5] = {matrix == CM, matrix != CD, matrix != CD, matrix != CI, matrix != CI};
  int lo[5], hi[5], kb[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const int h = c.L.bhdr + ((sj[j] >= 0 ? sj[j] : 0) * 3 + cj[j]) * 3;
    lo[j] = c.pv[h];
    hi[j] = c.pv[h + 1];
    kb[j] = c.pv[h + 2];
  }
  bool in[5];
  int off[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const int l = uni(lo[j]), u = uni(hi[j]), kk = k + dk[j];
    in[j] = en[j] && sj[j] >= 0 && l != kAbsent && l <= kk && kk <= u;
    off[j] = ld_off(c, in[j] ? uni(kb[j]) + kk : ws_data(c));
  }
  long long m = LLONG_MIN;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    if (!en[j]) continue;
    const long long v = in[j] ? ((long long)(uni(off[j]) + ad[j]) * 16 | ty[j]) : (long long)kOffsetNull;
    m = max64(m, v);
  }
  return m;
}

// wavefront_backtrace_affine (wavefront_backtrace.c:320-529); appends the
// operations to the output in forward order.  Returns 0 or -1 on a trace error.
template <class K>
DI int backtrace(Ctx<K>& c, const Dir& b, int score) {
  int matrix = b.ce;
  int k = b.tlen - b.plen, off = b.tlen;
  int h = off, v = off - k;
  RevRuns rr{0, -1, 0};
  if (matrix == CM) {
    if (v < b.plen) rev_emit(c, rr, kOpD, b.plen - v);
    if (h < b.tlen) rev_emit(c, rr, kOpI, b.tlen - h);
  }
  while (v > 0 && h > 0 && score > 0) {
    const int mismatch = score - c.x;
    const int gap_open1 = score - c.o - c.e;
    const int gap_extend1 = score - c.e;
    long long max_all;
    if constexpr (!K::kLds) {
      max_all = bt_max_mem(c, score, matrix, k);
    } else if (matrix == CM) {
      const long long misms = bt_src(c, b, mismatch, CM, k, 1, kBtM);
      const long long io = bt_src(c, b, gap_open1, CM, k - 1, 1, kBtI1Open);
      const long long ix = bt_src(c, b, gap_extend1, CI, k - 1, 1, kBtI1Ext);
      const long long dop = bt_src(c, b, gap_open1, CM, k + 1, 0, kBtD1Open);
      const long long dx = bt_src(c, b, gap_extend1, CD, k + 1, 0, kBtD1Ext);
      max_all = max64(misms, max64(max64(io, ix), max64(dop, dx)));
    } else if (matrix == CI) {
      max_all = max64(bt_src(c, b, gap_open1, CM, k - 1, 1, kBtI1Open),
                      bt_src(c, b, gap_extend1, CI, k - 1, 1, kBtI1Ext));
    } else {
      max_all = max64(bt_src(c, b, gap_open1, CM, k + 1, 0, kBtD1Open),
                      bt_src(c, b, gap_extend1, CD, k + 1, 0, kBtD1Ext));
    }
    if (max_all < 0) break;
    if (matrix == CM) {
      const int max_off = (int)(max_all >> 4);
      rev_emit(c, rr, kOpM, off - max_off);
      off = max_off;
      v = off - k;
      h = off;
      if (v <= 0 || h <= 0) break;
    }
    switch ((int)(max_all & 15)) {
      case kBtM: score = mismatch; matrix = CM; rev_emit(c, rr, kOpX, 1); --off; break;
      case kBtI1Open: score = gap_open1; matrix = CM; rev_emit(c, rr, kOpI, 1); --k; --off; break;
      case kBtI1Ext: score = gap_extend1; matrix = CI; rev_emit(c, rr, kOpI, 1); --k; --off; break;
      case kBtD1Open: score = gap_open1; matrix = CM; rev_emit(c, rr, kOpD, 1); ++k; break;
      case kBtD1Ext: score = gap_extend1; matrix = CD; rev_emit(c, rr, kOpD, 1); ++k; break;
      default: return -1;
    }
    if (c.spill) return 0;
    v = off - k;
    h = off;
  }
  if (matrix == CM) {
    if (v > 0 && h > 0) {
      const int nm = imin(v, h);
      rev_emit(c, rr, kOpM, nm);
      v -= nm;
      h -= nm;
    }
    rev_emit(c, rr, kOpD, v);
    rev_emit(c, rr, kOpI, h);
  } else if (v != 0 || h != 0 || (score != 0 && pen_m(c) == 0)) {
    return -1;
  }
  if (c.spill) return 0;
  wave_sync<K::kLds>();
  // forward order: the pending run first, then the stored runs backwards
  out_append(c, rr.op, rr.len);
  if constexpr (K::kLds) {
    for (int i = rr.n - 1; i >= 0; --i) {
      const uint32_t run = (uint32_t)uni(c.pv[runs_off(c) + i]);
      out_append(c, (int)(run & 15u), (int)(run >> 4));
    }
  } else {   // 64 runs per memory round trip
    for (int i0 = rr.n - 1; i0 >= 0; i0 -= kWave) {
      const int i = i0 - c.lane;
      const int v = i >= 0 ? c.pv[runs_off(c) + i] : 0;
      const int cnt = imin(kWave, i0 + 1);
      for (int j = 0; j < cnt; ++j) {
        const uint32_t run = (uint32_t)rdl(v, j);
        out_append(c, (int)(run & 15u), (int)(run >> 4));
      }
    }
  }
  return 0;
}

// Unidirectional WFA over the sub-problem with all wavefronts kept
// (wavefront_unialign.c:242-273).  Returns kStOk or kStUnattainable.
template <class K>
DI int base_case(Ctx<K>& c, int pb, int pe, int tb, int te, int cb, int ce, int* score_out) {
  Dir b{pb, pe - pb, tb, te - tb, 0, -1, cb, ce, 0, kStOk, 0};
  // a team's slower waves may still read the previous sub-problem's workspace
  if constexpr (K::kNW > 1) __syncthreads();
  // score 0: the begin component at k=0 (wavefront_aligner.c:311-383)
  c.arena_top = 1;
  set_hdr(c, b, 0, 0, CM, cb == CM ? 0 : kAbsent, 0, ws_data(c));
  set_hdr(c, b, 0, 0, CI, cb == CI ? 0 : kAbsent, 0, ws_data(c));
  set_hdr(c, b, 0, 0, CD, cb == CD ? 0 : kAbsent, 0, ws_data(c));
  st_off(c, ws_data(c), 0);
  team_sync<K>();
  int score = 0, ak;
  const int guard = score_guard(c, b.plen, b.tlen);
  int done = extend_step(c, b, 0, 0, &ak);
  while (!done) {
    ++score;
    done = compute_extend<K, false>(c, b, score, 0, &ak);
    if (c.spill) return kStCapacity;
    if (score >= c.max_steps) return kStUnattainable;   // reported as unattainable by the base
    if (score > guard) return kStGuard;
  }
  if (b.status != kStEndReached) return kStUnattainable;
  if (backtrace(c, b, score) != 0) return kStUnattainable;
  *score_out = score;
  return kStOk;
}

// ---------------------------------------------------------------------------
// recursion driver: wavefront_bialign_alignment (wavefront_bialign.c:581-658)

struct Sub {
  int pair, pb, pe, tb, te, cb, ce, rem, lvl;
};
template <class K>
DI void push(Ctx<K>& c, int& sp, const Sub& s) {
  if (sp >= (K::kLds ? kLdsScap : c.L.scap)) { c.spill = 1; return; }
  int* e = c.pv + stack_off(c) + sp * kSubInts;
  e[0] = s.pair; e[1] = s.pb; e[2] = s.pe; e[3] = s.tb; e[4] = s.te;
  e[5] = s.cb; e[6] = s.ce; e[7] = s.rem; e[8] = s.lvl;
  ++sp;
}
template <class K>
DI Sub pop(Ctx<K>& c, int& sp) {
  --sp;
  const int* e = c.pv + stack_off(c) + sp * kSubInts;
  return Sub{uni(e[0]), uni(e[1]), uni(e[2]), uni(e[3]), uni(e[4]), uni(e[5]), uni(e[6]), uni(e[7]),
             uni(e[8])};
}

constexpr int kSubLeaf = 1;    // output emitted
constexpr int kSubSplit = 2;   // breakpoint found: two children

// One sub-problem of wavefront_bialign_alignment (wavefront_bialign.c:581-658):
// trivial cases, the base case below the score threshold, or the breakpoint
// search with its exception path (:520-547).  Returns kSubLeaf, kSubSplit (the
// children in *left / *right) or a failure status.  *wf_score receives the
// level-0 score.
template <class K>
DI int run_sub(Ctx<K>& c, const Sub& s, Sub* left, Sub* right, int* wf_score) {
  const int pl = s.pe - s.pb, tl = s.te - s.tb;
  if (tl == 0) { out_append(c, kOpD, pl); return kSubLeaf; }
  if (pl == 0) { out_append(c, kOpI, tl); return kSubLeaf; }
  int bscore = 0, st;
  if (s.rem <= kFallbackMinScore) {
    st = base_case(c, s.pb, s.pe, s.tb, s.te, s.cb, s.ce, &bscore);
    if (c.spill) return kStCapacity;
    if (s.lvl == 0) *wf_score = bscore;
    return st == kStOk ? kSubLeaf : st;
  }
  Dir f{s.pb, pl, s.tb, tl, 0, 0, s.cb, s.ce, 0, kStOk, 0};
  Dir r{s.pb, pl, s.tb, tl, 1, 1, s.ce, s.cb, 0, kStOk, 0};
  BP bp;
  st = find_breakpoint(c, f, r, bp);
  if (c.spill) return kStCapacity;
  if (st != kStOk) {   // wavefront_bialign_find_breakpoint_exception (:520-547)
    if (st != kStEndReached) return st;
    const int reached = (f.status == kStEndReached) ? f.status_score : r.status_score;
    if (reached > kRecoveryMinScore) return kStEndUnreachable;
    st = base_case(c, s.pb, s.pe, s.tb, s.te, s.cb, s.ce, &bscore);
    if (c.spill) return kStCapacity;
    if (s.lvl == 0) *wf_score = bscore;
    return st == kStOk ? kSubLeaf : st;
  }
  if (s.lvl == 0) *wf_score = bp.score;
  const int bh = bp.of, bv = bp.of - bp.kf;
  *right = Sub{s.pair, s.pb + bv, s.pe, s.tb + bh, s.te, bp.comp, s.ce, bp.sr, s.lvl + 1};
  *left = Sub{s.pair, s.pb, s.pb + bv, s.tb, s.tb + bh, s.cb, bp.comp, bp.sf, s.lvl + 1};
  return kSubSplit;
}

// Final status of a failed alignment (wavefront_bialign.c:722-729).
DI int final_status(int st) {
  if (st == kStMaxSteps || st == kStGuard || st == kStCapacity) return st;
  return kStUnattainable;
}

// wavefront_bialign_compute_score (wavefront_bialign.c:662-699)
template <class K>
DI int score_pair(Ctx<K>& c, int plen, int tlen, int* wf_score) {
  *wf_score = INT_MIN;
  Dir f{0, plen, 0, tlen, 0, 0, CM, CM, 0, kStOk, 0};
  Dir r{0, plen, 0, tlen, 1, 1, CM, CM, 0, kStOk, 0};
  BP bp;
  const int st = find_breakpoint(c, f, r, bp);
  if (c.spill) return kStCapacity;
  if (st != kStOk && st != kStEndReached)
    return (st == kStMaxSteps || st == kStGuard) ? st : kStUnattainable;
  int s = bp.score;
  if (st == kStEndReached) s = (f.status == kStEndReached) ? f.status_score : r.status_score;
  *wf_score = s;
  const int um = pen_m(c);
  c.alg_score = (um == 0) ? -s : ((-um) * (plen + tlen) - s) / 2;
  return 0;
}

// Aligns the staged pair depth first on this aligner alone, emitting the runs
// in order.  Returns the final WFA status (wavefront_bialign.c:722-729).
template <class K>
DI int align_pair(Ctx<K>& c, int pair, int plen, int tlen, int score_only, int* wf_score) {
  if (score_only) return score_pair(c, plen, tlen, wf_score);
  *wf_score = INT_MIN;
  int sp = 0;
  const bool min_length = imax(plen, tlen) <= kFallbackMinLength;
  push(c, sp, Sub{pair, 0, plen, 0, tlen, CM, CM, min_length ? 0 : INT_MAX, 0});
  wave_sync<K::kLds>();
  int status = kStOk;
  int iters = 0;
  const int max_iters = 4 * (plen + tlen) + 64;
  while (sp > 0) {
    if (++iters > max_iters) { status = kStGuard; break; }
    const Sub s = pop(c, sp);
    Sub l, r;
    const int res = run_sub(c, s, &l, &r, wf_score);
    if (c.spill) return kStCapacity;
    if (res == kSubLeaf) continue;
    if (res != kSubSplit) { status = res; break; }
    push(c, sp, r);
    push(c, sp, l);
    if (c.spill) return kStCapacity;
    wave_sync<K::kLds>();
  }
  out_flush(c);
  if (status == kStOk) return 0;
  return final_status(status);
}

// ---------------------------------------------------------------------------
// shared recursion tasks (global tiers, score + CIGAR)
//
// A pair is taken from the tier's list by one aligner, which runs its root
// sub-problem.  At every breakpoint the right child is published to the tier's
// item queue (when it has room; else kept) and the left child stays on the
// aligner's own stack, so idle aligners take over halves of long pairs.
//   * Output: a leaf covering [pb,pe) x [tb,te) has at most (pe-pb)+(te-tb)
//     runs and writes them at pb+tb of the pair's run buffer (zeroed when the
//     pair starts); leaves tile the alignment path in order, so their slots
//     never overlap and the zero gaps can be squeezed out afterwards.
//   * pend[pair] counts unfinished sub-problems (+2 per split before the
//     children exist, -1 per finished one); whoever takes it to zero merges the
//     runs (maximal runs, as the sequential emitter produces them) and writes
//     the pair's results.
//   * Failures: the leftmost failing sub-problem is the one depth-first order
//     meets first (its left neighbours are finished leaves, its ancestors
//     succeeded), so its status is the pair's status.
//   * Publishing: tail claimed with an atomic, fields written, agent-scope
//     fence, then the ready flag set to the launch's epoch; a taker spins on
//     the flag and fences before reading.

constexpr int kFlagSpill = 1, kFlagRuns = 2;
// A claimed item's producer writes it within microseconds of claiming its slot;
// a taker that waits longer than KParams::ready_ticks (100 MHz clock, 60 s by
// default) for the ready flag gives up and raises the tier's error flag (a
// defect, never seen in a correct run).  Every aligner of the tier then leaves
// its work loop at its next poll, so the launch ends and the host reports the
// flag.  Pairs shorter than KParams::task_min_length (p + t) are aligned by one
// aligner from start to end with the runs emitted in order: sharing their
// halves costs more (zeroing, merging, agent-scope fences) than it saves.

DI int atomic_ld(const int32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// the leader's values a, b to the whole team
template <class K>
DI void team_bcast2(Ctx<K>& c, int& a, int& b) {
  if constexpr (K::kNW == 1) {
    a = uni(__shfl(a, 0, kWave));
    b = uni(__shfl(b, 0, kWave));
  } else {
    int32_t* buf = team_buf(c);
    if (c.leader) { buf[0] = a; buf[1] = b; }
    __syncthreads();
    a = uni(buf[0]);
    b = uni(buf[1]);
  }
}

// Stages the pair's sequences in the aligner's LDS area when they fit (same
// padded layout as the LDS tier), else points at HBM.
template <class K>
DI void stage_pair(Ctx<K>& c, const KParams& p, int pair, int32_t* lds, int wv) {
  const int plen = uni(p.p_len[pair]), tlen = uni(p.t_len[pair]);
  const uint64_t* gP = p.words + p.p_woff[pair];
  const uint64_t* gT = p.words + p.t_woff[pair];
  c.P = gP;
  c.T = gT;
  c.seq_lds = 0;
  const int pw = (plen + 31) >> 5, tw = (tlen + 31) >> 5;
  const int sw = p.seq_lds_words;
  if (pw + 3 <= sw && tw + 3 <= sw) {
    uint64_t* sP = reinterpret_cast<uint64_t*>(lds) + (K::kNW > 1 ? 0 : 2 * sw * wv);
    uint64_t* sT = sP + sw;
    const int t0 = K::kNW > 1 ? (int)threadIdx.x : c.lane;
    team_sync<K>();   // the previous pair's readers are done
    for (int i = t0; i < pw + 3; i += K::kNW * kWave) sP[i] = (i >= 1 && i <= pw) ? gP[i - 1] : 0ull;
    for (int i = t0; i < tw + 3; i += K::kNW * kWave) sT[i] = (i >= 1 && i <= tw) ? gT[i - 1] : 0ull;
    team_sync<K>();
    c.P = sP + 1;
    c.T = sT + 1;
    c.seq_lds = 1;
  }
}

// Publishes sub-problem s to the tier's queue; returns 0 when the queue is full.
template <class K>
DI int publish(Ctx<K>& c, const KParams& p, const Sub& s) {
  int ok = 0, unused = 0;
  if (c.leader) {
    const int slot = atomicAdd(&p.qctl[3], 1);
    if (slot < p.item_cap) {
      int32_t* it = p.items + (int64_t)slot * kItemInts;
      it[1] = s.pair; it[2] = s.pb; it[3] = s.pe; it[4] = s.tb; it[5] = s.te;
      it[6] = s.cb; it[7] = s.ce; it[8] = s.rem; it[9] = s.lvl;
      __threadfence();
      if (!p.withhold_ready) __hip_atomic_store(it, c.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      ok = 1;
    }
  }
  team_bcast2(c, ok, unused);
  return ok;
}

// Merges the pair's leaf runs in place (zero slots dropped, equal neighbours
// joined) and writes its results.  Wave 0 alone reads and rewrites the buffer.
template <class K>
DI void finish_pair(Ctx<K>& c, const KParams& p, int pair) {
  const int plen = uni(p.p_len[pair]), tlen = uni(p.t_len[pair]);
  const int flags = uni(atomic_ld(&p.pflag[pair]));
  const unsigned long long fail = __hip_atomic_load(&p.pfail[pair], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int fail_lo = uni((int)(uint32_t)(fail & 0xffffffffull)), fail_hi = uni((int)(uint32_t)(fail >> 32));
  const bool failed = !(fail_lo == -1 && fail_hi == -1);
  if (flags & kFlagSpill) {
    if (p.next_list) {      // re-run by the next tier with larger capacities
      if (c.leader) {
        const int slot = atomicAdd(p.next_count, 1);
        p.next_list[slot] = pair;
        atomicSub(&p.qctl[4], 1);   // no longer in flight here
      }
      return;
    }
  }
  int st = 0;
  if (flags & kFlagSpill) st = kStCapacity;
  else if (failed) st = final_status(-(fail_lo & 0xffff));
  else if (flags & kFlagRuns) st = kStCapacity;
  uint32_t* base = p.runs + p.runs_off[pair];
  c.out = base;
  c.out_cap = plen + tlen + 1;
  c.out_n = 0; c.pend_op = -1; c.pend_len = 0; c.alg_score = 0; c.cap_err = 0;
  if (st == 0 && c.wid == 0) {
    const int total = plen + tlen;
    for (int i0 = 0; i0 < total; i0 += kWave) {
      const int i = i0 + c.lane;
      const uint32_t v = i < total ? base[i] : 0u;
      uint64_t b = __ballot(v != 0u);
      while (b) {
        const int j = __builtin_ctzll(b);
        b &= b - 1;
        const uint32_t run = (uint32_t)rdl((int)v, j);
        out_append(c, (int)(run & 15u), (int)(run >> 4));
      }
    }
    out_flush(c);
  }
  if (c.leader) {
    p.status[pair] = st;
    p.score[pair] = (st == 0 && c.out_n > 0) ? c.alg_score : -1;
    p.nruns[pair] = st == 0 ? c.out_n : 0;
    atomicSub(&p.qctl[4], 1);   // no longer in flight
  }
  team_sync<K>();
}

// One finished sub-problem of `pair`; the last one finishes the pair.
template <class K>
DI void sub_done(Ctx<K>& c, const KParams& p, int pair) {
  int old = 0, unused = 0;
  if (c.leader) {
    __threadfence();
    old = atomicSub(&p.pend[pair], 1);
  }
  team_bcast2(c, old, unused);
  if (old == 1) {
    __threadfence();
    finish_pair(c, p, pair);
  }
}

// Runs `first` and every sub-problem it leaves on this aligner's stack.
template <class K>
DI void run_task(Ctx<K>& c, const KParams& p, const Sub& first, int& staged, int32_t* lds, int wv) {
  int sp = 0;
  push(c, sp, first);
  wave_sync<K::kLds>();
  int iters = 0;
  const int max_iters = 8 * ((first.pe - first.pb) + (first.te - first.tb)) + 64;
  while (sp > 0) {
    const Sub s = pop(c, sp);
    if (s.pair != staged) {
      stage_pair(c, p, s.pair, lds, wv);
      staged = s.pair;
    }
    c.out = p.runs + p.runs_off[s.pair] + s.pb + s.tb;
    c.out_cap = (s.pe - s.pb) + (s.te - s.tb);
    c.out_n = 0; c.pend_op = -1; c.pend_len = 0; c.alg_score = 0;
    c.spill = 0; c.cap_err = 0; c.arena_top = 0;
    int wf = INT_MIN;
    Sub l, r;
    int res = (++iters > max_iters) ? kStGuard : run_sub(c, s, &l, &r, &wf);
    if (c.spill) res = kStCapacity;
    if (res == kSubLeaf) {
      out_flush(c);
      if (c.cap_err && c.leader) atomicOr(&p.pflag[s.pair], kFlagRuns);
    } else if (res == kSubSplit) {
      if (c.leader) atomicAdd(&p.pend[s.pair], 2);
      int lost = 0;   // children the full stack could not take (finished as failed)
      if (!publish(c, p, r)) {
        push(c, sp, r);
        if (c.spill) { ++lost; c.spill = 0; }
      }
      push(c, sp, l);
      if (c.spill) { ++lost; c.spill = 0; }
      wave_sync<K::kLds>();
      if (lost) {
        if (c.leader) atomicOr(&p.pflag[s.pair], kFlagSpill);
        for (int q = 0; q < lost; ++q) sub_done(c, p, s.pair);
      }
    } else if (c.leader) {
      if (res == kStCapacity) {
        atomicOr(&p.pflag[s.pair], kFlagSpill);
      } else {
        const unsigned long long key = ((unsigned long long)(uint32_t)(s.pb + s.tb) << 16) |
                                       (unsigned long long)((uint32_t)(-res) & 0xffffu);
        atomicMin(&p.pfail[s.pair], key);
      }
    }
    if (s.lvl == 0 && wf != INT_MIN && c.leader) p.wf_score[s.pair] = wf;
    sub_done(c, p, s.pair);
  }
}

// Sets a pair up for shared-task alignment and runs its root (score-only: the
// whole alignment, wavefront_bialign_compute_score).
template <class K>
DI void start_pair(Ctx<K>& c, const KParams& p, int pair, int& staged, int32_t* lds, int wv) {
  const int plen = uni(p.p_len[pair]), tlen = uni(p.t_len[pair]);
  if (p.score_only) {
    if (pair != staged) {
      stage_pair(c, p, pair, lds, wv);
      staged = pair;
    }
    c.spill = 0; c.arena_top = 0;
    int wf = INT_MIN;
    const int st = score_pair(c, plen, tlen, &wf);
    if (c.spill && p.next_list) {
      if (c.leader) {
        const int slot = atomicAdd(p.next_count, 1);
        p.next_list[slot] = pair;
      }
      return;
    }
    if (c.leader) {
      p.status[pair] = c.spill ? kStCapacity : st;
      p.wf_score[pair] = wf;
      p.score[pair] = c.alg_score;
      p.nruns[pair] = 0;
    }
    return;
  }
  if (plen + tlen < p.task_min_length) {   // short pair: depth first on this aligner
    if (pair != staged) {
      stage_pair(c, p, pair, lds, wv);
      staged = pair;
    }
    c.out = p.runs + p.runs_off[pair];
    c.out_cap = plen + tlen + 1;
    c.out_n = 0; c.pend_op = -1; c.pend_len = 0; c.alg_score = 0;
    c.spill = 0; c.cap_err = 0; c.arena_top = 0;
    int wf = INT_MIN;
    int st = align_pair(c, pair, plen, tlen, 0, &wf);
    if (c.spill) {
      if (p.next_list) {
        if (c.leader) {
          const int slot = atomicAdd(p.next_count, 1);
          p.next_list[slot] = pair;
        }
        return;
      }
      st = kStCapacity;
    }
    if (c.cap_err) st = kStCapacity;
    if (c.leader) {
      p.status[pair] = st;
      p.wf_score[pair] = wf;
      p.score[pair] = (st == 0 && c.out_n > 0) ? c.alg_score : -1;
      p.nruns[pair] = st == 0 ? c.out_n : 0;
    }
    return;
  }
  uint32_t* base = p.runs + p.runs_off[pair];
  const int t0 = K::kNW > 1 ? (int)threadIdx.x : c.lane;
  for (int i = t0; i < plen + tlen + 1; i += K::kNW * kWave) base[i] = 0u;
  if (c.leader) {
    atomicAdd(&p.qctl[4], 1);   // a shared-task pair in flight: it may publish items
    p.pend[pair] = 1;
    p.pflag[pair] = 0;
    p.pfail[pair] = ~0ull;
    p.wf_score[pair] = INT_MIN;
  }
  __threadfence();
  team_sync<K>();
  const bool min_length = imax(plen, tlen) <= kFallbackMinLength;
  run_task(c, p, Sub{pair, 0, plen, 0, tlen, CM, CM, min_length ? 0 : INT_MAX, 0}, staged, lds, wv);
}

// Work loop of a global tier in shared-task mode: queued sub-problems first,
// then new pairs; ends when the list is drained, the queue is empty and no
// shared-task pair is in flight (only those publish items; the in-flight count
// drops after a pair's last sub-problem, which published nothing more).
template <class K>
DI void task_loop(Ctx<K>& c, const KParams& p, int count, int32_t* lds, int wv) {
  int32_t* q = p.qctl;
  int staged = -1;
  int idle = 0;
  for (;;) {
    int kind = 0, arg = 0;   // 1: queued item (arg = slot), 2: new pair (arg = id), -1: done
    if (c.leader) {
      for (;;) {
        // an aligner gave up on a claimed item: its pair can never finish, so
        // the in-flight count never drops; everyone leaves and the host reports it
        if (atomic_ld(&q[5]) != 0) { kind = -1; break; }
        const int h = atomic_ld(&q[2]);
        const int t = imin(atomic_ld(&q[3]), p.item_cap);
        if (h < t) {
          if (atomicCAS(&q[2], h, h + 1) == h) { kind = 1; arg = h; break; }
          continue;
        }
        if (atomic_ld(&q[0]) < count) {
          const int b = atomicAdd(&q[0], 1);
          if (b < count) { kind = 2; arg = p.list[b]; break; }
        }
        // list drained: more work can only come from shared-task pairs still in
        // flight (their items); with none, this aligner is done
        const int inflight = atomic_ld(&q[4]);
        __threadfence();
        const int h2 = atomic_ld(&q[2]);
        const int t2 = imin(atomic_ld(&q[3]), p.item_cap);
        kind = (inflight == 0 && h2 >= t2) ? -1 : 0;
        break;
      }
    }
    team_bcast2(c, kind, arg);
    if (kind < 0) break;
    if (kind == 0) {   // wait for items, backing off (every poll hits the same counters)
      for (int z = 0; z < idle; ++z) __builtin_amdgcn_s_sleep(127);
      idle = imin(2 * idle + 1, 32);
      continue;
    }
    idle = 0;
    if (kind == 1) {
      int32_t* it = p.items + (int64_t)arg * kItemInts;
      int ready = 1, unused = 0;
      if (c.leader) {
        const uint64_t t0 = wall_clock64();
        while (atomic_ld(it) != c.epoch) {
          if (wall_clock64() - t0 > (uint64_t)p.ready_ticks || atomic_ld(&q[5]) != 0) { ready = 0; break; }
          __builtin_amdgcn_s_sleep(2);
        }
      }
      team_bcast2(c, ready, unused);
      if (!ready) {
        if (c.leader) atomicOr(&q[5], 2);
        break;
      }
      __threadfence();
      const Sub s{uni(it[1]), uni(it[2]), uni(it[3]), uni(it[4]), uni(it[5]),
                  uni(it[6]), uni(it[7]), uni(it[8]), uni(it[9])};
      run_task(c, p, s, staged, lds, wv);
    } else {
      start_pair(c, p, arg, staged, lds, wv);
    }
  }
}

// ---------------------------------------------------------------------------
// kernels

// Every instantiation is held to 128 vector registers, 4 waves per SIMD: the
// team kernels then spill a few dozen registers outside their step loops, which
// costs less than the latency the fourth wave hides.
template <class K>
__global__ __launch_bounds__(K::kNW > 1 ? K::kNW * kWave : kWave * kWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(4, 8)))
void biwfa_persistent(KParams p) {
  extern __shared__ int32_t lds[];
  __shared__ int32_t tred[K::kNW > 1 ? 2 * K::kNW * kRedSlots : 1];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  const int wid = K::kNW > 1 ? wv : 0;
  const int64_t team = K::kNW > 1 ? (int64_t)blockIdx.x : (int64_t)blockIdx.x * (blockDim.x / kWave) + wv;
  int32_t* ws = K::kLds ? lds + wv * p.L.total : p.gws + team * p.gws_stride;
  Ctx<K> c;
  c.ws = ws;
  c.wd = ws;   // the layout puts the offsets first (L.data == 0)
  c.pv = ws + p.L.priv + wid * p.L.priv_ints;
  c.lane = lane;
  c.wid = wid;
  c.leader = wid == 0 && lane == 0;
  c.tred = tred;
  c.tph = 0;
  c.L = p.L;
  if constexpr (K::kDefPen) {   // compile-time constants; the launch checks the aligner's penalties
    c.x = kDefX; c.o = kDefO; c.e = kDefE; c.scope = kDefScope;
  } else {
    c.x = p.x; c.o = p.o; c.e = p.e; c.scope = p.scope;
  }
  c.max_steps = p.max_steps;
  if (K::kLds) {
    c.upen = lane == 0 ? p.um : lane == 1 ? p.ux : lane == 2 ? p.uo : p.ue;
    // Kept as vector values: as compares they would become scalar lane masks
    // hoisted out of every loop, and the kernel has no scalar registers to spare.
    const int cq = lane <= 1 ? CM : lane == 2 ? CI : CD;
    int m0 = -(int)(cq == CM), m1 = -(int)(cq == CI), m2 = -(int)(cq == CD);
    int al = lane == 0 ? 0 : lane == 2 ? 1 : -1, ah = lane == 0 ? 0 : lane == 3 ? -1 : 1;
    int qd = lane == 0 ? c.x : lane == 1 ? c.o + c.e : c.e, qc = cq;
    const int oc = lane % 3;
    int o0 = -(int)(oc == 0), o1 = -(int)(oc == 1), o2 = -(int)(oc == 2);
    asm volatile("" : "+v"(m0), "+v"(m1), "+v"(m2), "+v"(al), "+v"(ah));
    asm volatile("" : "+v"(qd), "+v"(qc), "+v"(o0), "+v"(o1), "+v"(o2));
    c.qm0 = m0; c.qm1 = m1; c.qm2 = m2;
    c.qlo = al; c.qhi = ah;
    c.qdelta = qd; c.qcomp = qc;
    c.om0 = o0; c.om1 = o1; c.om2 = o2;
  } else {
    c.um = p.um; c.ux = p.ux; c.uo = p.uo; c.ue = p.ue;
  }
  c.gmask = K::kDefPen ? kDefGmask : p.gmask;
  c.epoch = (uni(*(volatile const int32_t*)p.calls) << 3) | p.epoch;
  const int count = uni(*(volatile const int32_t*)p.count);
  if constexpr (!K::kLds) {
    // The list is final at launch (the tier below has finished); with no pairs
    // there are no shared tasks either, so an empty tier ends at once instead
    // of every aligner polling the queue counters.
    if (count == 0) return;
    task_loop(c, p, count, lds, wv);
  } else {
  // LDS tier: one wave per pair, depth-first recursion on the wave alone
  for (;;) {
    int base = 0;
    if (c.leader) base = atomicAdd(p.queue, p.chunk);
    base = uni(__shfl(base, 0, kWave));
    if (base >= count) break;
    const int end = imin(base + p.chunk, count);
    for (int w = base; w < end; ++w) {
      const int pair = uni(p.list[w]);
      const int plen = uni(p.p_len[pair]), tlen = uni(p.t_len[pair]);
      const uint64_t* gP = p.words + p.p_woff[pair];
      const uint64_t* gT = p.words + p.t_woff[pair];
      c.arena_top = 0;
      for (int q = 0; q < 3; ++q) {
        c.rlo[q] = kAbsent; c.rhi[q] = 0;
        c.blo[q] = kAbsent; c.bhi[q] = 0; c.bkb[q] = 0;
      }
      c.out = p.runs + p.runs_off[pair];
      c.out_cap = (int)(p.runs_off[pair + 1] - p.runs_off[pair]);
      c.out_n = 0; c.pend_op = -1; c.pend_len = 0; c.alg_score = 0;
      c.spill = 0; c.cap_err = 0; c.guard = 0;
      const int pw = (plen + 31) >> 5, tw = (tlen + 31) >> 5;
      // stage both sequences: [pad][words...][pad][pad] (the pads keep the
      // 64-bit windows in bounds at either end)
      if (pw + 3 > p.L.seq_words || tw + 3 > p.L.seq_words) {
        c.spill = 1;
      } else {
        uint64_t* sP = reinterpret_cast<uint64_t*>(ws + p.L.seq);
        uint64_t* sT = sP + p.L.seq_words;
        for (int i = lane; i < pw + 3; i += kWave) {
          sP[i] = (i >= 1 && i <= pw) ? gP[i - 1] : 0ull;
        }
        for (int i = lane; i < tw + 3; i += kWave) {
          sT[i] = (i >= 1 && i <= tw) ? gT[i - 1] : 0ull;
        }
        c.P = sP + 1;
        c.T = sT + 1;
        c.seq_lds = 1;
        wave_sync<K::kLds>();
      }
      int wf = INT_MIN, st = kStCapacity;
      if (!c.spill) st = align_pair(c, pair, plen, tlen, p.score_only, &wf);
      if (c.spill) {
        if (p.next_list) {      // re-run by the next tier with larger capacities
          if (c.leader) {
            const int slot = atomicAdd(p.next_count, 1);
            p.next_list[slot] = pair;
          }
          continue;
        }
        st = kStCapacity;
      }
      if (c.cap_err) st = kStCapacity;
      if (c.leader) {
        p.status[pair] = st;
        p.wf_score[pair] = wf;
        p.score[pair] = (p.score_only || c.out_n > 0) ? c.alg_score : -1;
        p.nruns[pair] = p.score_only ? 0 : c.out_n;
      }
    }
  }
  }
}

// ---------------------------------------------------------------------------
// Explicit instantiations and the small kernels.  The build compiles this file
// once per group (BW_PART = 0 .. kParts-1) so the groups build in parallel;
// without BW_PART one translation unit holds everything.
#if !defined(BW_PART)
#define BW_GROUP(n) 1
#else
#define BW_GROUP(n) (BW_PART == (n))
#endif

#if BW_GROUP(0)
// Sets every tier's counters for a new launch sequence (a kernel rather than a
// memcpy or memset, so that the sequence can be captured in a graph) and advances the call
// counter at counters[n] (the ready-flag epochs of the sequence derive from
// it, on the device, so a captured sequence replays with fresh epochs).
__global__ void biwfa_reset(int32_t* counters, const int32_t* init, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = init[i];
  if (threadIdx.x == 0) counters[n] = (counters[n] + 1) & 0x0fffffff;
}

// No-allocation path (bw_align_device_noalloc): zeroes its two tiers' counters
// (the admit kernel counts the LDS tier's pairs) and advances the call counter
// at counters[n].  A kernel, not a memset, so a caller may capture the calls.
__global__ void biwfa_noalloc_reset(int32_t* counters, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0;
  if (threadIdx.x == 0) counters[n] = (counters[n] + 1) & 0x0fffffff;
}

// No-allocation path: admits every well-formed pair with p + t <= max_pt to the
// LDS tier's list (in whatever order the waves reach it: results do not depend
// on the order).  A pair with a bad length, word offset or run capacity
// (runs_off[i+1] - runs_off[i] < p + t + 1) gets BW_STATUS_INVALID_ARG, a longer
// one BW_STATUS_DEVICE_CAPACITY; neither is aligned.
__global__ __launch_bounds__(256) void biwfa_noalloc_admit(KParams p, int64_t n, int64_t num_words, int max_pt,
                                                           int32_t* list, int32_t* count) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    int st = 1;   // 1: not a pair of this call
    if (i < n) {
      const int pl = p.p_len[i], tl = p.t_len[i];
      const int64_t pw = p.p_woff[i], tw = p.t_woff[i];
      const int64_t r0 = p.runs_off[i], r1 = p.runs_off[i + 1];
      st = 0;
      if (pl < 0 || tl < 0 || pw < 1 || tw < 1 || pw + (pl + 31) / 32 + 2 > num_words ||
          tw + (tl + 31) / 32 + 2 > num_words || r0 < 0 || r1 - r0 < (int64_t)pl + tl + 1)
        st = kStInvalid;
      else if ((int64_t)pl + tl > max_pt)
        st = kStCapacity;
      if (st != 0) {
        p.status[i] = st;
        p.score[i] = -1;
        p.wf_score[i] = INT_MIN;
        p.nruns[i] = 0;
      }
    }
    // one atomic per wave for its admitted pairs
    const uint64_t adm = __ballot(st == 0);
    int first = 0;
    if ((threadIdx.x & (kWave - 1)) == 0 && adm) first = atomicAdd(count, __builtin_popcountll(adm));
    first = __shfl(first, 0, kWave);
    if (st == 0) {
      const int r = __builtin_amdgcn_mbcnt_hi((uint32_t)(adm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)adm, 0));
      list[first + r] = (int32_t)i;
    }
  }
}

// Copies each pair's runs into the compacted layout (one wave per pair).
__global__ __launch_bounds__(256) void biwfa_compact(const uint32_t* runs, const int64_t* runs_off,
                                                     const int32_t* nruns, const int64_t* out_off,
                                                     uint32_t* out, int n) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (wave >= n) return;
  const uint32_t* src = runs + runs_off[wave];
  uint32_t* dst = out + out_off[wave];
  const int cnt = nruns[wave];
  for (int i = lane; i < cnt; i += kWave) dst[i] = src[i];
}

template __global__ void biwfa_persistent<Cfg<true, 1>>(KParams);
#endif
#if BW_GROUP(1)
template __global__ void biwfa_persistent<Cfg<true, 1, kLdsDefaultWidth>>(KParams);
#endif
#if BW_GROUP(2)
template __global__ void biwfa_persistent<Cfg<true, 1, kLdsDefaultWidth, true>>(KParams);
#endif
#if BW_GROUP(3)
template __global__ void biwfa_persistent<Cfg<false, 1>>(KParams);
#endif
#if BW_GROUP(4)
template __global__ void biwfa_persistent<Cfg<false, 2>>(KParams);
#endif
#if BW_GROUP(5)
template __global__ void biwfa_persistent<Cfg<false, 4>>(KParams);
#endif
#if BW_GROUP(6)
template __global__ void biwfa_persistent<Cfg<false, 8>>(KParams);
#endif
#if BW_GROUP(7)
template __global__ void biwfa_persistent<Cfg<false, 16>>(KParams);
#endif

}  // namespace bw
