// biwfa_affine2p.hip -- gfx950 bidirectional WFA under gap-affine-2p penalties
// (score + CIGAR): one wave64 per pair.
//
// Behaviour follows WFA2-lib v2.3 `wavefront_align` in exact BiWFA mode with
// distance_metric = gap_affine_2p (memory_mode = ultralow, heuristic = none):
//   recursion driver      wavefront/wavefront_bialign.c:581-658, 703-730
//   breakpoint search     wavefront/wavefront_bialign.c:411-519 (overlap bound
//                         max(o1, o2): wavefront_bialign_overlap_gopen_adjust, :396-410)
//   overlap detection     wavefront/wavefront_bialign.c:189-392 (D2, I2, D1, I1, M)
//   affine-2p compute     wavefront/wavefront_compute_affine2p.c:45-108, 334-367
//                         wavefront/wavefront_compute.c:40-86, 306-352, 420-493, 579-632
//   extend / termination  wavefront/wavefront_extend.c:86-211, wavefront_termination.c:37-114
//   base case + backtrace wavefront/wavefront_bialign.c:155-185,
//                         wavefront/wavefront_unialign.c:54-273,
//                         wavefront/wavefront_backtrace.c:38-140, 320-529
//
// Device mapping:
//   * A persistent grid of 4-wave workgroups; every wave is an independent
//     aligner that takes pairs from an atomic counter over the schedule order.
//   * Diagonals are spread over the 64 lanes in chunks of 64; scores, limits
//     and the recursion state are wave-uniform (scalar registers).
//   * Each of the seven inputs of a step (M[s-x], M[s-o1-e1], M[s-o2-e2],
//     I1/D1[s-e1], I2/D2[s-e2]) is read once per lane; the five outputs are
//     written once; M is extended in the same pass (32 bases per 64-bit XOR of
//     2-bit packed words); trims, the maximum antidiagonal and the end-diagonal
//     offset come out of per-chunk ballots and one wave reduction per step.
//   * Forward and reverse offsets live in modular rings (scope scores x 5
//     components x wcap diagonals, k indexed modulo wcap) in the wave's global
//     slab -- O(s) memory per pair -- with their lo/hi headers in LDS.  The base
//     case keeps every wavefront of its sub-problem in a bump arena over the same
//     slab, headers in the slab's private region.
//   * Uniform bookkeeping in memory (base-case headers, the backtrace's run
//     buffer, the recursion stack, output runs) is written by every lane with
//     the same value, so each lane's later read sees its own write.
#include <hip/hip_runtime.h>
#include <limits.h>

#include "biwfa_device.hpp"

namespace bw2p {

using bw::A2pParams;
using bw::kOffsetNull;
using bw::kWave;

#define DI __device__ __forceinline__

enum { CM = 0, CI1 = 1, CD1 = 2, CI2 = 3, CD2 = 4, NC = 5 };
constexpr int kAbsent = INT_MAX;   // header lo of a component WFA2 has not allocated
// uint32 run op codes (SAM encoding with mismatches shown, cigar.c:38-56)
constexpr int kOpM = 7, kOpX = 8, kOpI = 1, kOpD = 2;
// backtrace source types (wavefront_backtrace.c:49-59)
constexpr int kBtM = 9, kBtD2Ext = 8, kBtD2Open = 7, kBtD1Ext = 6, kBtD1Open = 5, kBtI2Ext = 4,
              kBtI2Open = 3, kBtI1Ext = 2, kBtI1Open = 1;
constexpr int kSubLeaf = 1, kSubSplit = 2;
constexpr int kHdrInts = 2 * bw::kA2pMaxScope * NC * 2;   // ring headers per wave (LDS)

DI int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
DI int rdl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
DI int imax(int a, int b) { return a >= b ? a : b; }
DI int imin(int a, int b) { return a <= b ? a : b; }

// Orders this wave's workspace writes before later reads by other lanes.
DI void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

DI int wave_max(int v) {
  for (int d = 32; d >= 1; d >>= 1) v = imax(v, __shfl_xor(v, d, kWave));
  return uni(v);
}

// A wavefront component as readers see it: offsets for k in [lo, hi] at
// wd[base + (k & mask)] (mask = wcap-1 in a ring, -1 in the arena).  lo > hi is
// the null wavefront (WFA2's wavefront_null has lo=1, hi=-1).
struct Wf {
  int lo, hi, base, mask;
};
DI Wf wf_null() { return Wf{1, -1, 0, 0}; }

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

struct Sub {
  int pb, pe, tb, te, cb, ce, rem, lvl;
};

struct Ctx {
  int32_t* wd;              // rings / base-case arena (slab base)
  int32_t* bh;              // base-case headers: hcap x NC x {lo, hi, base}
  int32_t* rt;              // reversed runs of one backtrace
  int32_t* stk;             // recursion stack
  int32_t* rh;              // ring headers (LDS): 2 x scope x NC x {lo, hi}
  const uint64_t* P;
  const uint64_t* T;
  int lane;
  int x, o1, e1, o2, e2, scope, max_steps;
  int um, ux, uo1, ue1, uo2, ue2;
  int wcap, arena_cap, hcap, rcap, scap;
  int arena_top;
  uint32_t* out;
  int out_cap, out_n, pend_op, pend_len, alg_score, cap_err;
  int spill;                // a workspace capacity was exceeded
  int guard;                // score bound of the current sub-problem
};

// ---------------------------------------------------------------------------
// headers

DI int ring_slot_base(const Ctx& c, int ring, int slot, int comp) {
  return ((ring * c.scope + slot) * NC + comp) * c.wcap;
}
DI int ring_hdr(const Ctx& c, int ring, int slot, int comp) { return ((ring * c.scope + slot) * NC + comp) * 2; }

// lo == kAbsent marks an unallocated component
DI void set_hdr(Ctx& c, const Dir& d, int score, int slot, int comp, int lo, int hi, int base) {
  if (d.ring >= 0) {
    const int i = ring_hdr(c, d.ring, slot, comp);
    c.rh[i] = lo;
    c.rh[i + 1] = hi;
  } else {
    int32_t* h = c.bh + (score * NC + comp) * 3;
    h[0] = lo;
    h[1] = hi;
    h[2] = base;
  }
}

// Component `comp` at `score` (ring slot `slot`): wavefront_compute_get_*wavefront
// (wavefront_compute.c:266-305).  *present reports a non-NULL WFA2 pointer
// (possibly with lo > hi).
DI Wf get_wf(const Ctx& c, const Dir& d, int score, int slot, int comp, int* present) {
  *present = 0;
  if (score < 0) return wf_null();
  int lo, hi, base, mask;
  if (d.ring >= 0) {
    const int i = ring_hdr(c, d.ring, slot, comp);
    lo = uni(c.rh[i]);
    hi = uni(c.rh[i + 1]);
    base = ring_slot_base(c, d.ring, slot, comp);
    mask = c.wcap - 1;
  } else {
    const int32_t* h = c.bh + (score * NC + comp) * 3;
    lo = uni(h[0]);
    hi = uni(h[1]);
    base = uni(h[2]);
    mask = -1;
  }
  if (lo == kAbsent) return wf_null();
  *present = 1;
  return Wf{lo, hi, base, mask};
}

// A step input: score - delta (ring slot slot - delta); an absent or empty
// wavefront reads as the null wavefront (lo=1, hi=-1) and *nonempty is 0.
DI Wf get_input(const Ctx& c, const Dir& d, int s, int slot, int delta, int comp, bool* nonempty) {
  int sl = slot - delta;
  if (sl < 0) sl += c.scope;
  int present;
  const Wf w = get_wf(c, d, s - delta, d.ring >= 0 ? sl : 0, comp, &present);
  *nonempty = present && w.lo <= w.hi;
  return *nonempty ? w : wf_null();
}

DI int rd(const Ctx& c, const Wf& w, int k) {
  int v = kOffsetNull;
  if (k >= w.lo && k <= w.hi) v = c.wd[w.base + (k & w.mask)];
  return v;
}

// ---------------------------------------------------------------------------
// sequences: 2-bit packed, base j at bits 2*(j%32) of word j/32

// Bases [32w + sh, 32w + sh + 32) from the two words lo = word w, hi = word w+1.
DI uint64_t window2(uint64_t lo, uint64_t hi, int sh) {
  const int b = sh << 1;
  return (lo >> b) | ((hi << 1) << (63 - b));
}

// Longest common extension from (v, h) of the sub-problem, 32 bases per
// 64-bit compare (forward: first mismatch = lowest differing base; reverse:
// highest), clamped to the sub-problem's end.
DI int lcp(const Ctx& c, const Dir& d, int v, int h) {
  const int lim = imin(d.plen - v, d.tlen - h);
  const bool rev = d.rev != 0;
  const int pp = rev ? d.pb + d.plen - 32 - v : d.pb + v;
  const int tp = rev ? d.tb + d.tlen - 32 - h : d.tb + h;
  const int sp = pp & 31, st = tp & 31, dw = rev ? -1 : 1;
  int wp = pp >> 5, wt = tp >> 5;
  int n = 0;
  for (;;) {
    const uint64_t x = window2(c.P[wp], c.P[wp + 1], sp) ^ window2(c.T[wt], c.T[wt + 1], st);
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

// ---------------------------------------------------------------------------
// one step: wavefront_compute_affine2p (wavefront_compute_affine2p.c:334-367)
// fused with wavefront_extend_end2end(_max) (wavefront_extend.c:86-211) and the
// termination test (wavefront_termination.c:37-114).  The lane that computes
// M[s][k] extends it at once; extension only reads M[s][k] and the sequences
// and the trims are taken before it, so the result equals compute-all-then-
// extend.  Returns 1 when the aligner is done (end reached, or END_UNREACHABLE
// after scope+1 null steps); *max_ak gets the maximum antidiagonal.

DI bool in_bounds(int v, int k, uint32_t tl, uint32_t pl) {
  return (uint32_t)v <= tl && (uint32_t)(v - k) <= pl;
}

DI int step(Ctx& c, Dir& d, int s, int slot, int* max_ak) {
  *max_ak = 0;
  if (d.ring < 0 && s >= c.hcap) { c.spill = 1; return 0; }
  bool n_mm, n_mo1, n_mo2, n_i1, n_d1, n_i2, n_d2;
  const Wf mm = get_input(c, d, s, slot, c.x, CM, &n_mm);
  const Wf mo1 = get_input(c, d, s, slot, c.o1 + c.e1, CM, &n_mo1);
  const Wf mo2 = get_input(c, d, s, slot, c.o2 + c.e2, CM, &n_mo2);
  const Wf i1 = get_input(c, d, s, slot, c.e1, CI1, &n_i1);
  const Wf d1 = get_input(c, d, s, slot, c.e1, CD1, &n_d1);
  const Wf i2 = get_input(c, d, s, slot, c.e2, CI2, &n_i2);
  const Wf d2 = get_input(c, d, s, slot, c.e2, CD2, &n_d2);
  if (!(n_mm || n_mo1 || n_mo2 || n_i1 || n_d1 || n_i2 || n_d2)) {   // null step (:346-357)
    d.null_steps++;
    for (int q = 0; q < NC; ++q) set_hdr(c, d, s, slot, q, kAbsent, 0, 0);
    if (d.null_steps > c.scope) {               // nothing left to extend (wavefront_extend.c)
      d.status = bw::kStEndUnreachable;
      d.status_score = s;
      return 1;
    }
    return 0;
  }
  d.null_steps = 0;
  // wavefront_compute_limits_input (wavefront_compute.c:40-86): null inputs
  // contribute the null wavefront's lo=1, hi=-1
  int lo = mm.lo, hi = mm.hi;
  lo = imin(lo, mo1.lo - 1); hi = imax(hi, mo1.hi + 1);
  lo = imin(lo, i1.lo + 1);  hi = imax(hi, i1.hi + 1);
  lo = imin(lo, d1.lo - 1);  hi = imax(hi, d1.hi - 1);
  lo = imin(lo, mo2.lo - 1); hi = imax(hi, mo2.hi + 1);
  lo = imin(lo, i2.lo + 1);  hi = imax(hi, i2.hi + 1);
  lo = imin(lo, d2.lo - 1);  hi = imax(hi, d2.hi - 1);
  // outputs allocated when one of their inputs exists (wavefront_compute.c:454-493)
  const bool has_i1 = n_mo1 || n_i1, has_d1 = n_mo1 || n_d1;
  const bool has_i2 = n_mo2 || n_i2, has_d2 = n_mo2 || n_d2;
  const int width = hi - lo + 1;
  int bm, bi1, bd1, bi2, bd2, mask;
  if (d.ring >= 0) {
    if (width > c.wcap) { c.spill = 1; return 0; }
    bm = ring_slot_base(c, d.ring, slot, CM);
    bi1 = ring_slot_base(c, d.ring, slot, CI1);
    bd1 = ring_slot_base(c, d.ring, slot, CD1);
    bi2 = ring_slot_base(c, d.ring, slot, CI2);
    bd2 = ring_slot_base(c, d.ring, slot, CD2);
    mask = c.wcap - 1;
  } else {
    const int need = width * (1 + (int)has_i1 + (int)has_d1 + (int)has_i2 + (int)has_d2);
    if (c.arena_top + need > c.arena_cap) { c.spill = 1; return 0; }
    int top = c.arena_top;
    bm = top - lo; top += width;
    bi1 = top - lo; if (has_i1) top += width;
    bd1 = top - lo; if (has_d1) top += width;
    bi2 = top - lo; if (has_i2) top += width;
    bd2 = top - lo; if (has_d2) top += width;
    c.arena_top = top;
    mask = -1;
  }
  const uint32_t tl = (uint32_t)d.tlen, pl = (uint32_t)d.plen;
  const int k_end = d.tlen - d.plen;
  const int ce = d.ce;
  int f[NC], l[NC];
  for (int q = 0; q < NC; ++q) { f[q] = INT_MAX; l[q] = INT_MIN; }
  int lmax = 0, endu = INT_MIN;
  for (int kb = lo; kb <= hi; kb += kWave) {
    const int k = kb + c.lane;
    const bool act = k <= hi;
    // wavefront_compute_affine2p_idm (wavefront_compute_affine2p.c:45-108)
    const int ins1 = imax(rd(c, mo1, k - 1), rd(c, i1, k - 1)) + 1;
    const int ins2 = imax(rd(c, mo2, k - 1), rd(c, i2, k - 1)) + 1;
    const int del1 = imax(rd(c, mo1, k + 1), rd(c, d1, k + 1));
    const int del2 = imax(rd(c, mo2, k + 1), rd(c, d2, k + 1));
    int mx = imax(imax(del1, del2), imax(rd(c, mm, k) + 1, imax(ins1, ins2)));
    const bool live = act && in_bounds(mx, k, tl, pl);
    if (live) {                                 // extension along the diagonal
      mx += lcp(c, d, mx - k, mx);
      lmax = imax(lmax, 2 * mx - k);
    } else {
      mx = kOffsetNull;
    }
    if (act) {
      c.wd[bm + (k & mask)] = mx;
      if (has_i1) c.wd[bi1 + (k & mask)] = ins1;
      if (has_d1) c.wd[bd1 + (k & mask)] = del1;
      if (has_i2) c.wd[bi2 + (k & mask)] = ins2;
      if (has_d2) c.wd[bd2 + (k & mask)] = del2;
    }
    // in-bounds flags for wavefront_compute_trim_ends (wavefront_compute.c:579-613)
    const uint64_t b[NC] = {__ballot(live), __ballot(act && in_bounds(ins1, k, tl, pl)),
                            __ballot(act && in_bounds(del1, k, tl, pl)), __ballot(act && in_bounds(ins2, k, tl, pl)),
                            __ballot(act && in_bounds(del2, k, tl, pl))};
    for (int q = 0; q < NC; ++q) {
      if (b[q]) {
        if (f[q] == INT_MAX) f[q] = kb + __builtin_ctzll(b[q]);
        l[q] = kb + 63 - __builtin_clzll(b[q]);
      }
    }
    if (k_end >= kb && k_end <= imin(hi, kb + kWave - 1)) {
      const int ev = ce == CM ? mx : ce == CI1 ? ins1 : ce == CD1 ? del1 : ce == CI2 ? ins2 : del2;
      endu = rdl(ev, k_end - kb);
    }
  }
  wave_sync();
  const bool has[NC] = {true, has_i1, has_d1, has_i2, has_d2};
  const int base[NC] = {bm, bi1, bd1, bi2, bd2};
  int tlo[NC], thi[NC];
  for (int q = 0; q < NC; ++q) {
    // trimmed limits: hi = last in-bounds (else lo-1), lo = first in-bounds (else lo)
    tlo[q] = l[q] == INT_MIN ? lo : f[q];
    thi[q] = l[q] == INT_MIN ? lo - 1 : l[q];
    set_hdr(c, d, s, slot, q, has[q] ? tlo[q] : kAbsent, thi[q], base[q]);
  }
  lmax = wave_max(lmax);
  // wavefront_termination_end2end: the end component holds k_end with offset >= tlen
  if (has[ce] && tlo[ce] <= k_end && k_end <= thi[ce] && endu >= d.tlen) {
    d.status = bw::kStEndReached;
    d.status_score = s;
    return 1;
  }
  *max_ak = lmax;
  return 0;
}

// Extension and termination test of the wavefronts at score 0 (the begin
// component's single offset); returns 1 when the aligner is done.
DI int extend_first(Ctx& c, Dir& d, int* max_ak) {
  *max_ak = 0;
  int present;
  const Wf m = get_wf(c, d, 0, 0, CM, &present);
  if (!present) return 0;   // no M wavefront: only the null-step count is tested (it is 0)
  int lmax = 0;
  for (int kb = m.lo; kb <= m.hi; kb += kWave) {
    const int k = kb + c.lane;
    if (k <= m.hi) {
      const int a = m.base + (k & m.mask);
      int off = c.wd[a];
      if (off != kOffsetNull) {
        off += lcp(c, d, off - k, off);
        c.wd[a] = off;
        lmax = imax(lmax, 2 * off - k);
      }
    }
  }
  wave_sync();
  lmax = wave_max(lmax);
  const int k_end = d.tlen - d.plen;
  int pe;
  const Wf e = get_wf(c, d, 0, 0, d.ce, &pe);
  if (pe && e.lo <= k_end && k_end <= e.hi && uni(c.wd[e.base + (k_end & e.mask)]) >= d.tlen) {
    d.status = bw::kStEndReached;
    d.status_score = 0;
    return 1;
  }
  *max_ak = lmax;
  return 0;
}

// ---------------------------------------------------------------------------
// breakpoint detection (wavefront_bialign.c:189-392)

// First k0 (ascending) of the overlapping interval that meets the overlap
// condition; comp CM: M-to-M, else indel-to-indel paying gap_open.
DI void scan(const Ctx& c, bool forward, int s0, int s1, const Wf& w0, const Wf& w1, int comp, int gap_open,
             int tl, int pl, BP& bp) {
  const int lo0 = w0.lo, hi0 = w0.hi;
  const int lo1 = tl - pl - w1.hi, hi1 = tl - pl - w1.lo;   // WAVEFRONT_K_INVERSE
  if (hi1 < lo0 || hi0 < lo1) return;
  const int min_hi = imin(hi0, hi1), max_lo = imax(lo0, lo1);
  const int sc = comp == CM ? s0 + s1 : s0 + s1 - gap_open;
  if (!(sc < bp.score)) return;
  for (int kb = max_lo; kb <= min_hi; kb += kWave) {
    const int k0 = kb + c.lane;
    const int k1 = tl - pl - k0;
    bool ok = false;
    int o0 = 0, o1 = 0;
    if (k0 <= min_hi) {
      o0 = c.wd[w0.base + (k0 & w0.mask)];
      o1 = c.wd[w1.base + (k1 & w1.mask)];
      ok = o0 + o1 >= tl;
      if (ok && comp != CM)   // indel: forward coordinates must be in bounds (:226-240)
        ok = forward ? !(o0 - k0 > pl || o0 > tl) : !(o1 - k1 > pl || o1 > tl);
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

// wavefront_bialign_overlap (wavefront_bialign.c:314-392) under gap-affine-2p:
// for every score of a1's scope the I2/D2 candidates first (skipping the rest
// of that score when even the second opening cannot improve), then I1/D1, then M.
DI void overlap(const Ctx& c, const Dir& a0, const Dir& a1, int s0, int slot0, int s1, int slot1, bool forward,
                BP& bp) {
  const int tl = a0.tlen, pl = a0.plen;
  int pm0, pi10, pd10, pi20, pd20;
  const Wf m0 = get_wf(c, a0, s0, slot0, CM, &pm0);
  if (!pm0) return;
  const Wf i10 = get_wf(c, a0, s0, slot0, CI1, &pi10);
  const Wf d10 = get_wf(c, a0, s0, slot0, CD1, &pd10);
  const Wf i20 = get_wf(c, a0, s0, slot0, CI2, &pi20);
  const Wf d20 = get_wf(c, a0, s0, slot0, CD2, &pd20);
  for (int i = 0; i < c.scope; ++i) {
    const int si = s1 - i;
    if (si < 0) break;
    int sl = slot1 - i;
    if (sl < 0) sl += c.scope;
    int p1;
    if (s0 + si - c.o2 >= bp.score) continue;
    const Wf dd2 = get_wf(c, a1, si, sl, CD2, &p1);
    if (pd20 && p1) scan(c, forward, s0, si, d20, dd2, CD2, c.o2, tl, pl, bp);
    const Wf ii2 = get_wf(c, a1, si, sl, CI2, &p1);
    if (pi20 && p1) scan(c, forward, s0, si, i20, ii2, CI2, c.o2, tl, pl, bp);
    if (s0 + si - c.o1 >= bp.score) continue;
    const Wf dd1 = get_wf(c, a1, si, sl, CD1, &p1);
    if (pd10 && p1) scan(c, forward, s0, si, d10, dd1, CD1, c.o1, tl, pl, bp);
    const Wf ii1 = get_wf(c, a1, si, sl, CI1, &p1);
    if (pi10 && p1) scan(c, forward, s0, si, i10, ii1, CI1, c.o1, tl, pl, bp);
    if (s0 + si >= bp.score) continue;
    const Wf mm1 = get_wf(c, a1, si, sl, CM, &p1);
    if (p1) scan(c, forward, s0, si, m0, mm1, CM, 0, tl, pl, bp);
  }
}

// wavefront_aligner_init_wf (wavefront_aligner.c:311-383) for a ring aligner:
// every slot absent, then the begin component's offset 0 at k=0, score 0.
DI void ring_init(Ctx& c, Dir& d) {
  for (int i = c.lane; i < c.scope * NC; i += kWave) c.rh[(d.ring * c.scope * NC + i) * 2] = kAbsent;
  wave_sync();
  set_hdr(c, d, 0, 0, d.cb, 0, 0, 0);
  c.wd[ring_slot_base(c, d.ring, 0, d.cb)] = 0;
  wave_sync();
  d.null_steps = 0;
  d.status = bw::kStOk;
  d.status_score = 0;
}

DI int next_slot(int slot, int scope) { return slot + 1 == scope ? 0 : slot + 1; }

// Upper bound of the optimal score of a plen x tlen sub-problem plus the
// slack the breakpoint search may run past it; a search beyond it can only be
// a defect and stops with kStGuard instead of looping.
DI int score_guard(const Ctx& c, int plen, int tlen) {
  const int mn = imin(plen, tlen), df = imax(plen, tlen) - mn;
  return mn * c.x + imin(c.o1 + df * c.e1, c.o2 + df * c.e2) + c.o1 + c.o2 + 2 * c.scope + 16;
}

// wavefront_bialign_find_breakpoint (wavefront_bialign.c:411-519)
DI int find_breakpoint(Ctx& c, Dir& f, Dir& r, BP& bp) {
  ring_init(c, f);
  ring_init(c, r);
  const int max_antidiagonal = f.plen + f.tlen - 1;
  int sf = 0, sr = 0, qf = 0, qr = 0, f_max = 0, r_max = 0, ak = 0;
  bp.score = INT_MAX;
  c.guard = score_guard(c, f.plen, f.tlen);
  if (extend_first(c, f, &f_max)) return f.status;
  if (extend_first(c, r, &r_max)) return r.status;
  bool last_forward = false;
  for (;;) {                                       // phase 1 (:447-480)
    if (f_max + r_max >= max_antidiagonal) break;
    if (sf + sr > c.guard) return bw::kStGuard;
    ++sf; qf = next_slot(qf, c.scope);
    const int fq = step(c, f, sf, qf, &ak);
    if (c.spill) return bw::kStCapacity;
    if (f_max < ak) f_max = ak;
    last_forward = true;
    if (fq) return f.status;
    if (f_max + r_max >= max_antidiagonal) break;
    ++sr; qr = next_slot(qr, c.scope);
    const int rq = step(c, r, sr, qr, &ak);
    if (c.spill) return bw::kStCapacity;
    if (r_max < ak) r_max = ak;
    last_forward = false;
    if (rq) return r.status;
    if (sr + sf >= c.max_steps) return bw::kStMaxSteps;
  }
  const int gap_open = imax(c.o1, c.o2);           // phase 2 (:482-516)
  for (;;) {
    if (last_forward) {
      const int min_r = (sr > c.scope - 1) ? sr - (c.scope - 1) : 0;
      if (sf + min_r - gap_open >= bp.score) break;
      overlap(c, f, r, sf, qf, sr, qr, true, bp);
      ++sr; qr = next_slot(qr, c.scope);
      const int rq = step(c, r, sr, qr, &ak);
      if (c.spill) return bw::kStCapacity;
      if (rq) return r.status;
    }
    const int min_f = (sf > c.scope - 1) ? sf - (c.scope - 1) : 0;
    if (min_f + sr - gap_open >= bp.score) break;
    if (sf + sr > c.guard) return bw::kStGuard;
    overlap(c, r, f, sr, qr, sf, qf, false, bp);
    ++sf; qf = next_slot(qf, c.scope);
    const int fq = step(c, f, sf, qf, &ak);
    if (c.spill) return bw::kStCapacity;
    if (fq) return f.status;
    if (sr + sf >= c.max_steps) return bw::kStMaxSteps;
    last_forward = true;
  }
  return bw::kStOk;
}

// ---------------------------------------------------------------------------
// output runs: merged into maximal runs as they are appended; the .alg score
// is cigar_score_gap_affine2p (cigar.c:304-342): a gap run costs the cheaper of
// its two models.

DI void out_flush(Ctx& c) {
  if (c.pend_len <= 0) return;
  if (c.out_n >= c.out_cap) {
    c.cap_err = 1;
  } else {
    c.out[c.out_n] = ((uint32_t)c.pend_len << 4) | (uint32_t)c.pend_op;
    c.out_n++;
  }
  const int n = c.pend_len;
  if (c.pend_op == kOpM) c.alg_score -= c.um * n;
  else if (c.pend_op == kOpX) c.alg_score -= c.ux * n;
  else c.alg_score -= imin(c.uo1 + c.ue1 * n, c.uo2 + c.ue2 * n);
  c.pend_len = 0;
  c.pend_op = -1;
}
DI void out_append(Ctx& c, int op, int n) {
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
DI void rev_emit(Ctx& c, RevRuns& rr, int op, int n) {
  if (n <= 0) return;
  if (op == rr.op) { rr.len += n; return; }
  if (rr.len > 0) {
    if (rr.n >= c.rcap) { c.spill = 1; return; }
    c.rt[rr.n] = (int)(((uint32_t)rr.len << 4) | (uint32_t)rr.op);
    rr.n++;
  }
  rr.op = op;
  rr.len = n;
}

// ---------------------------------------------------------------------------
// base case: wavefront_bialign_base (wavefront_bialign.c:155-185)

// The packed (offset << 4) | type maximum over the sources of (matrix, score, k)
// (wavefront_backtrace.c:64-219, 369-415): lane j < 9 evaluates source j (its
// header, then its offset), and the enabled sources are folded over lanes 0..8.
DI long long bt_max(const Ctx& c, int score, int matrix, int k) {
  const int j = c.lane;
  //         j:  0     1       2     3       4     5       6     7       8
  // source:     misms I1-open I1-ext D1-open D1-ext I2-open I2-ext D2-open D2-ext
  const int dscore = j == 0 ? c.x : (j == 1 || j == 3) ? c.o1 + c.e1 : (j == 2 || j == 4) ? c.e1
                   : (j == 5 || j == 7) ? c.o2 + c.e2 : c.e2;
  const int comp = j == 2 ? CI1 : j == 4 ? CD1 : j == 6 ? CI2 : j == 8 ? CD2 : CM;
  const int dk = j == 0 ? 0 : (j == 1 || j == 2 || j == 5 || j == 6) ? -1 : 1;
  const int add = (j == 0 || j == 1 || j == 2 || j == 5 || j == 6) ? 1 : 0;
  const int type = j == 0 ? kBtM : j == 1 ? kBtI1Open : j == 2 ? kBtI1Ext : j == 3 ? kBtD1Open : j == 4 ? kBtD1Ext
                 : j == 5 ? kBtI2Open : j == 6 ? kBtI2Ext : j == 7 ? kBtD2Open : kBtD2Ext;
  const bool en = j < 9 && (matrix == CM || (matrix == CI1 && (j == 1 || j == 2)) ||
                            (matrix == CD1 && (j == 3 || j == 4)) || (matrix == CI2 && (j == 5 || j == 6)) ||
                            (matrix == CD2 && (j == 7 || j == 8)));
  const int sj = score - dscore;
  long long v = LLONG_MIN;
  if (en) {
    v = (long long)kOffsetNull;
    if (sj >= 0) {
      const int32_t* h = c.bh + (sj * NC + comp) * 3;
      const int lo = h[0], hi = h[1], base = h[2];
      const int kk = k + dk;
      if (lo != kAbsent && lo <= kk && kk <= hi) v = (long long)(c.wd[base + kk] + add) * 16 | type;
    }
  }
  const int vlo = (int)(uint32_t)(unsigned long long)v, vhi = (int)((unsigned long long)v >> 32);
  long long m = LLONG_MIN;
  for (int q = 0; q < 9; ++q) {
    const long long w = (long long)(((unsigned long long)(uint32_t)rdl(vhi, q) << 32) | (uint32_t)rdl(vlo, q));
    m = w > m ? w : m;
  }
  return m;
}

// wavefront_backtrace_affine (wavefront_backtrace.c:320-529) with the I2/D2
// matrices; appends the operations to the output in forward order.  Returns 0,
// or -1 on a trace error.
DI int backtrace(Ctx& c, const Dir& b, int score) {
  int matrix = b.ce;
  int k = b.tlen - b.plen, off = b.tlen;
  int h = off, v = off - k;
  RevRuns rr{0, -1, 0};
  if (matrix == CM) {
    if (v < b.plen) rev_emit(c, rr, kOpD, b.plen - v);
    if (h < b.tlen) rev_emit(c, rr, kOpI, b.tlen - h);
  }
  while (v > 0 && h > 0 && score > 0) {
    const long long max_all = bt_max(c, score, matrix, k);
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
      case kBtM: score -= c.x; matrix = CM; rev_emit(c, rr, kOpX, 1); --off; break;
      case kBtI1Open: score -= c.o1 + c.e1; matrix = CM; rev_emit(c, rr, kOpI, 1); --k; --off; break;
      case kBtI1Ext: score -= c.e1; matrix = CI1; rev_emit(c, rr, kOpI, 1); --k; --off; break;
      case kBtI2Open: score -= c.o2 + c.e2; matrix = CM; rev_emit(c, rr, kOpI, 1); --k; --off; break;
      case kBtI2Ext: score -= c.e2; matrix = CI2; rev_emit(c, rr, kOpI, 1); --k; --off; break;
      case kBtD1Open: score -= c.o1 + c.e1; matrix = CM; rev_emit(c, rr, kOpD, 1); ++k; break;
      case kBtD1Ext: score -= c.e1; matrix = CD1; rev_emit(c, rr, kOpD, 1); ++k; break;
      case kBtD2Open: score -= c.o2 + c.e2; matrix = CM; rev_emit(c, rr, kOpD, 1); ++k; break;
      case kBtD2Ext: score -= c.e2; matrix = CD2; rev_emit(c, rr, kOpD, 1); ++k; break;
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
  } else if (v != 0 || h != 0 || (score != 0 && c.um == 0)) {
    return -1;   // "I?/D?-Beginning backtrace error" (wavefront_backtrace.c:517-524)
  }
  if (c.spill) return 0;
  // forward order: the pending run first, then the stored runs backwards
  out_append(c, rr.op, rr.len);
  for (int i = rr.n - 1; i >= 0; --i) {
    const uint32_t run = (uint32_t)uni(c.rt[i]);
    out_append(c, (int)(run & 15u), (int)(run >> 4));
  }
  return 0;
}

// Unidirectional WFA over the sub-problem with all wavefronts kept
// (wavefront_unialign.c:242-273).  Returns kStOk or kStUnattainable.
DI int base_case(Ctx& c, const Sub& s, int* score_out) {
  Dir b{s.pb, s.pe - s.pb, s.tb, s.te - s.tb, 0, -1, s.cb, s.ce, 0, bw::kStOk, 0};
  // score 0: the begin component at k=0 (wavefront_aligner.c:311-383)
  c.arena_top = 1;
  for (int q = 0; q < NC; ++q) set_hdr(c, b, 0, 0, q, q == s.cb ? 0 : kAbsent, 0, 0);
  c.wd[0] = 0;
  wave_sync();
  int score = 0, ak;
  const int guard = score_guard(c, b.plen, b.tlen);
  int done = extend_first(c, b, &ak);
  while (!done) {
    ++score;
    done = step(c, b, score, 0, &ak);
    if (c.spill) return bw::kStCapacity;
    if (score >= c.max_steps) return bw::kStUnattainable;   // reported as unattainable by the base
    if (score > guard) return bw::kStGuard;
  }
  if (b.status != bw::kStEndReached) return bw::kStUnattainable;
  if (backtrace(c, b, score) != 0) return bw::kStUnattainable;
  *score_out = score;
  return bw::kStOk;
}

// ---------------------------------------------------------------------------
// recursion driver: wavefront_bialign_alignment (wavefront_bialign.c:581-658)

DI void push(Ctx& c, int& sp, const Sub& s) {
  if (sp >= c.scap) { c.spill = 1; return; }
  int32_t* e = c.stk + sp * bw::kSubInts;
  e[0] = s.pb; e[1] = s.pe; e[2] = s.tb; e[3] = s.te;
  e[4] = s.cb; e[5] = s.ce; e[6] = s.rem; e[7] = s.lvl;
  ++sp;
}
DI Sub pop(Ctx& c, int& sp) {
  --sp;
  const int32_t* e = c.stk + sp * bw::kSubInts;
  return Sub{uni(e[0]), uni(e[1]), uni(e[2]), uni(e[3]), uni(e[4]), uni(e[5]), uni(e[6]), uni(e[7])};
}

// One sub-problem: trivial cases, the base case below the score threshold, or
// the breakpoint search with its exception path (:520-547).  Returns kSubLeaf,
// kSubSplit (children in *left / *right) or a failure status; *wf_score gets
// the level-0 score.
DI int run_sub(Ctx& c, const Sub& s, Sub* left, Sub* right, int* wf_score) {
  const int pl = s.pe - s.pb, tl = s.te - s.tb;
  if (tl == 0) { out_append(c, kOpD, pl); return kSubLeaf; }
  if (pl == 0) { out_append(c, kOpI, tl); return kSubLeaf; }
  int bscore = 0, st;
  if (s.rem <= bw::kFallbackMinScore) {
    st = base_case(c, s, &bscore);
    if (c.spill) return bw::kStCapacity;
    if (s.lvl == 0) *wf_score = bscore;
    return st == bw::kStOk ? kSubLeaf : st;
  }
  Dir f{s.pb, pl, s.tb, tl, 0, 0, s.cb, s.ce, 0, bw::kStOk, 0};
  Dir r{s.pb, pl, s.tb, tl, 1, 1, s.ce, s.cb, 0, bw::kStOk, 0};
  BP bp;
  st = find_breakpoint(c, f, r, bp);
  if (c.spill) return bw::kStCapacity;
  if (st != bw::kStOk) {   // wavefront_bialign_find_breakpoint_exception (:520-547)
    if (st != bw::kStEndReached) return st;
    const int reached = (f.status == bw::kStEndReached) ? f.status_score : r.status_score;
    if (reached > bw::kRecoveryMinScore) return bw::kStEndUnreachable;
    st = base_case(c, s, &bscore);
    if (c.spill) return bw::kStCapacity;
    if (s.lvl == 0) *wf_score = bscore;
    return st == bw::kStOk ? kSubLeaf : st;
  }
  if (s.lvl == 0) *wf_score = bp.score;
  const int bh = bp.of, bv = bp.of - bp.kf;
  *right = Sub{s.pb + bv, s.pe, s.tb + bh, s.te, bp.comp, s.ce, bp.sr, s.lvl + 1};
  *left = Sub{s.pb, s.pb + bv, s.tb, s.tb + bh, s.cb, bp.comp, bp.sf, s.lvl + 1};
  return kSubSplit;
}

// Final status of a failed alignment (wavefront_bialign.c:722-729).
DI int final_status(int st) {
  if (st == bw::kStMaxSteps || st == bw::kStGuard || st == bw::kStCapacity) return st;
  return bw::kStUnattainable;
}

// wavefront_bialign_compute_score (wavefront_bialign.c:662-699)
DI int score_pair(Ctx& c, int plen, int tlen, int* wf_score) {
  *wf_score = INT_MIN;
  Dir f{0, plen, 0, tlen, 0, 0, CM, CM, 0, bw::kStOk, 0};
  Dir r{0, plen, 0, tlen, 1, 1, CM, CM, 0, bw::kStOk, 0};
  BP bp;
  const int st = find_breakpoint(c, f, r, bp);
  if (c.spill) return bw::kStCapacity;
  if (st != bw::kStOk && st != bw::kStEndReached)
    return (st == bw::kStMaxSteps || st == bw::kStGuard) ? st : bw::kStUnattainable;
  int s = bp.score;
  if (st == bw::kStEndReached) s = (f.status == bw::kStEndReached) ? f.status_score : r.status_score;
  *wf_score = s;
  // wavefront_compute_classic_score (wavefront_compute.c:108-120)
  c.alg_score = (c.um == 0) ? -s : ((-c.um) * (plen + tlen) - s) / 2;
  return 0;
}

// Aligns one pair depth first, emitting the runs in order.  Returns the final
// WFA status (wavefront_bialign.c:703-730).
DI int align_pair(Ctx& c, int plen, int tlen, int score_only, int* wf_score) {
  if (score_only) return score_pair(c, plen, tlen, wf_score);
  *wf_score = INT_MIN;
  int sp = 0;
  const bool min_length = imax(plen, tlen) <= bw::kFallbackMinLength;
  push(c, sp, Sub{0, plen, 0, tlen, CM, CM, min_length ? 0 : INT_MAX, 0});
  int status = bw::kStOk;
  int iters = 0;
  const int max_iters = 4 * (plen + tlen) + 64;
  while (sp > 0) {
    if (++iters > max_iters) { status = bw::kStGuard; break; }
    const Sub s = pop(c, sp);
    Sub l, r;
    const int res = run_sub(c, s, &l, &r, wf_score);
    if (c.spill) return bw::kStCapacity;
    if (res == kSubLeaf) continue;
    if (res != kSubSplit) { status = res; break; }
    push(c, sp, r);
    push(c, sp, l);
    if (c.spill) return bw::kStCapacity;
  }
  out_flush(c);
  if (status == bw::kStOk) return 0;
  return final_status(status);
}

// ---------------------------------------------------------------------------
// kernels

// Persistent grid: every wave takes pairs in schedule order until the list is
// drained.  A wave beyond p.aligners has no slab and exits at once.
__global__ __launch_bounds__(kWave * bw::kA2pWavesPerBlock) void biwfa2p_persistent(A2pParams p) {
  __shared__ int32_t hdr[bw::kA2pWavesPerBlock][kHdrInts];
  const int wv = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t gw = (int64_t)blockIdx.x * bw::kA2pWavesPerBlock + wv;
  if (gw >= p.aligners) return;
  Ctx c;
  c.wd = p.gws + gw * p.stride;
  c.bh = c.wd + p.bhdr;
  c.rt = c.wd + p.runs_tmp;
  c.stk = c.wd + p.stack;
  c.rh = hdr[wv];
  c.lane = lane;
  c.x = p.x; c.o1 = p.o1; c.e1 = p.e1; c.o2 = p.o2; c.e2 = p.e2; c.scope = p.scope;
  c.max_steps = p.max_steps > 0 ? p.max_steps : INT_MAX;
  c.um = p.um; c.ux = p.ux; c.uo1 = p.uo1; c.ue1 = p.ue1; c.uo2 = p.uo2; c.ue2 = p.ue2;
  c.wcap = p.wcap; c.arena_cap = p.arena_cap; c.hcap = p.hcap; c.rcap = p.rcap; c.scap = p.scap;
  // No branch on the lane inside this loop: with `if (lane == 0)` around both
  // the queue take and the result stores, the compiler joined the two across
  // the loop's back edge and ran the index broadcast on lanes 1..63 alone,
  // which then re-read a stale index forever.  Every lane issues the atomic
  // (lane 0 adds 1, the others 0) and every lane stores the same results.
  for (;;) {
    const int idx = uni(atomicAdd(p.queue, lane == 0 ? 1 : 0));
    if (idx >= p.n) break;
    const int pair = uni(p.order[idx]);
    const int plen = uni(p.p_len[pair]), tlen = uni(p.t_len[pair]);
    c.P = p.words + p.p_woff[pair];
    c.T = p.words + p.t_woff[pair];
    c.out = p.runs + p.runs_off[pair];
    c.out_cap = plen + tlen + 1;
    c.out_n = 0; c.pend_op = -1; c.pend_len = 0; c.alg_score = 0; c.cap_err = 0;
    c.spill = 0; c.arena_top = 0;
    int wf = INT_MIN;
    int st = align_pair(c, plen, tlen, p.score_only, &wf);
    if (c.spill || c.cap_err) st = bw::kStCapacity;
    p.status[pair] = st;
    p.wf_score[pair] = wf;
    p.score[pair] = st != 0 ? -1 : (p.score_only || c.out_n > 0) ? c.alg_score : -1;
    p.nruns[pair] = (st != 0 || p.score_only) ? 0 : c.out_n;
  }
}

// Sets the launch's counters {queue, count, .., error} = {0, n, .., 0} (a
// kernel rather than a memset, so that a caller may capture the launch sequence).
__global__ void biwfa2p_reset(int32_t* qctl, int n) {
  if (threadIdx.x == 0) {
    qctl[0] = 0;
    qctl[1] = n;
    qctl[5] = 0;
  }
}

}  // namespace bw2p
