// biwfa_device.hpp -- shared host/device definitions of the gfx950 BiWFA kernels.
#pragma once
#include <stdint.h>

namespace bw {

// Offset value of an absent diagonal (WFA2-lib wavefront_offset.h:44).
constexpr int32_t kOffsetNull = (int32_t)(-2147483647 - 1) / 2;

// BiWFA thresholds (wavefront_bialign.c:48-50).
constexpr int kFallbackMinScore = 250;
constexpr int kFallbackMinLength = 100;
constexpr int kRecoveryMinScore = 500;

// Internal status codes (wfa.h:52-55) and device-only ones.
constexpr int kStOk = -1;
constexpr int kStEndReached = -2;
constexpr int kStEndUnreachable = -3;
constexpr int kStMaxSteps = -100;
constexpr int kStUnattainable = -300;
constexpr int kStCapacity = -400;
constexpr int kStInvalid = -402;     // a malformed pair (no-allocation path's admission)
constexpr int kStGuard = -405;       // termination guard tripped (never in a correct run)

constexpr int kMaxScope = 32;          // max(o+e, x)+1 supported by the ring layout
constexpr int kWave = 64;              // gfx950 wavefront size
constexpr int kWavesPerBlock = 4;      // independent single-wave aligners per workgroup
constexpr int kMaxTeam = 16;            // waves cooperating on one pair (one workgroup)
// LDS-tier workspace capacities (compile-time in its kernel): base-case scores
// (one header lane each), backtrace runs, recursion stack entries
constexpr int kLdsHcap = 64, kLdsRcap = 128, kLdsScap = 16;
constexpr int kRedSlots = 3 * kMaxScope;   // team reduction values per wave
constexpr int kSubInts = 10;           // recursion stack entry (a sub-problem)
constexpr int kItemInts = 12;          // shared recursion task: ready flag + sub-problem
constexpr int kQctlInts = 6;           // per tier: {queue, count, head, tail, in-flight task pairs, error}

// Kernel configuration: kLds keeps the whole workspace and both sequences in
// LDS (one wave per pair); otherwise the workspace is a global slab shared by a
// team of kNW waves (kNW > 1: the team is one workgroup).  kW > 0 fixes the
// ring width at compile time (the LDS tier's default width); 0 reads it from
// the layout.  kDefPen fixes the penalties at WFA2's defaults (x=4, o=6, e=2,
// match 0); otherwise they come from the launch parameters.
template <bool L, int NW, int W = 0, bool P = false>
struct Cfg {
  static constexpr bool kLds = L;
  static constexpr int kNW = NW;
  static constexpr int kW = W;
  static constexpr bool kDefPen = P;
};
constexpr int kLdsDefaultWidth = 32;   // LDS-tier ring width unless tuned otherwise
constexpr int kDefX = 4, kDefO = 6, kDefE = 2;   // WFA2's default gap-affine penalties
constexpr int kDefScope = (kDefO + kDefE > kDefX ? kDefO + kDefE : kDefX) + 1;   // ring depth 9
constexpr int kDefGmask = 1;   // every reachable score is even

// Workspace layout of one aligner (a wave, or a team of waves), in int32 units
// from its workspace base.  The ring / arena is shared by the team; every wave
// has a private region (priv + wave * priv_ints) holding its copy of the
// base-case headers, the backtrace buffer and the recursion stack, so that the
// wave-uniform bookkeeping never needs cross-wave synchronisation.
struct WsLayout {
  int data;         // ring data (2 x scope x 3 x wcap)  UNION  base-case arena (offset 0)
  int wcap;         // ring width capacity (power of two)
  int arena_cap;    // ints available at `data` for the base-case arena
  int priv;         // first private region
  int priv_ints;    // ints per private region
  int bhdr;         // (private) base-case headers: hcap scores x 3 comps x {lo, hi, kbase}
  int hcap;         // max base-case score + 1
  int runs_tmp;     // (private) reversed runs of one backtrace (uint32)
  int rcap;
  int stack;        // (private) recursion stack: scap entries x 8 ints
  int scap;
  int seq;          // staged 2-bit sequences (LDS path): P then T, each with pad words
  int seq_words;    // words available for each of P and T (including 3 pad words)
  int total;        // ints per aligner
};

// Kernel parameters (passed by value).
struct KParams {
  // input batch (device)
  const uint64_t* words;
  const int64_t* p_woff;
  const int32_t* p_len;
  const int64_t* t_woff;
  const int32_t* t_len;
  // work source of this tier: pair ids list[0 .. *count), taken through `queue`
  const int32_t* list;
  const int32_t* count;          // device count (the host's initial part plus the
                                 // previous tier's spills)
  int32_t* queue;                // next list index (atomic)
  int32_t chunk;
  // next tier: pairs that outgrow this tier's capacities are appended here
  // (nullptr: last tier, such a pair ends with kStCapacity)
  int32_t* next_list;
  int32_t* next_count;
  // penalties: internal (after the match<0 transform) and user (for the .alg score)
  int32_t x, o, e, scope;
  int32_t um, ux, uo, ue;
  int32_t gmask;                 // g-1 when every penalty is a multiple of g = 2^j, else 0
  int32_t score_only;
  int32_t max_steps;
  // outputs (device, indexed by pair id)
  int32_t* score;
  int32_t* wf_score;
  int32_t* status;
  uint32_t* runs;
  const int64_t* runs_off;       // [N+1] run capacity prefix
  int32_t* nruns;
  // workspace
  int32_t* gws;                  // global workspace (global kernels)
  int64_t gws_stride;            // ints per aligner (wave or team)
  int32_t seq_lds_words;         // team kernels: LDS words per sequence for staging (0: none)
  // Global tiers with score + CIGAR run the BiWFA recursion as shared tasks:
  // a sub-problem split off at a breakpoint may be taken by any aligner of the
  // tier.  Leaves write their runs at position pb+tb of the pair's run buffer;
  // the aligner that finishes a pair's last sub-problem merges them.
  int32_t task;                  // 1: shared-task recursion
  int32_t* items;                // item_cap x kItemInts
  int32_t item_cap;
  int32_t epoch;                 // tier index; the ready-flag value is (*calls << 3) | epoch
  const int32_t* calls;          // call counter, advanced by biwfa_reset (never 0 after it)
  int32_t* qctl;                 // this tier's {queue, count, head, tail, in-flight task pairs, error}
  int32_t* pend;                 // per pair: unfinished sub-problems
  unsigned long long* pfail;     // per pair: leftmost failure, (pb+tb) << 16 | -status
  int32_t* pflag;                // per pair: 1 workspace capacity exceeded, 2 run overflow
  int32_t task_min_length;       // pairs with p + t below this run depth first on one
                                 // aligner (no shared tasks); INT32_MAX: never shared
  int64_t ready_ticks;           // longest wait (100 MHz ticks) for a claimed item's ready flag
  int32_t withhold_ready;        // failure-path test: publish items without their ready flag
  WsLayout L;
};

}  // namespace bw

namespace bw {

// ---------------------------------------------------------------------------
// Gap-affine-2p kernel (biwfa_affine2p.hip): one wave per pair, five
// components (M, I1, D1, I2, D2), rings and base-case arena in a per-wave
// global slab, ring headers in LDS.
constexpr int kA2pMaxScope = 64;        // max(o1+e1, o2+e2, x) + 1 the header table holds
constexpr int kA2pWavesPerBlock = 4;

struct A2pParams {
  // input batch (device)
  const uint64_t* words;
  const int64_t* p_woff;
  const int32_t* p_len;
  const int64_t* t_woff;
  const int32_t* t_len;
  const int32_t* order;          // pair ids in schedule order
  int32_t n;
  int32_t* queue;                // next order index (atomic; zeroed by biwfa2p_reset)
  // penalties: internal (after the match<0 transform) and user (for the .alg score)
  int32_t x, o1, e1, o2, e2, scope;
  int32_t um, ux, uo1, ue1, uo2, ue2;
  int32_t score_only;
  int32_t max_steps;
  // outputs (device, indexed by pair id)
  int32_t* score;
  int32_t* wf_score;
  int32_t* status;
  uint32_t* runs;
  const int64_t* runs_off;       // [N+1] run capacity prefix
  int32_t* nruns;
  // workspace: one slab of `stride` ints per wave
  int32_t* gws;
  int64_t stride;
  int64_t aligners;              // waves with a slab
  int32_t wcap;                  // ring width (power of two)
  int32_t arena_cap;             // ints at the slab base for the rings / base-case arena
  int32_t bhdr;                  // base-case headers: hcap x 5 x {lo, hi, base}
  int32_t hcap;
  int32_t runs_tmp;              // reversed runs of one backtrace
  int32_t rcap;
  int32_t stack;                 // recursion stack: scap x kSubInts
  int32_t scap;
};

}  // namespace bw
