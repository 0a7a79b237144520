// runtime.cpp -- host side of libbiwfa_mi355x: the C ABI declared in
// include/biwfa_mi355x.h on top of the gfx950 kernels in biwfa_kernel.hip.
//
// One aligner = one device + one HIP stream + reusable device buffers.  A batch
// is aligned by two launches of the persistent kernel: the LDS instantiation
// over every pair in schedule order, then the global-workspace instantiation
// over the pairs the first one spilled (its work count is read on the device,
// so no host round trip sits between the launches).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "biwfa_device.hpp"
#include "biwfa_mi355x.h"

namespace bw {
template <class K>
__global__ void biwfa_persistent(KParams p);
__global__ void biwfa_noalloc_reset(int32_t* counters, int n);
__global__ void biwfa_noalloc_admit(KParams p, int64_t n, int64_t num_words, int max_pt, int32_t* list,
                                    int32_t* count);
__global__ void biwfa_reset(int32_t* counters, const int32_t* init, int n);
__global__ void biwfa_compact(const uint32_t* runs, const int64_t* runs_off, const int32_t* nruns,
                              const int64_t* out_off, uint32_t* out, int n);
}  // namespace bw
namespace bw2p {
__global__ void biwfa2p_persistent(bw::A2pParams p);
__global__ void biwfa2p_reset(int32_t* qctl, int n);
}  // namespace bw2p

namespace {

thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                   \
      return BW_STATUS_HIP_ERROR;                                                      \
    }                                                                                  \
  } while (0)

template <class T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * count);
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Workspace layout of one aligner (int32 units): the shared ring / arena,
// then one private region per wave of the team.  Wavefront limits live in
// registers; only the global-workspace kernels keep base-case headers in
// memory (mem_base_hdr), because their base cases may exceed one register lane
// per score.
bw::WsLayout make_layout(int scope, int wcap, int arena_min, int hcap, int rcap, int scap,
                         int seq_words, bool mem_base_hdr, int team_waves) {
  bw::WsLayout L{};
  int off = 0;
  L.data = off;   // the kernels index the offsets from the workspace base
  L.wcap = wcap;
  const int ring_offsets = 2 * scope * 3 * wcap;
  L.arena_cap = std::max(ring_offsets, arena_min);
  off += L.arena_cap;
  off = (off + 15) & ~15;
  L.priv = off;
  int po = 0;
  L.bhdr = po;
  L.hcap = hcap;
  if (mem_base_hdr) po += hcap * 3 * 3;
  L.runs_tmp = po;
  L.rcap = rcap;
  po += rcap;
  L.stack = po;
  L.scap = scap;
  po += scap * bw::kSubInts;
  L.priv_ints = (po + 15) & ~15;
  off += team_waves * L.priv_ints;
  L.seq = off;
  L.seq_words = seq_words;
  off += seq_words > 0 ? 2 * seq_words * 2 : 0;
  L.total = (off + 3) & ~3;
  return L;
}

struct Penalties {
  int x, o, e, scope;     // internal
  int um, ux, uo, ue;     // user
};

int internal_penalties(const bw_affine_penalties_t& p, Penalties* out) {
  // wavefront_penalties_set_affine (wavefront_penalties.c:95-131)
  if (p.match > 0 || p.mismatch <= 0 || p.gap_opening < 0 || p.gap_extension <= 0) {
    set_error("penalties must satisfy match<=0, mismatch>0, gap_opening>=0, gap_extension>0");
    return BW_STATUS_INVALID_ARG;
  }
  Penalties q;
  q.um = p.match; q.ux = p.mismatch; q.uo = p.gap_opening; q.ue = p.gap_extension;
  if (p.match < 0) {
    q.x = 2 * p.mismatch - 2 * p.match;
    q.o = 2 * p.gap_opening;
    q.e = 2 * p.gap_extension - p.match;
  } else {
    q.x = p.mismatch; q.o = p.gap_opening; q.e = p.gap_extension;
  }
  // wavefront_components_dimensions_affine (wavefront_components.c:81-100)
  q.scope = std::max(q.o + q.e, q.x) + 1;
  if (q.scope > bw::kMaxScope) {
    set_error("max(o+e, x)+1 exceeds the supported ring depth (32)");
    return BW_STATUS_INVALID_ARG;
  }
  *out = q;
  return 0;
}

// Gap-affine-2p penalties, internal (after the match<0 transform) and user.
struct Pen2p {
  int x, o1, e1, o2, e2, scope;
  int um, ux, uo1, ue1, uo2, ue2;
};

int internal_penalties_2p(const bw_affine2p_penalties_t& p, Pen2p* out) {
  // wavefront_penalties_set_affine2p (wavefront_penalties.c:132-173)
  if (p.match > 0 || p.mismatch <= 0 || p.gap_opening1 < 0 || p.gap_extension1 <= 0 || p.gap_opening2 < 0 ||
      p.gap_extension2 <= 0) {
    set_error("gap-affine-2p penalties must satisfy match<=0, mismatch>0, gap_opening1/2>=0, gap_extension1/2>0");
    return BW_STATUS_INVALID_ARG;
  }
  Pen2p q;
  q.um = p.match; q.ux = p.mismatch;
  q.uo1 = p.gap_opening1; q.ue1 = p.gap_extension1; q.uo2 = p.gap_opening2; q.ue2 = p.gap_extension2;
  if (p.match < 0) {
    q.x = 2 * p.mismatch - 2 * p.match;
    q.o1 = 2 * p.gap_opening1;
    q.e1 = 2 * p.gap_extension1 - p.match;
    q.o2 = 2 * p.gap_opening2;
    q.e2 = 2 * p.gap_extension2 - p.match;
  } else {
    q.x = p.mismatch; q.o1 = p.gap_opening1; q.e1 = p.gap_extension1; q.o2 = p.gap_opening2; q.e2 = p.gap_extension2;
  }
  // wavefront_components_dimensions_affine2p (wavefront_components.c:101-124)
  q.scope = std::max(std::max(q.o1 + q.e1, q.o2 + q.e2), q.x) + 1;
  if (q.scope > bw::kA2pMaxScope) {
    set_error("max(o1+e1, o2+e2, x)+1 exceeds the supported ring depth (64)");
    return BW_STATUS_INVALID_ARG;
  }
  *out = q;
  return 0;
}

}  // namespace

// The other host sources (sequence files) report through the same text.
namespace bw_host {
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace bw_host

// One launch of the persistent kernel over a list of pairs with fixed capacities.
// Tier 0 keeps its workspace in LDS; later tiers use per-wave global slabs whose
// ring width grows geometrically.  A pair that outgrows its tier (ring width,
// base-case arena, headers, recursion stack or LDS staging) is appended to the
// next tier's list on the device; the last tier holds the batch's worst case.
struct Tier {
  bool lds = false;
  bw::WsLayout L{};
  int nw = 1;                // waves per aligner (team size)
  int waves_per_block = bw::kWavesPerBlock;
  int blocks = 0;
  int64_t aligners = 0;      // global tiers: workspace slabs in use
  int64_t stride = 0;        // global tiers: ints per aligner
  int chunk = 1;
  int seq_cap = 0;           // LDS tier: longest sequence it can stage
  int seq_lds_words = 0;     // team tiers: LDS words per staged sequence
  size_t dyn_lds = 0;        // dynamic LDS bytes per workgroup
};

// Kernel instantiation of a global tier with `nw` waves per aligner.
static const void* team_kernel(int nw) {
  switch (nw) {
    case 1: return reinterpret_cast<const void*>(&bw::biwfa_persistent<bw::Cfg<false, 1>>);
    case 2: return reinterpret_cast<const void*>(&bw::biwfa_persistent<bw::Cfg<false, 2>>);
    case 4: return reinterpret_cast<const void*>(&bw::biwfa_persistent<bw::Cfg<false, 4>>);
    case 8: return reinterpret_cast<const void*>(&bw::biwfa_persistent<bw::Cfg<false, 8>>);
    default: return reinterpret_cast<const void*>(&bw::biwfa_persistent<bw::Cfg<false, 16>>);
  }
}

struct bw_dbatch {
  int64_t n = 0, nwords = 0;
  uint64_t* d_words = nullptr;
  int64_t* d_pwoff = nullptr;
  int32_t* d_plen = nullptr;
  int64_t* d_twoff = nullptr;
  int32_t* d_tlen = nullptr;
  std::vector<int32_t> h_plen, h_tlen;
  int max_p = 0, max_t = 0, max_pt = 0;
  std::vector<Tier> tiers;
  int32_t* d_lists = nullptr;   // tiers x n pair ids
  int32_t* d_init = nullptr;    // [kQctlInts*kMaxTiers] counter initial values
  int32_t* d_score = nullptr;
  int32_t* d_wf = nullptr;
  int32_t* d_status = nullptr;
  int32_t* d_nruns = nullptr;
  uint32_t* d_runs = nullptr;
  int64_t* d_runs_off = nullptr;
  int64_t runs_cap = 0;
  int32_t* d_gws = nullptr;
  // shared recursion tasks of the global tiers
  int32_t* d_items = nullptr;   // item_cap x kItemInts (ready flags start at 0)
  int32_t item_cap = 0;
  int32_t* d_pend = nullptr;
  unsigned long long* d_pfail = nullptr;
  int32_t* d_pflag = nullptr;
  int64_t gws_bytes = 0;
  std::vector<int32_t> tier_count;   // initial pairs per tier (last schedule)
  // The aligner that uploaded the batch: its tiers were planned for that
  // aligner's penalties, tuning and device, so only it may align, schedule or
  // download the batch (bw_dbatch_align / bw_set_schedule / bw_dbatch_download).
  uint64_t owner = 0;
  // graph mode: the owner's launch sequence, captured once and replayed
  hipGraphExec_t gexec = nullptr;
  // gap-affine-2p batch: one launch of the single-wave 2p kernel over `d_order`
  bool a2p = false;
  int32_t* d_order = nullptr;
  bw::A2pParams p2{};
  int blocks2p = 0;
};

constexpr int kMaxTiers = 8;
static void drop_one(bw_aligner_t* a);
// Longest wait for a claimed shared task's ready flag: 60 s of the 100 MHz clock.
constexpr int64_t kReadyTicksDefault = 60ll * 100000000ll;
// Pairs with p + t below this run depth first on one aligner of a global tier.
constexpr int kTaskMinLength = 4096;
std::atomic<uint64_t> g_next_aligner_id{1};

struct bw_aligner {
  uint64_t id = 0;                 // unique over the process (device-batch ownership)
  bw_attr_t attr{};
  Penalties pen{};
  bool a2p = false;                // gap-affine-2p aligner (its own kernel; no tiers)
  Pen2p p2{};
  int dev = 0, cu_count = 0;
  size_t lds_per_block_max = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t tev[kMaxTiers + 1] = {};   // tier boundaries of the last launch sequence
  int last_ntiers = 0;
  float tier_ms[kMaxTiers] = {};
  int32_t tier_pairs[kMaxTiers] = {};   // pairs each tier took on (initial + spilled in)
  bw_tier_shape_t tier_shape[kMaxTiers] = {};   // launch shapes of the last sequence
  bw_tuning_t tuning{};
  bw_progress_fn progress = nullptr;
  void* progress_user = nullptr;
  int32_t* d_counters = nullptr;   // per tier: {queue, count, head, tail, in-flight task pairs, error},
                                   // then the call counter the ready-flag epochs derive from
  const int32_t* last_counters = nullptr;   // counters of the last launch sequence (d_counters or scratch)
  int64_t ready_ticks = kReadyTicksDefault; // bound of a shared-task ready-flag wait (100 MHz ticks)
  int withhold_ready = 0;                   // failure-path test hook (bw_test_set_ready_guard)
  int32_t* scratch = nullptr;               // this device's no-allocation scratch (bw_align_device_noalloc)
  int64_t ws_budget = 0;                    // bytes for the global tiers' workspace slabs
  bw_dbatch_t* one = nullptr;               // single-pair calls: reused device batch (bw_align)
  int graph = 0;                   // replay a captured launch sequence (bw_set_graph)
  bool last_graph = false;         // the last sequence was a graph replay (no per-tier events)
  float last_total_ms = 0.f, last_main_ms = 0.f, last_spill_ms = 0.f;
  // last single-pair result
  int status = 0, score = 0, wf_score = 0;
  std::vector<uint32_t> runs;
  std::vector<char> ops;
  std::vector<uint32_t> sam;
  int schedule = BW_SCHEDULE_LENGTH_SORTED;
};

// ---------------- scratch of the no-allocation entry point ----------------
// bw_align_device_noalloc keeps everything it needs beyond the caller's inputs
// and outputs in this array: the counters of its two tiers, the LDS tier's pair
// list, the spill tier's pair list and the spill tier's workspace slabs.  Each
// GPU holds its own copy for the life of the process; every call on a GPU
// reuses it (calls that overlap on one GPU race).
constexpr int64_t kNaMaxPairs = 1 << 20;          // pairs per call
constexpr int kNaMaxPT = 2040;                    // p + t of the longest pair it aligns
constexpr int64_t kNaScratchBytes = 64ll << 20;   // 64 MiB per GPU
constexpr int64_t kNaCounterInts = 64;            // 2 tiers x kQctlInts + call counter, padded
static_assert(kNaCounterInts >= 2 * bw::kQctlInts + 1, "no-allocation counter area too small");
static __device__ __align__(256) int32_t scratch[kNaScratchBytes / 4];

// Device address of `scratch` on each device (hipGetSymbolAddress with the
// device current, once per device; the host-side name is not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_ptr[kMaxDevices] = {};

static hipError_t scratch_for_device(int dev, int32_t** out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<int32_t*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

extern "C" {

const char* bw_last_error(void) { return g_last_error.c_str(); }

const char* bw_strerror(int status) {
  switch (status) {
    case BW_STATUS_ALG_COMPLETED: return "[WFA] Alignment completed successfully";
    case BW_STATUS_ALG_PARTIAL: return "[WFA] Alignment extension computed (partial alignment)";
    case BW_STATUS_MAX_STEPS_REACHED: return "[WFA] Alignment failed. Maximum WFA-steps limit reached";
    case BW_STATUS_OOM: return "[WFA] Alignment failed. Maximum memory limit reached";
    case BW_STATUS_UNATTAINABLE: return "[WFA] Alignment failed. Unattainable under configured heuristics";
    case BW_STATUS_DEVICE_CAPACITY: return "[BiWFA-MI355X] A per-pair device capacity was exceeded";
    case BW_STATUS_HIP_ERROR: return "[BiWFA-MI355X] HIP runtime error";
    case BW_STATUS_INVALID_ARG: return "[BiWFA-MI355X] Invalid argument";
    case BW_STATUS_NO_DEVICE: return "[BiWFA-MI355X] No usable gfx950 device";
    case BW_STATUS_BAD_SEQUENCE: return "[BiWFA-MI355X] Sequence byte outside ACGT";
    case BW_STATUS_INTERNAL: return "[BiWFA-MI355X] Device termination guard tripped";
    default: return "[WFA] Unknown error code";
  }
}

void bw_attr_default(bw_attr_t* attr) {
  std::memset(attr, 0, sizeof(*attr));
  attr->penalties = bw_affine_penalties_t{0, 4, 6, 2};
  attr->scope = BW_SCOPE_ALIGNMENT;
  attr->max_alignment_steps = 0;
  attr->device = 0;
  attr->schedule = BW_SCHEDULE_LENGTH_SORTED;
  attr->waves_per_cu = 0;
  attr->queue_chunk = 0;
  attr->distance_metric = BW_GAP_AFFINE;
  // align_benchmark's gap-affine-2p defaults (tools/align_benchmark/align_benchmark_params.c:62-69)
  attr->affine2p_penalties = bw_affine2p_penalties_t{0, 4, 6, 2, 24, 1};
}

bw_aligner_t* bw_aligner_new(const bw_attr_t* attr_in) {
  bw_attr_t attr;
  if (attr_in) attr = *attr_in; else bw_attr_default(&attr);
  Penalties pen;
  if (internal_penalties(attr.penalties, &pen) != 0) return nullptr;
  Pen2p p2{};
  if (attr.distance_metric != BW_GAP_AFFINE && attr.distance_metric != BW_GAP_AFFINE_2P) {
    set_error("unknown distance metric");
    return nullptr;
  }
  if (attr.distance_metric == BW_GAP_AFFINE_2P && internal_penalties_2p(attr.affine2p_penalties, &p2) != 0)
    return nullptr;
  if (attr.schedule < BW_SCHEDULE_INPUT_ORDER || attr.schedule > BW_SCHEDULE_LEARNED) {
    set_error("unknown schedule");
    return nullptr;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    set_error("no HIP device available");
    return nullptr;
  }
  if (attr.device < 0 || attr.device >= ndev) {
    set_error("device ordinal out of range");
    return nullptr;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, attr.device) != hipSuccess) {
    set_error("hipGetDeviceProperties failed");
    return nullptr;
  }
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
    return nullptr;
  }
  bw_aligner* a = new bw_aligner();
  a->id = g_next_aligner_id.fetch_add(1);
  a->attr = attr;
  a->pen = pen;
  a->a2p = attr.distance_metric == BW_GAP_AFFINE_2P;
  a->p2 = p2;
  a->dev = attr.device;
  a->cu_count = prop.multiProcessorCount;
  // workspace slabs of the global tiers: at most a third of the device memory, 64 GiB
  a->ws_budget = std::min<int64_t>((int64_t)64 << 30, (int64_t)(prop.totalGlobalMem / 3));
  a->lds_per_block_max = prop.sharedMemPerBlock;
  a->schedule = attr.schedule;
  if (hipSetDevice(a->dev) != hipSuccess || hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking) != hipSuccess ||
      dalloc(&a->d_counters, bw::kQctlInts * kMaxTiers + 1) != hipSuccess ||
      hipMemset(a->d_counters, 0, sizeof(int32_t) * (bw::kQctlInts * kMaxTiers + 1)) != hipSuccess) {
    set_error("stream / counter allocation failed");
    delete a;
    return nullptr;
  }
  {
    const hipError_t e = scratch_for_device(a->dev, &a->scratch);
    if (e != hipSuccess) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      (void)hipStreamDestroy(a->stream);
      (void)hipFree(a->d_counters);
      delete a;
      return nullptr;
    }
  }
  for (auto& e : a->ev) (void)hipEventCreate(&e);
  for (auto& e : a->tev) (void)hipEventCreate(&e);
  return a;
}

void bw_aligner_delete(bw_aligner_t* a) {
  if (!a) return;
  (void)hipSetDevice(a->dev);
  if (a->stream) (void)hipStreamSynchronize(a->stream);
  bw_dbatch_free(a->one);
  dfree(a->d_counters);
  for (auto& e : a->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : a->tev)
    if (e) (void)hipEventDestroy(e);
  if (a->stream) (void)hipStreamDestroy(a->stream);
  delete a;
}

void bw_dbatch_free(bw_dbatch_t* b) {
  if (!b) return;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  dfree(b->d_words); dfree(b->d_pwoff); dfree(b->d_plen); dfree(b->d_twoff); dfree(b->d_tlen);
  dfree(b->d_lists); dfree(b->d_init); dfree(b->d_score); dfree(b->d_wf); dfree(b->d_status);
  dfree(b->d_nruns); dfree(b->d_runs); dfree(b->d_runs_off); dfree(b->d_gws);
  dfree(b->d_items); dfree(b->d_pend); dfree(b->d_pfail); dfree(b->d_pflag);
  dfree(b->d_order);
  delete b;
}

}  // extern "C"

// Largest power of two dividing all of x, o, e (every reachable score is a
// multiple of it), as a mask for the kernel's null-step shortcut.
static int null_score_mask(const Penalties& q) {
  int g = q.x | q.o | q.e;
  g &= -g;                 // lowest set bit = largest power of two dividing all three
  return g > 1 ? g - 1 : 0;
}

// WFA2's default penalties (x=4, o=6, e=2), which the LDS kernel has a
// compile-time variant for.
static bool default_penalties(const Penalties& q) {
  return q.x == bw::kDefX && q.o == bw::kDefO && q.e == bw::kDefE && q.scope == bw::kDefScope &&
         null_score_mask(q) == bw::kDefGmask;
}

// LDS-tier kernel instantiation: ring width 32 and the default penalties as
// compile-time constants where they apply (fewer scalar registers live across
// the recursion), else the runtime-parameter kernel.
static const void* lds_kernel_for(int wcap, bool defpen) {
  if (wcap == bw::kLdsDefaultWidth && defpen)
    return reinterpret_cast<const void*>(&bw::biwfa_persistent<bw::Cfg<true, 1, bw::kLdsDefaultWidth, true>>);
  if (wcap == bw::kLdsDefaultWidth)
    return reinterpret_cast<const void*>(&bw::biwfa_persistent<bw::Cfg<true, 1, bw::kLdsDefaultWidth>>);
  return reinterpret_cast<const void*>(&bw::biwfa_persistent<bw::Cfg<true, 1>>);
}

// Upper bound of a base-case score: 250 below a breakpoint, 500 on the
// end-reached fallback, and the max(p,t) <= 100 whole-pair case
// (min(p,t)*x + o + |p-t|*e, the bound WFA2 sizes its wavefront table with,
// wavefront_components.c:81-100).
static int base_score_bound(const Penalties& q) {
  return std::max(bw::kRecoveryMinScore, 100 * q.x + q.o + 100 * q.e) + 1;
}

// Waves per aligner of the global tier with ring width `wcap` (the built-in
// rule, from bw_autotune runs on the mixed-length, long-pair and 1 kbp batches:
// a step is bound by the latency of its HBM round trips, and with enough pairs
// to fill the GPU several small teams per CU hide it better than one wide team).
static int team_waves(const bw_aligner_t* a, int wcap) {
  (void)a;
  if (wcap <= 1024) return 1;
  return 4;
}

// Tier 0: the whole workspace and both sequences in LDS, one wave per pair,
// ring width 32 and sequences up to 512 bases staged unless tuned otherwise.
static Tier make_lds_tier(const bw_aligner_t* a) {
  const Penalties& q = a->pen;
  const bw_tuning_t& tu = a->tuning;
  Tier t;
  t.lds = true;
  t.seq_cap = tu.lds_seq_cap > 0 ? tu.lds_seq_cap : 512;
  const int seq_words = (t.seq_cap + 31) / 32 + 3;
  const int lds_wcap = tu.lds_ring_width > 0 ? tu.lds_ring_width : bw::kLdsDefaultWidth;
  t.L = make_layout(q.scope, lds_wcap, 0, bw::kLdsHcap, bw::kLdsRcap, bw::kLdsScap, seq_words, false, 1);
  t.waves_per_block = bw::kWavesPerBlock;
  t.dyn_lds = (size_t)t.L.total * 4 * t.waves_per_block;
  int nb = 0;
  const void* lds_kernel = lds_kernel_for(lds_wcap, default_penalties(q));
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lds_kernel, bw::kWave * t.waves_per_block, t.dyn_lds) !=
          hipSuccess ||
      nb <= 0)
    nb = std::max<int>(1, (int)((160 * 1024) / t.dyn_lds));
  if (a->attr.waves_per_cu > 0) nb = std::max(1, std::min(nb, a->attr.waves_per_cu / t.waves_per_block));
  t.blocks = a->cu_count * nb;
  // 2 pairs per queue grab: on 1,000,000 short reads 0.4 % faster than 4 (fewer
  // grabs than 1, a shorter tail than 4 or more)
  t.chunk = tu.queue_chunk[0] > 0 ? tu.queue_chunk[0] : a->attr.queue_chunk > 0 ? a->attr.queue_chunk : 2;
  return t;
}

// Global tier gi (ring width wcap; `last`: the widest tier, whose base-case
// arena holds the worst case of the batch) for pairs of p + t <= max_pt and
// sequences of at most max_seq_words words (with pads).  `budget` bounds the
// bytes of all its workspace slabs together.  nw > 0 fixes the team size.
static Tier make_global_tier(const bw_aligner_t* a, int gi, int wcap, bool last, int max_pt, int max_seq_words,
                             int64_t budget, int nw = 0) {
  const Penalties& q = a->pen;
  const int scope = q.scope;
  const int smax = base_score_bound(q);
  const int g = null_score_mask(q) + 1;
  const bw_tuning_t& tu = a->tuning;
  Tier t;
  t.lds = false;
  t.nw = nw > 0 ? nw : tu.team_waves[gi] > 0 ? tu.team_waves[gi] : team_waves(a, wcap);
  // base-case arena: every score's three components, each padded to whole
  // groups of 4 offsets (up to 6 more) from a multiple-of-4 index (up to 3)
  const int width = std::min(2 * smax / q.e + 3, max_pt + 3);
  const int64_t worst = (int64_t)3 * ((smax + g) / g) * (width + 9);
  const int64_t arena =
      last ? worst : std::min<int64_t>(worst, (int64_t)3 * ((smax + g) / g) * (std::min(wcap, width) + 9));
  const int steps = std::min(max_pt, smax / std::max(1, std::min(q.x, q.e)) + 1);
  t.L = make_layout(scope, wcap, (int)std::min<int64_t>(arena, (int64_t)1 << 30), smax + 2,
                    2 * steps + 16, 64, 0, true, t.nw);
  t.stride = t.L.total;
  const int64_t per_aligner = t.stride * 4;
  const void* kern = team_kernel(t.nw);
  // Pairs are staged in LDS when the batch's longest sequence fits the
  // aligner's share of the CU's LDS (the kernels run 4 waves per SIMD, i.e.
  // 16 waves per CU).
  const int64_t lds_cu = 160 * 1024;
  if (t.nw == 1) {
    t.waves_per_block = bw::kWavesPerBlock;
    t.seq_lds_words = (max_seq_words * 16 <= lds_cu / 16) ? max_seq_words : 0;
    t.dyn_lds = (size_t)t.seq_lds_words * 16 * bw::kWavesPerBlock;
    const int64_t waves = std::max<int64_t>(bw::kWavesPerBlock,
                                            std::min<int64_t>(budget / per_aligner, (int64_t)a->cu_count * 16));
    t.blocks = (int)std::max<int64_t>(1, waves / bw::kWavesPerBlock);
    t.aligners = (int64_t)t.blocks * bw::kWavesPerBlock;
  } else {
    // one team per workgroup
    t.waves_per_block = t.nw;
    const int teams_cu = std::max(1, 16 / t.nw);
    const int64_t red_bytes = (int64_t)2 * t.nw * bw::kRedSlots * 4;
    const int64_t share = lds_cu / teams_cu - red_bytes;
    t.seq_lds_words = (max_seq_words * 16 <= share) ? max_seq_words : 0;
    t.dyn_lds = (size_t)t.seq_lds_words * 16;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, bw::kWave * t.nw, t.dyn_lds) != hipSuccess ||
        nb <= 0)
      nb = 1;
    const int64_t teams = std::max<int64_t>(1, std::min<int64_t>(budget / per_aligner, (int64_t)a->cu_count * nb));
    t.blocks = (int)teams;
    t.aligners = teams;
  }
  t.chunk = tu.queue_chunk[gi + 1] > 0 ? tu.queue_chunk[gi + 1] : 1;
  return t;
}

// The LDS kernel takes these layout values as compile-time constants.
static int check_lds_layout(const bw_aligner_t* a, const Tier& t0) {
  const bw::WsLayout& L0 = t0.L;
  const bool fixed = L0.wcap == bw::kLdsDefaultWidth && default_penalties(a->pen);   // compiled-in variant
  if (L0.data != 0 || L0.hcap != bw::kLdsHcap || L0.rcap != bw::kLdsRcap || L0.scap != bw::kLdsScap ||
      L0.runs_tmp != 0 || L0.stack != bw::kLdsRcap ||
      (fixed && L0.arena_cap != 2 * bw::kDefScope * 3 * bw::kLdsDefaultWidth)) {
    set_error("LDS-tier layout does not match the kernel's constants");
    return BW_STATUS_INTERNAL;
  }
  // a workgroup of the LDS tier holds kWavesPerBlock waves' workspaces
  if (t0.dyn_lds > (size_t)160 * 1024) {
    set_error("LDS tier: ring width " + std::to_string(L0.wcap) + " with sequences staged up to " +
              std::to_string(t0.seq_cap) + " bases needs " + std::to_string(t0.dyn_lds / 1024) +
              " KiB per workgroup, more than a CU's 160 KiB (tune a narrower ring or a shorter lds_seq_cap)");
    return BW_STATUS_INVALID_ARG;
  }
  return 0;
}

static std::vector<Tier> plan_tiers(const bw_aligner_t* a, const bw_dbatch_t* b, int64_t budget) {
  std::vector<Tier> tiers;
  tiers.push_back(make_lds_tier(a));
  // Global tiers: single-wave aligners with rings 256 and 1024 wide, then one
  // team tier as wide as the batch needs.  Team tiers of graded widths would
  // only save workspace memory (every team's slab holds its tier's ring), at
  // the price of a tail at every tier boundary and re-runs of the pairs that
  // outgrow a tier; one slab of the widest ring per team fits the budget.
  const int wmax = next_pow2(std::max(b->max_pt + 8, 256));
  const int max_seq_words = (std::max(b->max_p, b->max_t) + 31) / 32 + 3;
  for (int w = 256;; w *= 4) {
    const bool last = w >= wmax || w > 1024 || (int)tiers.size() + 1 >= kMaxTiers;
    const int wcap = last ? wmax : w;
    tiers.push_back(make_global_tier(a, (int)tiers.size() - 1, wcap, last, b->max_pt, max_seq_words, budget));
    if (last || (int)tiers.size() >= kMaxTiers) break;
  }
  return tiers;
}

// Initial tier of a pair: tier 0 when it can be staged in LDS, else the first
// global tier whose ring is at least `width_guess` wide.
static int initial_tier(const bw_dbatch_t* b, int pair, double width_guess) {
  const int p = b->h_plen[pair], t = b->h_tlen[pair];
  const auto& T = b->tiers;
  if (std::max(p, t) <= T[0].seq_cap && width_guess <= T[0].L.wcap) return 0;
  for (size_t i = 1; i < T.size(); ++i)
    if (T[i].L.wcap >= width_guess || i + 1 == T.size()) return (int)i;
  return (int)T.size() - 1;
}

// Gap-affine-2p batch: inputs, outputs and one workspace slab per wave of the
// single-wave 2p kernel.  A slab holds the two rings (2 x scope x 5 x wcap
// offsets, wcap a power of two >= p + t + 1 of the longest pair) or, in the
// same space, a base case's arena, then the base-case headers, the backtrace
// run buffer and the recursion stack.  The base-case capacities come from the
// largest score a base case can reach: 250 below a breakpoint, 500 on the
// end-reached fallback, the max(p, t) <= 100 whole-pair case (100 mismatches
// plus a 100-base gap), with slack.
static int upload_2p(bw_aligner_t* a, const bw_batch_t* in, bw_dbatch_t* b, const std::vector<int64_t>& runs_off,
                     int64_t budget, bw_dbatch_t** out) {
  const Pen2p& q = a->p2;
  const int64_t n = b->n;
  auto fail = [&](hipError_t e) {
    set_error(std::string("device allocation/copy failed: ") + hipGetErrorString(e));
    bw_dbatch_free(b);
    return BW_STATUS_HIP_ERROR;
  };
  b->a2p = true;
  const int wcap = next_pow2(std::max(b->max_pt + 1, 64));
  const int smax = std::max(bw::kRecoveryMinScore, 100 * q.x + std::min(q.o1 + 100 * q.e1, q.o2 + 100 * q.e2)) +
                   2 * q.scope + 64;
  // a base case's arena: five components per score, the score-s wavefront at
  // most 2s/e_min + 3 diagonals wide (its limits grow by one per gap extension)
  // and never wider than p + t + 3
  const int emin = std::max(1, std::min(q.e1, q.e2));
  int64_t arena_need = 0;
  for (int sc = 0; sc <= smax; ++sc)
    arena_need += 5 * std::min<int64_t>(2 * (int64_t)sc / emin + 3, (int64_t)b->max_pt + 3);
  const int64_t ring_ints = (int64_t)2 * q.scope * 5 * wcap;
  const int64_t arena = std::max<int64_t>(ring_ints, arena_need);
  if (arena > ((int64_t)1 << 30)) {
    set_error("gap-affine-2p workspace of one pair exceeds 4 GiB");
    bw_dbatch_free(b);
    return BW_STATUS_INVALID_ARG;
  }
  bw::A2pParams& p = b->p2;
  p = bw::A2pParams{};
  p.wcap = wcap;
  p.arena_cap = (int)arena;
  int64_t off = (arena + 63) & ~(int64_t)63;
  p.bhdr = (int)off;
  p.hcap = smax + 2;
  off += (int64_t)p.hcap * 5 * 3;
  p.runs_tmp = (int)off;
  p.rcap = 2 * smax + 16;
  off += p.rcap;
  p.stack = (int)off;
  p.scap = 64;
  off += (int64_t)p.scap * bw::kSubInts;
  p.stride = (off + 63) & ~(int64_t)63;
  const int64_t slab = p.stride * 4;
  const int64_t fit = budget / slab;
  if (fit < 1) {
    set_error("gap-affine-2p workspace of one pair exceeds the workspace budget");
    bw_dbatch_free(b);
    return BW_STATUS_DEVICE_CAPACITY;
  }
  // as many waves as the CUs hold at once (the kernel's occupancy), the pairs
  // need and the workspace budget allows
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&bw2p::biwfa2p_persistent),
                                                   bw::kWave * bw::kA2pWavesPerBlock, 0) != hipSuccess || per_cu <= 0)
    per_cu = 2;
  const int64_t resident = (int64_t)a->cu_count * per_cu * bw::kA2pWavesPerBlock;
  p.aligners = std::max<int64_t>(1, std::min<int64_t>({std::max<int64_t>(n, 1), resident, fit}));
  b->blocks2p = (int)((p.aligners + bw::kA2pWavesPerBlock - 1) / bw::kA2pWavesPerBlock);
  hipError_t e;
  if ((e = dalloc(&b->d_words, in->num_words)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_pwoff, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_plen, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_twoff, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_tlen, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_order, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_score, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_wf, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_status, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_nruns, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_runs, b->runs_cap)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_runs_off, n + 1)) != hipSuccess) return fail(e);
  b->gws_bytes = p.aligners * slab;
  if ((e = hipMalloc(reinterpret_cast<void**>(&b->d_gws), b->gws_bytes)) != hipSuccess) return fail(e);
  hipStream_t s = a->stream;
  if ((e = hipMemcpyAsync(b->d_words, in->words, sizeof(uint64_t) * in->num_words, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
  if (n > 0) {
    if ((e = hipMemcpyAsync(b->d_pwoff, in->pattern_word_offset, sizeof(int64_t) * n, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(b->d_plen, in->pattern_length, sizeof(int32_t) * n, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(b->d_twoff, in->text_word_offset, sizeof(int64_t) * n, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(b->d_tlen, in->text_length, sizeof(int32_t) * n, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
  }
  if ((e = hipMemcpyAsync(b->d_runs_off, runs_off.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
  p.words = b->d_words;
  p.p_woff = b->d_pwoff; p.p_len = b->d_plen;
  p.t_woff = b->d_twoff; p.t_len = b->d_tlen;
  p.order = b->d_order;
  p.n = (int32_t)n;
  p.queue = a->d_counters;
  p.x = q.x; p.o1 = q.o1; p.e1 = q.e1; p.o2 = q.o2; p.e2 = q.e2; p.scope = q.scope;
  p.um = q.um; p.ux = q.ux; p.uo1 = q.uo1; p.ue1 = q.ue1; p.uo2 = q.uo2; p.ue2 = q.ue2;
  p.score_only = a->attr.scope == BW_SCOPE_SCORE;
  p.max_steps = a->attr.max_alignment_steps > 0 ? a->attr.max_alignment_steps : INT32_MAX;
  p.score = b->d_score; p.wf_score = b->d_wf; p.status = b->d_status;
  p.runs = b->d_runs; p.runs_off = b->d_runs_off; p.nruns = b->d_nruns;
  p.gws = b->d_gws;
  *out = b;
  const int sched = a->schedule == BW_SCHEDULE_LEARNED ? BW_SCHEDULE_LENGTH_SORTED : a->schedule;
  const int st = bw_set_schedule(a, b, sched, nullptr);
  if (st != 0) {
    bw_dbatch_free(b);
    *out = nullptr;
    return st;
  }
  return 0;
}

extern "C" {

static int upload_impl(bw_aligner_t* a, const bw_batch_t* in, bw_dbatch_t** out, int64_t budget = 0) {
  *out = nullptr;
  if (!a || !in || in->num_pairs < 0 || (in->num_pairs > 0 && (!in->words || !in->pattern_length ||
      !in->text_length || !in->pattern_word_offset || !in->text_word_offset))) {
    set_error("invalid batch");
    return BW_STATUS_INVALID_ARG;
  }
  if (in->num_pairs > INT32_MAX / kMaxTiers) {
    set_error("too many pairs in one batch");
    return BW_STATUS_INVALID_ARG;
  }
  HIP_TRY(hipSetDevice(a->dev));
  bw_dbatch* b = new bw_dbatch();
  const int64_t n = in->num_pairs;
  b->n = n;
  b->nwords = in->num_words;
  b->h_plen.assign(in->pattern_length, in->pattern_length + n);
  b->h_tlen.assign(in->text_length, in->text_length + n);
  std::vector<int64_t> runs_off(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int p = b->h_plen[i], t = b->h_tlen[i];
    const int64_t pw = in->pattern_word_offset[i], tw = in->text_word_offset[i];
    if (p < 0 || t < 0 || (int64_t)p + t > (1 << 28) || pw < 1 || tw < 1 ||
        pw + (p + 31) / 32 + 2 > in->num_words || tw + (t + 31) / 32 + 2 > in->num_words) {
      set_error("pair " + std::to_string(i) + ": bad length or word offset (need one pad word "
                "before and two after every sequence)");
      bw_dbatch_free(b);
      return BW_STATUS_INVALID_ARG;
    }
    b->max_p = std::max(b->max_p, p);
    b->max_t = std::max(b->max_t, t);
    b->max_pt = std::max(b->max_pt, p + t);
    runs_off[i + 1] = runs_off[i] + (int64_t)p + t + 1;
  }
  b->runs_cap = runs_off[n];
  b->owner = a->id;
  if (a->a2p) return upload_2p(a, in, b, runs_off, budget > 0 ? std::min(budget, a->ws_budget) : a->ws_budget, out);
  b->tiers = plan_tiers(a, b, budget > 0 ? std::min(budget, a->ws_budget) : a->ws_budget);
  if (const int st = check_lds_layout(a, b->tiers[0])) {
    bw_dbatch_free(b);
    return st;
  }
  auto fail = [&](hipError_t e) {
    set_error(std::string("device allocation/copy failed: ") + hipGetErrorString(e));
    bw_dbatch_free(b);
    return BW_STATUS_HIP_ERROR;
  };
  hipError_t e;
  const size_t nt = b->tiers.size();
  if ((e = dalloc(&b->d_words, in->num_words)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_pwoff, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_plen, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_twoff, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_tlen, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_lists, nt * (size_t)std::max<int64_t>(n, 1))) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_init, bw::kQctlInts * kMaxTiers)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_pend, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_pfail, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_pflag, n)) != hipSuccess) return fail(e);
  b->item_cap = 1 << 20;
  if ((e = dalloc(&b->d_items, (size_t)b->item_cap * bw::kItemInts)) != hipSuccess) return fail(e);
  if ((e = hipMemsetAsync(b->d_items, 0, sizeof(int32_t) * (size_t)b->item_cap * bw::kItemInts, a->stream)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_score, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_wf, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_status, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_nruns, n)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_runs, b->runs_cap)) != hipSuccess) return fail(e);
  if ((e = dalloc(&b->d_runs_off, n + 1)) != hipSuccess) return fail(e);
  int64_t gbytes = 0;
  for (const Tier& t : b->tiers)
    if (!t.lds) gbytes = std::max<int64_t>(gbytes, t.aligners * t.stride * 4);
  if ((e = hipMalloc(reinterpret_cast<void**>(&b->d_gws), std::max<int64_t>(gbytes, 64))) != hipSuccess) return fail(e);
  b->gws_bytes = gbytes;
  hipStream_t s = a->stream;
  if ((e = hipMemcpyAsync(b->d_words, in->words, sizeof(uint64_t) * in->num_words, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
  if (n > 0) {
    if ((e = hipMemcpyAsync(b->d_pwoff, in->pattern_word_offset, sizeof(int64_t) * n, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(b->d_plen, in->pattern_length, sizeof(int32_t) * n, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(b->d_twoff, in->text_word_offset, sizeof(int64_t) * n, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(b->d_tlen, in->text_length, sizeof(int32_t) * n, hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
  }
  if ((e = hipMemcpyAsync(b->d_runs_off, runs_off.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, s)) != hipSuccess) return fail(e);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
  *out = b;
  const int sched = a->schedule == BW_SCHEDULE_LEARNED ? BW_SCHEDULE_LENGTH_SORTED : a->schedule;
  const int st = bw_set_schedule(a, b, sched, nullptr);
  if (st != 0) {
    bw_dbatch_free(b);
    *out = nullptr;
    return st;
  }
  return 0;
}

bw_dbatch_t* bw_dbatch_upload(bw_aligner_t* a, const bw_batch_t* in) {
  bw_dbatch_t* b = nullptr;
  upload_impl(a, in, &b);
  return b;
}

// The batch's tiers were planned for the aligner that uploaded it.
static int check_owner(const bw_aligner_t* a, const bw_dbatch_t* b) {
  if (b->owner != a->id) {
    set_error("the device batch was uploaded by another aligner (its tiers are planned for that "
              "aligner's penalties, tuning and device)");
    return BW_STATUS_INVALID_ARG;
  }
  return 0;
}

int bw_set_schedule(bw_aligner_t* a, bw_dbatch_t* b, int schedule, const float* predicted_score) {
  if (!a || !b) return BW_STATUS_INVALID_ARG;
  if (const int st = check_owner(a, b)) return st;
  if (schedule == BW_SCHEDULE_LEARNED && !predicted_score) {
    set_error("the learned schedule needs one predicted score per pair");
    return BW_STATUS_INVALID_ARG;
  }
  if (schedule < BW_SCHEDULE_INPUT_ORDER || schedule > BW_SCHEDULE_LEARNED) {
    set_error("unknown schedule");
    return BW_STATUS_INVALID_ARG;
  }
  const int64_t n = b->n;
  if (b->a2p) {   // one list, input order or longest (p + t) first
    if (schedule == BW_SCHEDULE_LEARNED) {
      set_error("the learned schedule is for gap-affine batches");
      return BW_STATUS_INVALID_ARG;
    }
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    if (schedule == BW_SCHEDULE_LENGTH_SORTED)
      std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return (int64_t)b->h_plen[x] + b->h_tlen[x] > (int64_t)b->h_plen[y] + b->h_tlen[y];
      });
    a->schedule = schedule;
    HIP_TRY(hipSetDevice(a->dev));
    if (n > 0) HIP_TRY(hipMemcpyAsync(b->d_order, order.data(), 4 * n, hipMemcpyHostToDevice, a->stream));
    HIP_TRY(hipStreamSynchronize(a->stream));
    return 0;
  }
  const Penalties& q = a->pen;
  // cost key (larger first) and initial tier of every pair
  std::vector<double> key(n);
  std::vector<int> tier(n);
  for (int64_t i = 0; i < n; ++i) {
    const double pt = (double)b->h_plen[i] + b->h_tlen[i];
    if (schedule == BW_SCHEDULE_LEARNED) {
      // BiWFA work grows with the square of the score (both directions of every
      // recursion level cover ~s/e diagonals for ~s/2 scores); the widest ring a
      // pair needs is about its per-direction score over e.
      const double s = std::max(0.0, (double)predicted_score[i]) * ((q.um < 0) ? 2.0 : 1.0);
      key[i] = s * s / q.e + 4.0 * pt;
      tier[i] = initial_tier(b, (int)i, std::min(pt + 8.0, 1.25 * s / q.e + 8.0));
    } else {
      // no score estimate: pairs that fit the LDS staging start in tier 0, longer
      // ones in the first global tier as wide as a quarter of their length
      key[i] = (schedule == BW_SCHEDULE_LENGTH_SORTED) ? pt : 0.0;
      const bool stage = std::max(b->h_plen[i], b->h_tlen[i]) <= b->tiers[0].seq_cap;
      tier[i] = initial_tier(b, (int)i, stage ? 0.0 : std::min(pt + 8.0, 0.25 * pt + 64.0));
    }
  }
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  if (schedule != BW_SCHEDULE_INPUT_ORDER)
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return key[x] > key[y]; });
  const size_t nt = b->tiers.size();
  std::vector<std::vector<int32_t>> lists(nt);
  for (int32_t i : order) lists[tier[i]].push_back(i);
  // Tiers run one after another, so a single-wave global tier that holds fewer
  // than two pairs per aligner ends in a tail where most of the GPU idles.
  // Under the learned schedule such a tier hands its pairs to the next wider
  // one when that one has pairs too, so that one launch replaces two
  // under-filled ones (the predicted widths of a batch of like pairs straddle a
  // tier boundary: on 10,000 x 1 kbp pairs the split into two half-full
  // launches was 21 % slower than one launch).  The wider ring only costs memory.
  if (schedule == BW_SCHEDULE_LEARNED) {
    for (size_t t = 1; t + 1 < nt; ++t) {
      const Tier& T = b->tiers[t];
      if (T.lds || T.nw != 1 || lists[t].empty() || lists[t + 1].empty() ||
          (int64_t)lists[t].size() >= 2 * T.aligners)
        continue;
      lists[t + 1].insert(lists[t + 1].end(), lists[t].begin(), lists[t].end());
      lists[t].clear();
      std::stable_sort(lists[t + 1].begin(), lists[t + 1].end(), [&](int32_t x, int32_t y) { return key[x] > key[y]; });
    }
  }
  // Team tiers take their pairs longest first under the learned schedule too.
  // In predicted-cost order a team tier's first pairs are all the batch's
  // costliest, and their wide root searches then load HBM at the same time:
  // on the 200,000-pair mixed batch that order made the team tier 2.5 % slower
  // than length order with the same pairs in it.  The single-wave tiers keep
  // predicted-cost order, which shortens their tails.
  if (schedule == BW_SCHEDULE_LEARNED) {
    for (size_t t = 0; t < lists.size(); ++t) {
      if (b->tiers[t].lds || b->tiers[t].nw == 1) continue;
      std::stable_sort(lists[t].begin(), lists[t].end(), [&](int32_t x, int32_t y) {
        return (int64_t)b->h_plen[x] + b->h_tlen[x] > (int64_t)b->h_plen[y] + b->h_tlen[y];
      });
    }
  }
  a->schedule = schedule;
  HIP_TRY(hipSetDevice(a->dev));
  std::vector<int32_t> init(bw::kQctlInts * kMaxTiers, 0);
  b->tier_count.assign(nt, 0);
  for (size_t t = 0; t < nt; ++t) {
    init[bw::kQctlInts * t + 1] = (int32_t)lists[t].size();
    b->tier_count[t] = (int32_t)lists[t].size();
    if (!lists[t].empty())
      HIP_TRY(hipMemcpyAsync(b->d_lists + t * (size_t)n, lists[t].data(), 4 * lists[t].size(),
                             hipMemcpyHostToDevice, a->stream));
  }
  HIP_TRY(hipMemcpyAsync(b->d_init, init.data(), 4 * init.size(), hipMemcpyHostToDevice, a->stream));
  HIP_TRY(hipStreamSynchronize(a->stream));
  return 0;
}

static bw::KParams base_params(const bw_aligner_t* a, const bw_dbatch_t* b) {
  bw::KParams p{};
  p.words = b->d_words;
  p.p_woff = b->d_pwoff; p.p_len = b->d_plen;
  p.t_woff = b->d_twoff; p.t_len = b->d_tlen;
  p.x = a->pen.x; p.o = a->pen.o; p.e = a->pen.e; p.scope = a->pen.scope;
  p.um = a->pen.um; p.ux = a->pen.ux; p.uo = a->pen.uo; p.ue = a->pen.ue;
  p.gmask = null_score_mask(a->pen);
  p.score_only = a->attr.scope == BW_SCOPE_SCORE;
  p.max_steps = a->attr.max_alignment_steps > 0 ? a->attr.max_alignment_steps : INT32_MAX;
  p.score = b->d_score; p.wf_score = b->d_wf; p.status = b->d_status;
  p.runs = b->d_runs; p.runs_off = b->d_runs_off; p.nruns = b->d_nruns;
  p.task_min_length = kTaskMinLength;
  p.ready_ticks = a->ready_ticks;
  p.withhold_ready = a->withhold_ready;
  return p;
}

// The per-call launch sequence: reset of the tiers' counters (which also
// advances the call counter), then one persistent launch per tier.  Tier
// boundary events are recorded unless the sequence is being captured.
static int launch_sequence(bw_aligner_t* a, bw_dbatch_t* b, bool events) {
  hipStream_t s = a->stream;
  const int nt = (int)b->tiers.size();
  hipLaunchKernelGGL(bw::biwfa_reset, dim3(1), dim3(64), 0, s, a->d_counters, b->d_init,
                     bw::kQctlInts * kMaxTiers);
  HIP_TRY(hipGetLastError());
  if (events) {
    HIP_TRY(hipEventRecord(a->ev[1], s));
    HIP_TRY(hipEventRecord(a->tev[0], s));
  }
  for (int t = 0; t < nt; ++t) {
    const Tier& T = b->tiers[t];
    bw::KParams p = base_params(a, b);
    p.list = b->d_lists + (size_t)t * b->n;
    p.qctl = a->d_counters + bw::kQctlInts * t;
    p.queue = p.qctl;
    p.count = p.qctl + 1;
    p.task = 1;
    p.items = b->d_items;
    p.item_cap = b->item_cap;
    p.epoch = t;
    p.calls = a->d_counters + bw::kQctlInts * kMaxTiers;
    p.pend = b->d_pend;
    p.pfail = b->d_pfail;
    p.pflag = b->d_pflag;
    p.chunk = T.chunk;
    p.next_list = (t + 1 < nt) ? b->d_lists + (size_t)(t + 1) * b->n : nullptr;
    p.next_count = (t + 1 < nt) ? a->d_counters + bw::kQctlInts * (t + 1) + 1 : nullptr;
    p.L = T.L;
    p.gws = b->d_gws;
    p.gws_stride = T.stride;
    p.seq_lds_words = T.seq_lds_words;
    if (T.lds) {
      void* args[] = {&p};
      HIP_TRY(hipLaunchKernel(lds_kernel_for(T.L.wcap, default_penalties(a->pen)), dim3(T.blocks),
                              dim3(bw::kWave * T.waves_per_block), args, T.dyn_lds, s));
      HIP_TRY(hipGetLastError());
    } else {
      void* args[] = {&p};
      HIP_TRY(hipLaunchKernel(team_kernel(T.nw), dim3(T.blocks), dim3(bw::kWave * T.waves_per_block), args,
                              T.dyn_lds, s));
    }
    if (events) {
      HIP_TRY(hipEventRecord(a->tev[t + 1], s));
      if (t == 0) HIP_TRY(hipEventRecord(a->ev[2], s));
    }
  }
  return 0;
}

// Names of the kernel instantiations as profilers print them.
static const char* kernel_name(const Tier& T, bool defpen) {
  if (T.lds) {
    if (T.L.wcap == bw::kLdsDefaultWidth && defpen) return "bw::biwfa_persistent<bw::Cfg<true, 1, 32, true>>";
    if (T.L.wcap == bw::kLdsDefaultWidth) return "bw::biwfa_persistent<bw::Cfg<true, 1, 32, false>>";
    return "bw::biwfa_persistent<bw::Cfg<true, 1, 0, false>>";
  }
  switch (T.nw) {
    case 1: return "bw::biwfa_persistent<bw::Cfg<false, 1, 0, false>>";
    case 2: return "bw::biwfa_persistent<bw::Cfg<false, 2, 0, false>>";
    case 4: return "bw::biwfa_persistent<bw::Cfg<false, 4, 0, false>>";
    case 8: return "bw::biwfa_persistent<bw::Cfg<false, 8, 0, false>>";
    default: return "bw::biwfa_persistent<bw::Cfg<false, 16, 0, false>>";
  }
}

static void record_shape(bw_aligner_t* a, int t, const Tier& T) {
  const bool defpen = default_penalties(a->pen);
  bw_tier_shape_t& sh = a->tier_shape[t];
  sh = bw_tier_shape_t{};
  {
    sh.lds = T.lds ? 1 : 0;
    sh.waves_per_aligner = T.lds ? 1 : T.nw;
    sh.ring_width = T.L.wcap;
    sh.blocks = T.blocks;
    sh.waves_per_block = T.waves_per_block;
    sh.aligners = T.lds ? (int64_t)T.blocks * T.waves_per_block : T.aligners;
    std::snprintf(sh.kernel, sizeof(sh.kernel), "%s", kernel_name(T, defpen));
  }
}

static void record_shapes(bw_aligner_t* a, const bw_dbatch_t* b) {
  for (int t = 0; t < (int)b->tiers.size() && t < kMaxTiers; ++t) record_shape(a, t, b->tiers[t]);
}

// Gap-affine-2p: a counter reset and one launch of the single-wave 2p kernel
// (eager launches also in graph mode).
static int launch_2p(bw_aligner_t* a, bw_dbatch_t* b) {
  hipStream_t s = a->stream;
  a->last_counters = a->d_counters;
  bw_tier_shape_t& sh = a->tier_shape[0];
  sh = bw_tier_shape_t{};
  sh.lds = 0;
  sh.waves_per_aligner = 1;
  sh.ring_width = b->p2.wcap;
  sh.blocks = b->blocks2p;
  sh.waves_per_block = bw::kA2pWavesPerBlock;
  sh.aligners = b->p2.aligners;
  std::snprintf(sh.kernel, sizeof(sh.kernel), "%s", "bw2p::biwfa2p_persistent(bw::A2pParams)");
  HIP_TRY(hipEventRecord(a->ev[0], s));
  HIP_TRY(hipEventRecord(a->ev[1], s));
  HIP_TRY(hipEventRecord(a->tev[0], s));
  hipLaunchKernelGGL(bw2p::biwfa2p_reset, dim3(1), dim3(64), 0, s, a->d_counters, (int)b->n);
  HIP_TRY(hipGetLastError());
  bw::A2pParams p = b->p2;
  p.queue = a->d_counters;
  void* args[] = {&p};
  HIP_TRY(hipLaunchKernel(reinterpret_cast<const void*>(&bw2p::biwfa2p_persistent), dim3(b->blocks2p),
                          dim3(bw::kWave * bw::kA2pWavesPerBlock), args, 0, s));
  HIP_TRY(hipEventRecord(a->tev[1], s));
  HIP_TRY(hipEventRecord(a->ev[2], s));
  HIP_TRY(hipEventRecord(a->ev[3], s));
  a->last_ntiers = 1;
  a->last_graph = false;
  return 0;
}

int bw_dbatch_align(bw_aligner_t* a, bw_dbatch_t* b) {
  if (!a || !b) return BW_STATUS_INVALID_ARG;
  if (const int st = check_owner(a, b)) return st;
  HIP_TRY(hipSetDevice(a->dev));
  if (b->n == 0) return 0;
  hipStream_t s = a->stream;
  if (b->a2p) return launch_2p(a, b);
  const int nt = (int)b->tiers.size();
  a->last_counters = a->d_counters;
  if (a->graph) {
    if (!b->gexec) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      const int st = launch_sequence(a, b, false);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
      if (st != 0) {
        if (g) (void)hipGraphDestroy(g);
        return st;
      }
      HIP_TRY(e);
      const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HIP_TRY(ei);
    }
    HIP_TRY(hipEventRecord(a->ev[0], s));
    HIP_TRY(hipEventRecord(a->ev[1], s));
    record_shapes(a, b);
    HIP_TRY(hipGraphLaunch(b->gexec, s));
    HIP_TRY(hipEventRecord(a->ev[2], s));
    HIP_TRY(hipEventRecord(a->ev[3], s));
    a->last_ntiers = nt;
    a->last_graph = true;
    return 0;
  }
  record_shapes(a, b);
  HIP_TRY(hipEventRecord(a->ev[0], s));
  const int st = launch_sequence(a, b, true);
  if (st != 0) return st;
  HIP_TRY(hipEventRecord(a->ev[3], s));
  a->last_ntiers = nt;
  a->last_graph = false;
  return 0;
}

int bw_set_graph(bw_aligner_t* a, int enable) {
  if (!a) return BW_STATUS_INVALID_ARG;
  a->graph = enable ? 1 : 0;
  return 0;
}

int bw_synchronize(bw_aligner_t* a) {
  if (!a) return BW_STATUS_INVALID_ARG;
  HIP_TRY(hipSetDevice(a->dev));
  HIP_TRY(hipStreamSynchronize(a->stream));
  float total = 0.f, main_ms = 0.f, spill_ms = 0.f;
  if (hipEventElapsedTime(&total, a->ev[0], a->ev[3]) == hipSuccess) a->last_total_ms = total;
  if (hipEventElapsedTime(&main_ms, a->ev[1], a->ev[2]) == hipSuccess) a->last_main_ms = main_ms;
  if (hipEventElapsedTime(&spill_ms, a->ev[2], a->ev[3]) == hipSuccess) a->last_spill_ms = spill_ms;
  for (int t = 0; t < a->last_ntiers; ++t) {
    float ms = 0.f;
    a->tier_ms[t] = !a->last_graph && hipEventElapsedTime(&ms, a->tev[t], a->tev[t + 1]) == hipSuccess ? ms : 0.f;
  }
  int32_t cnt[bw::kQctlInts * kMaxTiers] = {};
  if (a->last_counters)
    HIP_TRY(hipMemcpy(cnt, a->last_counters, sizeof(int32_t) * bw::kQctlInts * a->last_ntiers,
                      hipMemcpyDeviceToHost));
  for (int t = 0; t < kMaxTiers; ++t) a->tier_pairs[t] = t < a->last_ntiers ? cnt[bw::kQctlInts * t + 1] : 0;
  for (int t = 0; t < a->last_ntiers; ++t) {
    if (cnt[bw::kQctlInts * t + 5] != 0) {
      set_error("tier " + std::to_string(t) + ": a shared recursion task made no progress (device flag " +
                std::to_string(cnt[bw::kQctlInts * t + 5]) + ")");
      return BW_STATUS_INTERNAL;
    }
  }
  return 0;
}

int bw_last_timing(const bw_aligner_t* a, float* total_ms, float* main_ms, float* spill_ms) {
  if (!a) return BW_STATUS_INVALID_ARG;
  if (total_ms) *total_ms = a->last_total_ms;
  if (main_ms) *main_ms = a->last_main_ms;
  if (spill_ms) *spill_ms = a->last_spill_ms;
  return 0;
}

int bw_dbatch_download(bw_aligner_t* a, bw_dbatch_t* b, bw_result_t* r) {
  if (!a || !b || !r) return BW_STATUS_INVALID_ARG;
  if (const int st = check_owner(a, b)) return st;
  HIP_TRY(hipSetDevice(a->dev));
  const int64_t n = b->n;
  hipStream_t s = a->stream;
  HIP_TRY(hipStreamSynchronize(s));
  if (n == 0) {
    if (r->cigar_offset) r->cigar_offset[0] = 0;
    r->cigar_total = 0;
    return 0;
  }
  if (r->score) HIP_TRY(hipMemcpyAsync(r->score, b->d_score, 4 * n, hipMemcpyDeviceToHost, s));
  if (r->wf_score) HIP_TRY(hipMemcpyAsync(r->wf_score, b->d_wf, 4 * n, hipMemcpyDeviceToHost, s));
  if (r->status) HIP_TRY(hipMemcpyAsync(r->status, b->d_status, 4 * n, hipMemcpyDeviceToHost, s));
  std::vector<int32_t> nruns(n);
  HIP_TRY(hipMemcpyAsync(nruns.data(), b->d_nruns, 4 * n, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  std::vector<int64_t> off(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) off[i + 1] = off[i] + std::max(0, nruns[i]);
  r->cigar_total = off[n];
  if (r->cigar_offset) std::memcpy(r->cigar_offset, off.data(), sizeof(int64_t) * (n + 1));
  if (r->cigar_runs) {
    if (off[n] > r->cigar_capacity) {
      set_error("cigar_capacity too small: need " + std::to_string(off[n]));
      return BW_STATUS_INVALID_ARG;
    }
    if (off[n] > 0) {
      int64_t* d_off = nullptr;
      uint32_t* d_out = nullptr;
      HIP_TRY(dalloc(&d_off, n + 1));
      HIP_TRY(dalloc(&d_out, off[n]));
      HIP_TRY(hipMemcpyAsync(d_off, off.data(), 8 * (n + 1), hipMemcpyHostToDevice, s));
      const int threads = 256, waves_per_block = threads / bw::kWave;
      const int blocks = (int)((n + waves_per_block - 1) / waves_per_block);
      hipLaunchKernelGGL(bw::biwfa_compact, dim3(blocks), dim3(threads), 0, s, b->d_runs,
                         b->d_runs_off, b->d_nruns, d_off, d_out, (int)n);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipMemcpyAsync(r->cigar_runs, d_out, 4 * off[n], hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      dfree(d_off);
      dfree(d_out);
    }
  }
  return 0;
}

int bw_align_batch(bw_aligner_t* a, const bw_batch_t* batch, bw_result_t* result) {
  bw_dbatch_t* b = nullptr;
  int st = upload_impl(a, batch, &b);
  if (st != 0) return st;
  st = bw_dbatch_align(a, b);
  if (st == 0) st = bw_synchronize(a);
  if (st == 0) st = bw_dbatch_download(a, b, result);
  bw_dbatch_free(b);
  return st;
}

// ---------------- launch tuning ----------------

int bw_set_tuning(bw_aligner_t* a, const bw_tuning_t* t) {
  if (!a || !t) return BW_STATUS_INVALID_ARG;
  const int w = t->lds_ring_width;
  if (w != 0 && w != 16 && w != 32 && w != 64 && w != 128) {
    set_error("lds_ring_width must be 0, 16, 32, 64 or 128");
    return BW_STATUS_INVALID_ARG;
  }
  if (t->lds_seq_cap < 0 || t->lds_seq_cap > 4096) {
    set_error("lds_seq_cap must be in [0, 4096]");
    return BW_STATUS_INVALID_ARG;
  }
  for (int i = 0; i < BW_MAX_TIERS; ++i) {
    const int v = t->team_waves[i];
    if (v != 0 && v != 1 && v != 2 && v != 4 && v != 8 && v != 16) {
      set_error("team_waves must be 0, 1, 2, 4, 8 or 16");
      return BW_STATUS_INVALID_ARG;
    }
    if (t->queue_chunk[i] < 0 || t->queue_chunk[i] > 1024) {
      set_error("queue_chunk must be in [0, 1024]");
      return BW_STATUS_INVALID_ARG;
    }
  }
  a->tuning = *t;
  drop_one(a);
  return 0;
}

int bw_get_tuning(const bw_aligner_t* a, bw_tuning_t* t) {
  if (!a || !t) return BW_STATUS_INVALID_ARG;
  *t = a->tuning;
  return 0;
}

int bw_last_tier_timing(const bw_aligner_t* a, float* ms, int capacity) {
  if (!a) return BW_STATUS_INVALID_ARG;
  for (int t = 0; t < a->last_ntiers && t < capacity; ++t) ms[t] = a->tier_ms[t];
  return a->last_ntiers;
}

int bw_last_tier_pairs(const bw_aligner_t* a, int32_t* pairs, int capacity) {
  if (!a) return BW_STATUS_INVALID_ARG;
  for (int t = 0; t < a->last_ntiers && t < capacity; ++t) pairs[t] = a->tier_pairs[t];
  return a->last_ntiers;
}

int bw_last_tier_shape(const bw_aligner_t* a, int tier, bw_tier_shape_t* shape) {
  if (!a || !shape || tier < 0 || tier >= a->last_ntiers) return BW_STATUS_INVALID_ARG;
  *shape = a->tier_shape[tier];
  return 0;
}

static void report(const bw_aligner_t* a, const std::string& msg) {
  if (a->progress) a->progress(msg.c_str(), a->progress_user);
}

// Runs the sample under tuning `t` (upload, one warm-up, `repeats` timed runs)
// and returns the mean per-tier and total milliseconds.
static int time_tuning(bw_aligner_t* a, const bw_batch_t* sample, const float* pred, const bw_tuning_t& t,
                       int repeats, std::vector<float>* tier_ms, float* total_ms) {
  a->tuning = t;
  bw_dbatch_t* b = nullptr;
  int st = upload_impl(a, sample, &b);
  if (st != 0) return st;
  if (pred) {
    st = bw_set_schedule(a, b, BW_SCHEDULE_LEARNED, pred);
    if (st != 0) { bw_dbatch_free(b); return st; }
  }
  tier_ms->assign(kMaxTiers, 0.f);
  *total_ms = 0.f;
  for (int r = 0; r <= repeats && st == 0; ++r) {
    st = bw_dbatch_align(a, b);
    if (st == 0) st = bw_synchronize(a);
    // one line per run: a candidate on long pairs can take minutes
    if (st == 0)
      report(a, "autotune run " + std::to_string(r) + " of " + std::to_string(repeats) +
                    (r == 0 ? " (warm-up): " : ": ") + std::to_string(a->last_total_ms) + " ms");
    if (st != 0 || r == 0) continue;
    for (int i = 0; i < a->last_ntiers; ++i) (*tier_ms)[i] += a->tier_ms[i] / repeats;
    *total_ms += a->last_total_ms / repeats;
  }
  bw_dbatch_free(b);
  return st;
}

int bw_set_progress(bw_aligner_t* a, bw_progress_fn fn, void* user) {
  if (!a) return BW_STATUS_INVALID_ARG;
  a->progress = fn;
  a->progress_user = user;
  return 0;
}

// Smallest team that keeps a wave at no more than 256 chunks of 64 diagonals
// per step on a ring of width wcap (fewer lanes only make such a tier crawl).
static int min_team_for(int wcap) {
  int nw = 1;
  while (nw < 16 && wcap > 256 * 64 * nw) nw *= 2;
  return nw;
}

int bw_autotune(bw_aligner_t* a, const bw_batch_t* sample, const float* pred, int repeats,
                bw_tuning_t* best) {
  if (!a || !sample) return BW_STATUS_INVALID_ARG;
  if (a->a2p) {
    set_error("autotuning tunes the gap-affine tiers; a gap-affine-2p aligner has none");
    return BW_STATUS_INVALID_ARG;
  }
  drop_one(a);   // the single-pair batch was planned for the tuning in force
  const int sched0 = a->schedule;
  if (repeats < 1) repeats = 1;
  const bw_tuning_t start = a->tuning;
  bw_tuning_t cur = start;
  std::vector<float> tms;
  float total = 0.f;
  // 1. waves per pair of every global tier: each tier's time depends only on its
  //    own team size (which pairs reach a tier is decided by capacities)
  static const int kTeams[] = {1, 2, 4, 8, 16};
  float best_tier[kMaxTiers];
  for (float& v : best_tier) v = 3.4e38f;
  bw_tuning_t pick = cur;
  for (int nw : kTeams) {
    bw_tuning_t t = cur;
    bool allowed[BW_MAX_TIERS];
    for (int i = 0; i < BW_MAX_TIERS; ++i) {
      const int lo = min_team_for(256 << (2 * i));
      allowed[i] = nw >= lo;
      t.team_waves[i] = std::max(nw, lo);
    }
    int st = time_tuning(a, sample, pred, t, repeats, &tms, &total);
    if (st != 0) { a->tuning = start; a->schedule = sched0; return st; }
    std::string line = "autotune team_waves=" + std::to_string(nw) + ": total " + std::to_string(total) + " ms, tiers";
    for (int i = 0; i < a->last_ntiers; ++i) line += " " + std::to_string(tms[i]);
    report(a, line);
    for (int i = 1; i < kMaxTiers; ++i) {
      if (allowed[i - 1] && tms[i] < best_tier[i]) {
        best_tier[i] = tms[i];
        pick.team_waves[i - 1] = nw;
      }
    }
  }
  cur = pick;
  // 2. LDS tier: queue chunk, then ring width (whole-batch time: the ring width
  //    moves pairs between tiers).  A sample with no pair in the LDS tier (long
  //    pairs only) cannot tell these settings apart, so they are kept.
  if (a->tier_pairs[0] == 0) {
    report(a, "autotune lds tier: no sample pair starts there, its settings are kept");
    a->tuning = cur;
    a->schedule = sched0;
    if (best) *best = cur;
    return 0;
  }
  float best_total = 3.4e38f;
  static const int kChunks[] = {1, 2, 4, 8, 16};
  pick = cur;
  for (int ch : kChunks) {
    bw_tuning_t t = cur;
    t.queue_chunk[0] = ch;
    int st = time_tuning(a, sample, pred, t, repeats, &tms, &total);
    if (st != 0) { a->tuning = start; a->schedule = sched0; return st; }
    report(a, "autotune lds queue_chunk=" + std::to_string(ch) + ": total " + std::to_string(total) + " ms");
    if (total < best_total) { best_total = total; pick = t; }
  }
  cur = pick;
  static const int kWidths[] = {16, 32, 64};
  for (int w : kWidths) {
    bw_tuning_t t = cur;
    t.lds_ring_width = w;
    int st = time_tuning(a, sample, pred, t, repeats, &tms, &total);
    if (st != 0) { a->tuning = start; a->schedule = sched0; return st; }
    report(a, "autotune lds_ring_width=" + std::to_string(w) + ": total " + std::to_string(total) + " ms");
    if (total < best_total) { best_total = total; pick = t; }
  }
  a->tuning = pick;
  a->schedule = sched0;
  if (best) *best = pick;
  return 0;
}

// ---------------- single pair (WFA2-style) ----------------

// Workspace bytes of the reused single-pair batch: enough aligners for one
// pair's shared recursion, far fewer than a full batch's.
constexpr int64_t kOneBudget = (int64_t)512 << 20;

// The reused device batch of the single-pair calls, sized for pairs up to
// cap_p x cap_t (one zeroed pair of that shape is uploaded to plan it).
static int reserve_one(bw_aligner_t* a, int cap_p, int cap_t) {
  bw_dbatch_free(a->one);
  a->one = nullptr;
  const int pw = (cap_p + 31) / 32, tw = (cap_t + 31) / 32;
  std::vector<uint64_t> words((size_t)1 + pw + tw + 2, 0ull);
  const int64_t pwoff = 1, twoff = 1 + pw;
  bw_batch_t bt;
  bt.num_pairs = 1;
  bt.num_words = (int64_t)words.size();
  bt.words = words.data();
  bt.pattern_word_offset = &pwoff;
  bt.pattern_length = &cap_p;
  bt.text_word_offset = &twoff;
  bt.text_length = &cap_t;
  return upload_impl(a, &bt, &a->one, kOneBudget);
}

static void drop_one(bw_aligner_t* a) {
  bw_dbatch_free(a->one);
  a->one = nullptr;
}

// One pair through the reused device batch: its tiers and workspace are
// planned once for the largest pair seen so far (the capacity grows by
// doubling), and each call only copies the pair in, aligns and copies the
// result out (one reused aligner per thread, as align_benchmark.c:521-523 does
// with wavefront_aligner_new).
static int align_one(bw_aligner_t* a, const std::vector<uint64_t>& words, int plen, int tlen) {
  auto failed = [&](int st) {
    a->status = st;
    a->runs.clear();
    a->ops.clear();
    return st;
  };
  if (hipSetDevice(a->dev) != hipSuccess) return failed(BW_STATUS_HIP_ERROR);
  bw_dbatch_t* b = a->one;
  if (!b || plen > b->max_p || tlen > b->max_t) {
    int cp = std::max(plen, 1024), ct = std::max(tlen, 1024);
    if (b) {
      cp = std::max(cp, plen > b->max_p ? 2 * b->max_p : b->max_p);
      ct = std::max(ct, tlen > b->max_t ? 2 * b->max_t : b->max_t);
    }
    if (const int st = reserve_one(a, cp, ct)) return failed(st);
    b = a->one;
  }
  const int pw = (plen + 31) / 32;
  const int64_t pwoff = 1, twoff = 1 + pw;
  hipStream_t s = a->stream;
  if (hipMemcpyAsync(b->d_words, words.data(), 8 * words.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(b->d_pwoff, &pwoff, 8, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(b->d_twoff, &twoff, 8, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(b->d_plen, &plen, 4, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(b->d_tlen, &tlen, 4, hipMemcpyHostToDevice, s) != hipSuccess) {
    set_error("copy of the pair to the device failed");
    return failed(BW_STATUS_HIP_ERROR);
  }
  b->h_plen[0] = plen;
  b->h_tlen[0] = tlen;
  const int sched = a->schedule == BW_SCHEDULE_LEARNED ? BW_SCHEDULE_LENGTH_SORTED : a->schedule;
  int st = bw_set_schedule(a, b, sched, nullptr);     // (synchronizes the stream: the copies are done)
  if (st == 0) st = bw_dbatch_align(a, b);
  if (st == 0) st = bw_synchronize(a);
  if (st != 0) return failed(st);
  int32_t out[4] = {0, 0, 0, 0};
  if (hipMemcpyAsync(&out[0], b->d_score, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&out[1], b->d_wf, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&out[2], b->d_status, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipMemcpyAsync(&out[3], b->d_nruns, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("copy of the result from the device failed");
    return failed(BW_STATUS_HIP_ERROR);
  }
  // a single pair's runs start at the batch's first run slot (runs_off[0] = 0)
  const int nr = std::max(0, std::min<int>(out[3], (int)b->runs_cap));
  a->runs.assign((size_t)nr, 0u);
  if (nr > 0 && (hipMemcpyAsync(a->runs.data(), b->d_runs, 4 * (size_t)nr, hipMemcpyDeviceToHost, s) != hipSuccess ||
                 hipStreamSynchronize(s) != hipSuccess)) {
    set_error("copy of the CIGAR from the device failed");
    return failed(BW_STATUS_HIP_ERROR);
  }
  a->status = out[2];
  a->score = out[0];
  a->wf_score = out[1];
  a->ops.clear();
  for (uint32_t run : a->runs) {
    const int op = run & 15u;
    const char ch = op == BW_OP_MATCH ? 'M' : op == BW_OP_MISMATCH ? 'X' : op == BW_OP_INSERTION ? 'I' : 'D';
    a->ops.insert(a->ops.end(), run >> 4, ch);
  }
  return a->status;
}

static int pack_ascii(const char* seq, int len, uint64_t* w) {
  for (int b = 0; b < len; b += 32) {
    uint64_t acc = 0;
    for (int j = 0; j < 32 && b + j < len; ++j) {
      uint64_t code;
      switch (seq[b + j]) {
        case 'A': case 'a': code = 0; break;
        case 'C': case 'c': code = 1; break;
        case 'G': case 'g': code = 2; break;
        case 'T': case 't': code = 3; break;
        default: return -1;
      }
      acc |= code << (2 * j);
    }
    w[b / 32] = acc;
  }
  return 0;
}

int bw_align(bw_aligner_t* a, const char* pattern, int plen, const char* text, int tlen) {
  if (!a || plen < 0 || tlen < 0 || (plen > 0 && !pattern) || (tlen > 0 && !text)) return BW_STATUS_INVALID_ARG;
  const int pw = (plen + 31) / 32, tw = (tlen + 31) / 32;
  std::vector<uint64_t> words((size_t)1 + pw + tw + 2, 0ull);
  if (pack_ascii(pattern, plen, words.data() + 1) != 0 || pack_ascii(text, tlen, words.data() + 1 + pw) != 0) {
    set_error("sequence byte outside ACGT");
    a->status = BW_STATUS_BAD_SEQUENCE;
    return BW_STATUS_BAD_SEQUENCE;
  }
  return align_one(a, words, plen, tlen);
}

int bw_align_packed2bits(bw_aligner_t* a, const uint8_t* pattern, int plen, const uint8_t* text, int tlen) {
  if (!a || plen < 0 || tlen < 0 || (plen > 0 && !pattern) || (tlen > 0 && !text)) return BW_STATUS_INVALID_ARG;
  // 4 bases per byte, LSB first: byte j of the input is bits 8*(j%8).. of word j/8.
  const int pw = (plen + 31) / 32, tw = (tlen + 31) / 32;
  std::vector<uint64_t> words((size_t)1 + pw + tw + 2, 0ull);
  uint8_t* wp = reinterpret_cast<uint8_t*>(words.data() + 1);
  uint8_t* wt = reinterpret_cast<uint8_t*>(words.data() + 1 + pw);
  std::memcpy(wp, pattern, (size_t)(plen + 3) / 4);
  std::memcpy(wt, text, (size_t)(tlen + 3) / 4);
  return align_one(a, words, plen, tlen);
}

int bw_get_score(const bw_aligner_t* a) { return a ? a->score : 0; }
int bw_get_wf_score(const bw_aligner_t* a) { return a ? a->wf_score : 0; }
int bw_get_status(const bw_aligner_t* a) { return a ? a->status : BW_STATUS_INVALID_ARG; }

int bw_get_cigar_ops(const bw_aligner_t* a, const char** ops, int* len) {
  if (!a || !ops || !len) return BW_STATUS_INVALID_ARG;
  *ops = a->ops.data();
  *len = (int)a->ops.size();
  return 0;
}

int bw_get_sam_cigar(bw_aligner_t* a, int show_mismatches, const uint32_t** buf, int* len) {
  if (!a || !buf || !len) return BW_STATUS_INVALID_ARG;
  if (show_mismatches) {
    *buf = a->runs.data();
    *len = (int)a->runs.size();
    return 0;
  }
  // fold X into M (SAM op 0) and merge, as cigar_compute_CIGAR does
  // (cigar.c:181-229): the fold starts at the second operation, so a leading
  // mismatch stays one X run of length 1
  a->sam.clear();
  bool first = true;
  for (uint32_t run : a->runs) {
    uint32_t op = run & 15u, n = run >> 4;
    if (n == 0) continue;
    if (first && op == BW_OP_MISMATCH) {
      a->sam.push_back((1u << 4) | BW_OP_MISMATCH);
      if (--n == 0) {
        first = false;
        continue;
      }
    }
    first = false;
    if (op == BW_OP_MATCH || op == BW_OP_MISMATCH) op = 0;
    if (!a->sam.empty() && (a->sam.back() & 15u) == op) a->sam.back() += n << 4;
    else a->sam.push_back((n << 4) | op);
  }
  *buf = a->sam.data();
  *len = (int)a->sam.size();
  return 0;
}

int bw_cigar_sprint(const bw_aligner_t* a, char* buffer, int cap) {
  if (!a || !buffer || cap <= 0) return -1;
  int cur = 0;
  for (uint32_t run : a->runs) {
    const int op = run & 15u;
    const char ch = op == BW_OP_MATCH ? 'M' : op == BW_OP_MISMATCH ? 'X' : op == BW_OP_INSERTION ? 'I' : 'D';
    const int w = std::snprintf(buffer + cur, (size_t)(cap - cur), "%u%c", run >> 4, ch);
    if (w < 0 || cur + w >= cap) return -1;
    cur += w;
  }
  buffer[cur] = '\0';
  return cur;
}

int bw_align_pair(const char* pattern, int plen, const char* text, int tlen, int x, int o, int e,
                  int32_t* score, char* ops, int ops_cap, int* ops_len) {
  static thread_local bw_aligner_t* tl_aligner = nullptr;
  static thread_local bw_affine_penalties_t tl_pen{0, 0, 0, 0};
  if (!tl_aligner || tl_pen.mismatch != x || tl_pen.gap_opening != o || tl_pen.gap_extension != e) {
    if (tl_aligner) bw_aligner_delete(tl_aligner);
    bw_attr_t attr;
    bw_attr_default(&attr);
    attr.penalties = bw_affine_penalties_t{0, x, o, e};
    tl_aligner = bw_aligner_new(&attr);
    tl_pen = attr.penalties;
    if (!tl_aligner) return BW_STATUS_NO_DEVICE;
  }
  const int st = bw_align(tl_aligner, pattern, plen, text, tlen);
  if (score) *score = tl_aligner->score;
  if (ops_len) *ops_len = (int)tl_aligner->ops.size();
  if (ops) {
    if ((int)tl_aligner->ops.size() > ops_cap) return BW_STATUS_INVALID_ARG;
    std::memcpy(ops, tl_aligner->ops.data(), tl_aligner->ops.size());
  }
  return st;
}


// ---------------- failure-path test hook ----------------

int bw_test_set_ready_guard(bw_aligner_t* a, int64_t ready_ticks, int withhold_ready) {
  if (!a || ready_ticks < 0) return BW_STATUS_INVALID_ARG;
  a->ready_ticks = ready_ticks > 0 ? ready_ticks : kReadyTicksDefault;
  a->withhold_ready = withhold_ready ? 1 : 0;
  return 0;
}

// ---------------- no-allocation entry point ----------------

int bw_align_device_noalloc(bw_aligner_t* a, const bw_batch_t* in, const bw_device_result_t* out) {
  if (!a || !in || !out) {
    set_error("bw_align_device_noalloc: null aligner, batch or result");
    return BW_STATUS_INVALID_ARG;
  }
  if (a->a2p) {
    set_error("bw_align_device_noalloc aligns gap-affine pairs; this aligner is gap-affine-2p");
    return BW_STATUS_INVALID_ARG;
  }
  if (in->num_pairs < 0 || in->num_pairs > kNaMaxPairs) {
    set_error("bw_align_device_noalloc: num_pairs must be in [0, " + std::to_string(kNaMaxPairs) + "]");
    return BW_STATUS_INVALID_ARG;
  }
  const int64_t n = in->num_pairs;
  if (n > 0 && (!in->words || !in->pattern_word_offset || !in->pattern_length || !in->text_word_offset ||
                !in->text_length || !out->score || !out->wf_score || !out->status || !out->num_runs ||
                !out->runs || !out->run_offset)) {
    set_error("bw_align_device_noalloc: a batch or result pointer is null");
    return BW_STATUS_INVALID_ARG;
  }
  if (!a->scratch) {
    set_error("bw_align_device_noalloc: no scratch address for this device");
    return BW_STATUS_INTERNAL;
  }
  HIP_TRY(hipSetDevice(a->dev));
  if (n == 0) return 0;
  // Tiers: the LDS tier over every admitted pair, then one single-wave global
  // tier (ring width 2048, the widest wavefront of any pair with p + t <= 2040)
  // over the pairs it passes on, with its slabs in the scratch.
  Tier t0 = make_lds_tier(a);
  if (const int st = check_lds_layout(a, t0)) return st;
  Tier t1 = make_global_tier(a, 0, 2048, true, kNaMaxPT, (kNaMaxPT + 31) / 32 + 3, kNaScratchBytes, 1);
  int32_t* const counters = a->scratch;
  int32_t* const list0 = counters + kNaCounterInts;
  int32_t* const list1 = list0 + kNaMaxPairs;
  int32_t* const slabs = list1 + kNaMaxPairs;
  const int64_t slab_ints = kNaScratchBytes / 4 - (slabs - counters);
  // as many single-wave aligners as the slabs hold (blocks of 4, or of 1 when
  // fewer than 4 fit: penalties with a deep base-case bound need large arenas)
  int64_t aligners = std::min<int64_t>(t1.aligners, slab_ints / t1.stride);
  if (aligners < 1) {
    set_error("bw_align_device_noalloc: one spill-tier workspace for these penalties exceeds the scratch");
    return BW_STATUS_INVALID_ARG;
  }
  if (aligners >= bw::kWavesPerBlock) {
    aligners -= aligners % bw::kWavesPerBlock;
  } else {
    t1.waves_per_block = 1;
    t1.dyn_lds = (size_t)t1.seq_lds_words * 16;
  }
  t1.blocks = (int)(aligners / t1.waves_per_block);
  t1.aligners = aligners;
  hipStream_t s = a->stream;
  bw::KParams p{};
  p.words = in->words;
  p.p_woff = in->pattern_word_offset; p.p_len = in->pattern_length;
  p.t_woff = in->text_word_offset; p.t_len = in->text_length;
  p.x = a->pen.x; p.o = a->pen.o; p.e = a->pen.e; p.scope = a->pen.scope;
  p.um = a->pen.um; p.ux = a->pen.ux; p.uo = a->pen.uo; p.ue = a->pen.ue;
  p.gmask = null_score_mask(a->pen);
  p.score_only = a->attr.scope == BW_SCOPE_SCORE;
  p.max_steps = a->attr.max_alignment_steps > 0 ? a->attr.max_alignment_steps : INT32_MAX;
  p.score = out->score; p.wf_score = out->wf_score; p.status = out->status;
  p.runs = out->runs; p.runs_off = out->run_offset; p.nruns = out->num_runs;
  p.task = 1;
  p.task_min_length = INT32_MAX;   // every pair depth first: no shared tasks, no item pool
  p.ready_ticks = a->ready_ticks;
  p.calls = counters + bw::kQctlInts * 2;
  HIP_TRY(hipEventRecord(a->ev[0], s));
  hipLaunchKernelGGL(bw::biwfa_noalloc_reset, dim3(1), dim3(64), 0, s, counters, bw::kQctlInts * 2);
  HIP_TRY(hipGetLastError());
  {
    const int threads = 256;
    const int blocks = (int)std::min<int64_t>((n + threads - 1) / threads, 4096);
    hipLaunchKernelGGL(bw::biwfa_noalloc_admit, dim3(blocks), dim3(threads), 0, s, p, n, in->num_words,
                       kNaMaxPT, list0, counters + 1);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipEventRecord(a->ev[1], s));
  HIP_TRY(hipEventRecord(a->tev[0], s));
  const Tier* tiers[2] = {&t0, &t1};
  for (int t = 0; t < 2; ++t) {
    const Tier& T = *tiers[t];
    record_shape(a, t, T);
    bw::KParams q = p;
    q.list = t == 0 ? list0 : list1;
    q.qctl = counters + bw::kQctlInts * t;
    q.queue = q.qctl;
    q.count = q.qctl + 1;
    q.epoch = t;
    q.chunk = T.chunk;
    q.next_list = t == 0 ? list1 : nullptr;
    q.next_count = t == 0 ? counters + bw::kQctlInts + 1 : nullptr;
    q.L = T.L;
    q.gws = slabs;
    q.gws_stride = T.stride;
    q.seq_lds_words = T.seq_lds_words;
    void* args[] = {&q};
    const void* kern = T.lds ? lds_kernel_for(T.L.wcap, default_penalties(a->pen)) : team_kernel(1);
    HIP_TRY(hipLaunchKernel(kern, dim3(T.blocks), dim3(bw::kWave * T.waves_per_block), args, T.dyn_lds, s));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(a->tev[t + 1], s));
    if (t == 0) HIP_TRY(hipEventRecord(a->ev[2], s));
  }
  HIP_TRY(hipEventRecord(a->ev[3], s));
  a->last_counters = counters;
  a->last_ntiers = 2;
  a->last_graph = false;
  return 0;
}

}  // extern "C"
