// Seeded synthetic pair generator and 2-bit packer (host side of libbiwfa_mi355x).
//
// The reference's generators (src/generate_dataset*.cpp, WFA2-lib
// tools/generate_dataset/generate_dataset.c:330-400) seed from the clock and
// produce substitution-only or X/I/D 1:1:1 edits.  This generator is
// deterministic per (seed, pair index): pair i draws from its own
// xoshiro256** stream seeded through SplitMix64, so a batch can be generated in
// parallel and any pair regenerated alone.
//
// Per pair:
//   1. pattern length: fixed, or round(exp(U(ln len_min, ln len_max)));
//   2. pattern bases: uniform over ACGT;
//   3. number of edit events: fixed, or round(U(rate_min, rate_max) * plen);
//   4. event positions: distinct pattern coordinates (rejection sampling);
//   5. per event, in ascending position order: the kind (a run event with
//      probability run_frac, else single-base X:I:D = 1:1:1), the run length,
//      and the new bases.
// The text is built left to right: X replaces pattern[pos] by a different
// uniform base, I inserts `len` uniform bases before pattern[pos], D drops
// `len` pattern bases from pos on (events inside a dropped stretch are void).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "biwfa_mi355x.h"

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  Rng(uint64_t seed, uint64_t stream) {
    uint64_t x = seed * 0xD1B54A32D192ED03ull + stream * 0x9E3779B97F4A7C15ull + 1;
    for (int i = 0; i < 4; ++i) s[i] = splitmix(x);
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {  // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

int pattern_length(Rng& r, const bw_dataset_params_t& p) {
  if (p.len_max <= p.len_min) return p.len_min;
  const double lo = std::log((double)p.len_min), hi = std::log((double)p.len_max);
  return (int)std::lround(std::exp(lo + (hi - lo) * r.uniform()));
}

int run_length(Rng& r, const bw_dataset_params_t& p, bool run_event) {
  if (!run_event) {
    if (p.geometric_runs) {  // Geometric(p=0.5) on {1,2,...}
      int n = 1;
      while (r.next() & 1) ++n;
      return n;
    }
    return 1;
  }
  return p.run_min + (int)r.below((uint32_t)(p.run_max - p.run_min + 1));
}

// Generates pair `idx` into pat / txt (cleared first).
void gen_pair(const bw_dataset_params_t& prm, int64_t idx, std::vector<char>& pat,
              std::vector<char>& txt, std::vector<uint8_t>& mark, std::vector<int>& pos) {
  Rng r(prm.seed, (uint64_t)idx);
  const int plen = pattern_length(r, prm);
  pat.resize(plen);
  for (int i = 0; i < plen; i += 32) {
    uint64_t w = r.next();
    for (int j = 0; j < 32 && i + j < plen; ++j) pat[i + j] = kBases[(w >> (2 * j)) & 3];
  }
  int nedits = prm.num_edits;
  if (nedits < 0) {
    const double rate = prm.rate_min + (prm.rate_max - prm.rate_min) * r.uniform();
    nedits = (int)std::lround(rate * plen);
  }
  nedits = std::min(nedits, plen / 2);
  mark.assign(plen, 0);
  pos.clear();
  while ((int)pos.size() < nedits) {
    const int q = (int)r.below((uint32_t)plen);
    if (!mark[q]) { mark[q] = 1; pos.push_back(q); }
  }
  std::sort(pos.begin(), pos.end());
  txt.clear();
  txt.reserve(plen + plen / 4 + 16);
  int cursor = 0;  // next pattern base to copy
  for (int q : pos) {
    const bool run_event = prm.run_frac > 0.0 && r.uniform() < prm.run_frac;
    int kind;  // 0:X 1:I 2:D
    if (run_event) kind = 1 + (int)r.below(2);
    else kind = (int)r.below(3);
    const int len = (kind == 0) ? 1 : run_length(r, prm, run_event);
    uint64_t bases = r.next();
    if (q < cursor) continue;  // swallowed by an earlier deletion
    txt.insert(txt.end(), pat.begin() + cursor, pat.begin() + q);
    cursor = q;
    if (kind == 0) {
      const int old = (int)(std::find(kBases, kBases + 4, pat[q]) - kBases);
      txt.push_back(kBases[(old + 1 + (int)(bases % 3)) & 3]);
      cursor = q + 1;
    } else if (kind == 1) {
      for (int j = 0; j < len; ++j) {
        if ((j & 31) == 0 && j) bases = r.next();
        txt.push_back(kBases[(bases >> (2 * (j & 31))) & 3]);
      }
    } else {
      cursor = std::min(plen, q + len);
    }
  }
  txt.insert(txt.end(), pat.begin() + cursor, pat.end());
}

}  // namespace

extern "C" {

int bw_dataset_lengths(const bw_dataset_params_t* prm, int32_t* p_len, int32_t* t_len) {
  if (!prm || prm->num_pairs < 0) return BW_STATUS_INVALID_ARG;
  #pragma omp parallel
  {
    std::vector<char> pat, txt;
    std::vector<uint8_t> mark;
    std::vector<int> pos;
    #pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < prm->num_pairs; ++i) {
      gen_pair(*prm, i, pat, txt, mark, pos);
      p_len[i] = (int32_t)pat.size();
      t_len[i] = (int32_t)txt.size();
    }
  }
  return 0;
}

int bw_dataset_generate(const bw_dataset_params_t* prm, const int64_t* p_off,
                        const int64_t* t_off, char* seq) {
  if (!prm || prm->num_pairs < 0) return BW_STATUS_INVALID_ARG;
  #pragma omp parallel
  {
    std::vector<char> pat, txt;
    std::vector<uint8_t> mark;
    std::vector<int> pos;
    #pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < prm->num_pairs; ++i) {
      gen_pair(*prm, i, pat, txt, mark, pos);
      if (!pat.empty()) std::memcpy(seq + p_off[i], pat.data(), pat.size());
      if (!txt.empty()) std::memcpy(seq + t_off[i], txt.data(), txt.size());
    }
  }
  return 0;
}

// 2-bit packing: base j of a sequence at bits 2*(j%32) of word j/32 (A=0 C=1
// G=2 T=3, the code order of WFA2's wavefront_sequences.c:110-123 decoder).
// Bytes other than ACGT (either case) are rejected (returns the index of the
// first offending sequence as -(1000+i)).
int bw_pack_2bit(int64_t num_seqs, const char* ascii, const int64_t* off,
                 const int32_t* len, const int64_t* word_off, uint64_t* words) {
  int64_t bad = -1;
  #pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < num_seqs; ++i) {
    const char* s = ascii + off[i];
    uint64_t* w = words + word_off[i];
    const int n = len[i];
    for (int b = 0; b < n; b += 32) {
      uint64_t acc = 0;
      const int m = std::min(32, n - b);
      for (int j = 0; j < m; ++j) {
        uint64_t code;
        switch (s[b + j]) {
          case 'A': case 'a': code = 0; break;
          case 'C': case 'c': code = 1; break;
          case 'G': case 'g': code = 2; break;
          case 'T': case 't': code = 3; break;
          default:
            code = 0;
            #pragma omp critical
            { if (bad < 0 || i < bad) bad = i; }
        }
        acc |= code << (2 * j);
      }
      w[b >> 5] = acc;
    }
  }
  return bad >= 0 ? (int)(-(1000 + bad)) : 0;
}

}  // extern "C"
