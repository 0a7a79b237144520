// host_check.cpp -- exercises the library's host-only C/C++ (the seeded
// generator and 2-bit packer, the .seq reader, the .alg writer) and the C
// oracle under AddressSanitizer + UndefinedBehaviorSanitizer.  Built and run by
// tests/test_sanitizers.py with g++ -fsanitize=address,undefined (no GPU code
// is compiled into it).
//
//   host_check UTEST_SEQ GOLDEN_DIR OUT_DIR
//
// Aligns every utest pair with the oracle under gap-affine (full and score-only)
// and gap-affine-2p, writes each result set with bw_alg_write and compares the
// files byte for byte with WFA2's golden .alg files; generates, packs and reads
// back small seeded batches of every shape (including empty sequences and the
// repo's "N P T" file format).  Exit status 0 = everything matched.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "biwfa_mi355x.h"
#include "biwfa_oracle.h"

namespace bw_host {
// The runtime's error-text setter, which the sequence reader reports through.
void set_error(const std::string& msg) { std::fprintf(stderr, "[host_check] library error: %s\n", msg.c_str()); }
}  // namespace bw_host

namespace {

int failures = 0;

void check(bool ok, const std::string& what) {
  if (!ok) {
    std::fprintf(stderr, "[host_check] FAILED: %s\n", what.c_str());
    ++failures;
  }
}

std::string slurp(const std::string& path) {
  std::string s;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return s;
  char buf[65536];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
  std::fclose(f);
  return s;
}

struct Pairs {
  std::vector<std::string> p, t;
};

// Every pair of a .seq file through the product reader, and its packed view.
Pairs read_pairs(const char* path, int format, bw_batch_t* batch_out) {
  Pairs out;
  bw_seqs_t* s = bw_seqs_read(path, format);
  check(s != nullptr, std::string("bw_seqs_read ") + path);
  if (!s) return out;
  const int64_t n = bw_seqs_count(s);
  for (int64_t i = 0; i < n; ++i) {
    const char *p = nullptr, *t = nullptr;
    int pl = 0, tl = 0;
    check(bw_seqs_get(s, i, &p, &pl, &t, &tl) == 0, "bw_seqs_get");
    out.p.emplace_back(p, (size_t)pl);
    out.t.emplace_back(t, (size_t)tl);
  }
  bw_batch_t b{};
  check(bw_seqs_batch(s, &b) == 0, "bw_seqs_batch");
  check(b.num_pairs == n, "packed pair count");
  // unpack every base back from the words and compare with the ASCII
  static const char kBase[4] = {'A', 'C', 'G', 'T'};
  for (int64_t i = 0; i < n; ++i) {
    for (int side = 0; side < 2; ++side) {
      const std::string& seq = side ? out.t[(size_t)i] : out.p[(size_t)i];
      const int64_t w0 = side ? b.text_word_offset[i] : b.pattern_word_offset[i];
      for (size_t j = 0; j < seq.size(); ++j) {
        const uint64_t word = b.words[w0 + (int64_t)(j / 32)];
        if (kBase[(word >> (2 * (j % 32))) & 3u] != seq[j]) {
          check(false, "2-bit round trip of pair " + std::to_string(i));
          j = seq.size();
        }
      }
    }
  }
  if (batch_out) *batch_out = b;
  bw_seqs_free(s);
  return out;
}

// Aligns every pair with the oracle and writes the .alg file.
void align_and_write(const Pairs& pr, const int* pen, bool two_piece, bool score_only, const std::string& path) {
  const int64_t n = (int64_t)pr.p.size();
  std::vector<int32_t> score((size_t)n, 0);
  std::vector<int64_t> off((size_t)n + 1, 0);
  std::vector<uint32_t> runs;
  for (int64_t i = 0; i < n; ++i) {
    const std::string &P = pr.p[(size_t)i], &T = pr.t[(size_t)i];
    const int cap = (int)(P.size() + T.size()) + 1;
    std::vector<uint32_t> r((size_t)cap);
    int nr = 0, alg = 0, wf = 0;
    int st;
    if (two_piece) {
      // the batch entry, one thread, over this one pair
      const std::string buf = P + T;
      const int64_t po = 0, to = (int64_t)P.size();
      const int32_t pl = (int32_t)P.size(), tl = (int32_t)T.size();
      const int64_t roff[2] = {0, cap};
      int32_t a = 0, w = 0, s = 0, k = 0;
      int64_t cells = 0, words = 0;
      double el = 0;
      st = bwo_align_batch_2p(1, buf.data(), &po, &pl, &to, &tl, pen[0], pen[1], pen[2], pen[3], pen[4], pen[5],
                              score_only ? 1 : 0, INT32_MAX, 1, &a, &w, &s, r.data(), roff, &k, &cells, &words, &el);
      st = st != 0 ? st : s;
      alg = a;
      nr = k;
    } else {
      st = bwo_align(P.data(), (int)P.size(), T.data(), (int)T.size(), pen[0], pen[1], pen[2], pen[3],
                     score_only ? 1 : 0, INT32_MAX, r.data(), cap, &nr, &alg, &wf, nullptr);
    }
    check(st == BWO_STATUS_COMPLETED, "oracle status of pair " + std::to_string(i));
    score[(size_t)i] = alg;
    if (!score_only) runs.insert(runs.end(), r.begin(), r.begin() + nr);
    off[(size_t)i + 1] = (int64_t)runs.size();
  }
  check(bw_alg_write(path.c_str(), n, score.data(), off.data(), runs.data(), score_only ? 1 : 0) == 0,
        "bw_alg_write " + path);
}

// A small seeded batch: lengths, sequences, packing, and every base unpacked again.
void generate(const bw_dataset_params_t& prm) {
  const int64_t n = prm.num_pairs;
  std::vector<int32_t> pl((size_t)n), tl((size_t)n);
  check(bw_dataset_lengths(&prm, pl.data(), tl.data()) == 0, "bw_dataset_lengths");
  std::vector<int64_t> po((size_t)n), to((size_t)n), pw((size_t)n), tw((size_t)n);
  int64_t c = 0, w = 1;
  for (int64_t i = 0; i < n; ++i) {
    po[(size_t)i] = c; c += pl[(size_t)i];
    to[(size_t)i] = c; c += tl[(size_t)i];
    pw[(size_t)i] = w; w += (pl[(size_t)i] + 31) / 32;
    tw[(size_t)i] = w; w += (tl[(size_t)i] + 31) / 32;
  }
  std::vector<char> seq((size_t)c);
  std::vector<uint64_t> words((size_t)w + 2, 0ull);
  check(bw_dataset_generate(&prm, po.data(), to.data(), seq.data()) == 0, "bw_dataset_generate");
  check(bw_pack_2bit(n, seq.data(), po.data(), pl.data(), pw.data(), words.data()) == 0, "bw_pack_2bit P");
  check(bw_pack_2bit(n, seq.data(), to.data(), tl.data(), tw.data(), words.data()) == 0, "bw_pack_2bit T");
  static const char kBase[4] = {'A', 'C', 'G', 'T'};
  for (int64_t i = 0; i < n; ++i)
    for (int32_t j = 0; j < pl[(size_t)i]; ++j)
      if (kBase[(words[(size_t)(pw[(size_t)i] + j / 32)] >> (2 * (j % 32))) & 3u] != seq[(size_t)(po[(size_t)i] + j)]) {
        check(false, "generated pair " + std::to_string(i) + " packs");
        break;
      }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s UTEST_SEQ GOLDEN_DIR OUT_DIR\n", argv[0]);
    return 2;
  }
  const std::string seq = argv[1], gold = argv[2], out = argv[3];
  bw_batch_t b{};
  const Pairs utest = read_pairs(seq.c_str(), BW_SEQ_FORMAT_AUTO, &b);
  check(utest.p.size() == 305, "305 utest pairs");

  const int affine[4] = {0, 4, 6, 2};
  const int affine2p[6] = {0, 4, 6, 2, 24, 1};
  // gap-affine over every pair; gap-affine-2p (a deeper ring, slower under the
  // sanitizers) over the first 64, against the golden file's first 64 lines
  struct Case {
    const int* pen;
    bool two_piece, score_only;
    const char* golden;
    size_t pairs;
  } cases[] = {
      {affine, false, false, "test.biwfa.affine.alg", 305},
      {affine, false, true, "test.biwfa.score.affine.alg", 305},
      {affine2p, true, false, "test.biwfa.affine2p.alg", 64},
      {affine2p, true, true, "test.biwfa.score.affine2p.alg", 64},
  };
  for (const Case& c : cases) {
    const std::string path = out + "/" + c.golden;
    Pairs sub = utest;
    sub.p.resize(std::min(c.pairs, sub.p.size()));
    sub.t.resize(sub.p.size());
    align_and_write(sub, c.pen, c.two_piece, c.score_only, path);
    const std::string want = slurp(gold + "/" + c.golden);
    size_t end = 0;
    for (size_t i = 0; i < sub.p.size() && end != std::string::npos; ++i) end = want.find('\n', end) + 1;
    check(end != 0 && slurp(path) == want.substr(0, end), std::string("byte-equal ") + c.golden);
  }

  // the repo's "N P T" format (every pattern P and every text T bases long)
  const std::string repo = out + "/repo_format.txt";
  {
    FILE* f = std::fopen(repo.c_str(), "wb");
    std::fputs("3 8 7\nACGTACGT\nACGAACG\nGATTACAG\nGATACAG\nACGTTTTT\nACCTTTT\n", f);
    std::fclose(f);
  }
  const Pairs rp = read_pairs(repo.c_str(), BW_SEQ_FORMAT_REPO, nullptr);
  check(rp.p.size() == 3 && rp.p[2] == "ACGTTTTT" && rp.t[2] == "ACCTTTT", "repo-format pairs");
  // WFA2 format with an empty text and an empty pattern: one gap run each,
  // -(o + e * len) under match 0
  const std::string edge = out + "/edge.seq";
  {
    FILE* f = std::fopen(edge.c_str(), "wb");
    std::fputs(">ACGT\n<\n>\n<GGA\n>ACGT\n<ACGT\n", f);
    std::fclose(f);
  }
  const Pairs ep = read_pairs(edge.c_str(), BW_SEQ_FORMAT_WFA2, nullptr);
  check(ep.p.size() == 3 && ep.t[0].empty() && ep.p[1].empty(), "empty sequences read");
  align_and_write(ep, affine, false, false, out + "/edge.alg");
  check(slurp(out + "/edge.alg") == "-14\t4D\n-12\t3I\n0\t4M\n", "edge-case .alg lines");

  // seeded batches of each shape: fixed edit counts, rates, geometric and long indel runs
  bw_dataset_params_t shapes[4] = {};
  shapes[0] = bw_dataset_params_t{64, 1, 1000, 1000, -1, 0.05, 0.05, 0, 0.0, 0, 0};
  shapes[1] = bw_dataset_params_t{256, 2, 150, 150, -1, 0.02, 0.02, 0, 0.0, 0, 0};
  shapes[2] = bw_dataset_params_t{16, 3, 100, 50000, -1, 0.01, 0.15, 1, 0.0, 0, 0};
  shapes[3] = bw_dataset_params_t{4, 5, 100000, 100000, -1, 0.15, 0.15, 0, 0.10, 10, 200};
  for (const bw_dataset_params_t& s : shapes) generate(s);

  if (failures) {
    std::fprintf(stderr, "[host_check] %d check(s) failed\n", failures);
    return 1;
  }
  std::printf("host_check ok\n");
  return 0;
}
