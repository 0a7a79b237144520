// Sequence-file input and .alg output of libbiwfa_mi355x (host side).
//
// Input formats:
//   * WFA2's .seq (align_benchmark_read_input, tools/align_benchmark/
//     align_benchmark.c:313-338): two lines per pair, the first holds the
//     pattern and the second the text, each after a one-character marker
//     ('>' / '<'; the reader does not look at it, so generate_dataset's
//     '<' first lines for a longer pattern, generate_dataset.c:387-393, are
//     read the same way).  As there, a line's length minus two (marker and
//     newline) is the sequence length, so a last line without a newline loses
//     its last base.
//   * the repo driver's input (src/biWFA.cu:776-835): a header "N P T", then N
//     pairs of lines of which the first P (pattern) and T (text) characters
//     are the sequences.
// Output: align_benchmark's lite .alg lines (benchmark_print_output_lite and
// benchmark_print_output, tools/align_benchmark/benchmark/benchmark_utils.c:
// 102-117 and 143-179): "score<TAB>CIGAR", "-" for an empty CIGAR, and -1 as
// the score of a full alignment with an empty CIGAR.
#include <algorithm>
#include <cerrno>
#include <cinttypes>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "biwfa_mi355x.h"

namespace bw_host {
void set_error(const std::string& msg);   // runtime.cpp: the text bw_last_error returns
}
using bw_host::set_error;

struct bw_seqs {
  std::string ascii;                      // every sequence back to back
  std::vector<int64_t> p_off, t_off;      // offsets into ascii
  std::vector<int32_t> p_len, t_len;
  // 2-bit packed view (bw_seqs_batch)
  std::vector<uint64_t> words;
  std::vector<int64_t> p_woff, t_woff;
};

namespace {

struct File {
  FILE* f = nullptr;
  explicit File(const char* path) : f(std::fopen(path, "rb")) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// One line with its newline, as getline returns it; false at end of file.
bool read_line(FILE* f, std::string& line) {
  line.clear();
  int c;
  while ((c = std::fgetc(f)) != EOF) {
    line.push_back((char)c);
    if (c == '\n') break;
  }
  return !line.empty();
}

bool add_pair(bw_seqs* s, const char* p, int64_t pl, const char* t, int64_t tl, int64_t index) {
  if (pl < 0 || tl < 0 || pl > (1 << 27) || tl > (1 << 27)) {
    set_error("pair " + std::to_string(index) + ": bad sequence length");
    return false;
  }
  s->p_off.push_back((int64_t)s->ascii.size());
  s->p_len.push_back((int32_t)pl);
  s->ascii.append(p, (size_t)pl);
  s->t_off.push_back((int64_t)s->ascii.size());
  s->t_len.push_back((int32_t)tl);
  s->ascii.append(t, (size_t)tl);
  return true;
}

bool read_wfa2(FILE* f, bw_seqs* s) {
  std::string l1, l2;
  int64_t n = 0;
  while (read_line(f, l1)) {
    if (!read_line(f, l2)) {
      set_error("pair " + std::to_string(n) + ": pattern line without a text line");
      return false;
    }
    if (l1.size() < 2 || l2.size() < 2) {
      set_error("pair " + std::to_string(n) + ": line shorter than marker plus newline");
      return false;
    }
    const int64_t pl = (int64_t)l1.size() - 2, tl = (int64_t)l2.size() - 2;
    if (!add_pair(s, l1.data() + 1, pl, l2.data() + 1, tl, n)) return false;
    ++n;
  }
  return true;
}

bool read_repo(FILE* f, bw_seqs* s) {
  std::string h, l1, l2;
  if (!read_line(f, h)) {
    set_error("empty input file");
    return false;
  }
  long long n = 0, pl = 0, tl = 0;
  if (std::sscanf(h.c_str(), "%lld %lld %lld", &n, &pl, &tl) != 3 || n < 0 || pl < 0 || tl < 0) {
    set_error("header is not \"num_alignments pattern_length text_length\"");
    return false;
  }
  for (long long i = 0; i < n; ++i) {
    if (!read_line(f, l1) || !read_line(f, l2)) {
      set_error("alignment " + std::to_string(i) + ": missing pattern or text line");
      return false;
    }
    auto body = [](const std::string& l) { return (int64_t)(l.size() - (!l.empty() && l.back() == '\n')); };
    if (body(l1) < pl || body(l2) < tl) {
      set_error("alignment " + std::to_string(i) + ": line shorter than the header's length");
      return false;
    }
    if (!add_pair(s, l1.data(), pl, l2.data(), tl, i)) return false;
  }
  return true;
}

}  // namespace

extern "C" {

bw_seqs_t* bw_seqs_read(const char* path, int format) {
  if (!path || format < BW_SEQ_FORMAT_AUTO || format > BW_SEQ_FORMAT_REPO) {
    set_error("bw_seqs_read: invalid argument");
    return nullptr;
  }
  File in(path);
  if (!in.f) {
    set_error(std::string("cannot open ") + path + ": " + std::strerror(errno));
    return nullptr;
  }
  if (format == BW_SEQ_FORMAT_AUTO) {
    const int c = std::fgetc(in.f);
    format = (c == '>' || c == '<' || c == EOF) ? BW_SEQ_FORMAT_WFA2 : BW_SEQ_FORMAT_REPO;
    std::rewind(in.f);
  }
  bw_seqs* s = new bw_seqs();
  const bool ok = format == BW_SEQ_FORMAT_WFA2 ? read_wfa2(in.f, s) : read_repo(in.f, s);
  if (!ok || std::ferror(in.f)) {
    if (ok) set_error(std::string("read error on ") + path);
    delete s;
    return nullptr;
  }
  return s;
}

int64_t bw_seqs_count(const bw_seqs_t* s) { return s ? (int64_t)s->p_len.size() : 0; }

int bw_seqs_get(const bw_seqs_t* s, int64_t i, const char** pattern, int* pattern_length, const char** text,
                int* text_length) {
  if (!s || i < 0 || i >= (int64_t)s->p_len.size()) return BW_STATUS_INVALID_ARG;
  if (pattern) *pattern = s->ascii.data() + s->p_off[i];
  if (pattern_length) *pattern_length = s->p_len[i];
  if (text) *text = s->ascii.data() + s->t_off[i];
  if (text_length) *text_length = s->t_len[i];
  return 0;
}

int bw_seqs_batch(bw_seqs_t* s, bw_batch_t* batch) {
  if (!s || !batch) return BW_STATUS_INVALID_ARG;
  const int64_t n = (int64_t)s->p_len.size();
  // sequences word-aligned back to back after one zero word, two zero words at
  // the end (the bw_batch_t layout the Python packer writes too)
  s->p_woff.assign(n, 0);
  s->t_woff.assign(n, 0);
  int64_t w = 1;
  for (int64_t i = 0; i < n; ++i) {
    s->p_woff[i] = w;
    w += (s->p_len[i] + 31) / 32;
    s->t_woff[i] = w;
    w += (s->t_len[i] + 31) / 32;
  }
  w += 2;
  s->words.assign((size_t)w, 0ull);
  if (n > 0) {
    const int st = bw_pack_2bit(n, s->ascii.data(), s->p_off.data(), s->p_len.data(), s->p_woff.data(),
                                s->words.data());
    const int st2 = bw_pack_2bit(n, s->ascii.data(), s->t_off.data(), s->t_len.data(), s->t_woff.data(),
                                 s->words.data());
    if (st != 0 || st2 != 0) {
      const int bad = std::min(st != 0 ? -st - 1000 : INT32_MAX, st2 != 0 ? -st2 - 1000 : INT32_MAX);
      set_error("pair " + std::to_string(bad) + ": sequence byte outside ACGT");
      return BW_STATUS_BAD_SEQUENCE;
    }
  }
  batch->num_pairs = n;
  batch->num_words = w;
  batch->words = s->words.data();
  batch->pattern_word_offset = s->p_woff.data();
  batch->pattern_length = s->p_len.data();
  batch->text_word_offset = s->t_woff.data();
  batch->text_length = s->t_len.data();
  return 0;
}

void bw_seqs_free(bw_seqs_t* s) { delete s; }

int bw_alg_write(const char* path, int64_t num_pairs, const int32_t* score, const int64_t* cigar_offset,
                 const uint32_t* cigar_runs, int score_only) {
  if (!path || num_pairs < 0 || (num_pairs > 0 && !score) || (!score_only && num_pairs > 0 && !cigar_offset)) {
    set_error("bw_alg_write: invalid argument");
    return BW_STATUS_INVALID_ARG;
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    set_error(std::string("cannot open ") + path + ": " + std::strerror(errno));
    return BW_STATUS_INVALID_ARG;
  }
  std::string line;
  for (int64_t i = 0; i < num_pairs; ++i) {
    const int64_t b = score_only ? 0 : cigar_offset[i], e = score_only ? 0 : cigar_offset[i + 1];
    const bool empty = b >= e;
    const int sc = (score_only || !empty) ? score[i] : -1;
    line = std::to_string(sc);
    line.push_back('\t');
    if (empty) {
      line.push_back('-');
    } else {
      for (int64_t r = b; r < e; ++r) {
        const uint32_t run = cigar_runs[r], op = run & 15u;
        line += std::to_string(run >> 4);
        line.push_back(op == BW_OP_MATCH ? 'M' : op == BW_OP_MISMATCH ? 'X' : op == BW_OP_INSERTION ? 'I' : 'D');
      }
    }
    line.push_back('\n');
    if (std::fwrite(line.data(), 1, line.size(), f) != line.size()) {
      std::fclose(f);
      set_error(std::string("write error on ") + path);
      return BW_STATUS_INVALID_ARG;
    }
  }
  if (std::fclose(f) != 0) {
    set_error(std::string("write error on ") + path);
    return BW_STATUS_INVALID_ARG;
  }
  return 0;
}

}  // extern "C"
