// Board files of the solver command line: multi-threaded reading, sorting,
// de-duplication and writing, with the reference's reading rules
// (main.cu:49-67) kept exactly:
//   * the count is read like fscanf("%lld\n"), which also skips every
//     whitespace character after it;
//   * each board line is what fgets(buf, 32) returns: at most 31 bytes,
//     ending after a newline; a line that runs out before it starts means
//     "Failed to load boards" (exit 1);
//   * the first 16 bytes of a line are the base81 board.
// Decoding goes through a 256-entry table built from go2_board_from_text,
// so it cannot disagree with the library's decoder.
#pragma once
#include <go2.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace go2cli {

[[noreturn]] inline void die(int code, const std::string& msg) {
  fprintf(stderr, "error: %s\n", msg.c_str());
  exit(code);
}

inline bool board_less(const go2_board& a, const go2_board& b) {
  return a.player != b.player ? a.player < b.player : a.opponent < b.opponent;
}
inline bool board_eq(const go2_board& a, const go2_board& b) {
  return a.player == b.player && a.opponent == b.opponent;
}

inline unsigned worker_count(size_t n) {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 1;
  if (t > 16) t = 16;
  const size_t per = 1 << 16;  // small inputs stay on one thread
  if (n / per + 1 < t) t = (unsigned)(n / per + 1);
  return t;
}

// Runs f(k, begin, end) for the k-th of `t` contiguous ranges of [0, n).
template <class F>
inline void parallel_ranges(size_t n, unsigned t, F&& f) {
  if (t <= 1) {
    f(0u, (size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned k = 0; k < t; ++k)
    th.emplace_back([&, k] { f(k, n * k / t, n * (k + 1) / t); });
  for (auto& x : th) x.join();
}

// Per character: the player and opponent discs of its four squares, or
// invalid.
struct DecodeTable {
  uint8_t p[256], o[256], ok[256];
  DecodeTable() {
    char text[17];
    go2_board empty{0, 0};
    go2_board_to_text(&empty, text);
    for (int c = 0; c < 256; ++c) {
      text[0] = (char)c;
      go2_board b;
      ok[c] = c != 0 && go2_board_from_text(text, &b) == GO2_OK;
      p[c] = ok[c] ? (uint8_t)(b.player & 15) : 0;
      o[c] = ok[c] ? (uint8_t)(b.opponent & 15) : 0;
    }
  }
};

inline std::vector<char> read_file(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) die(2, std::string("cannot open input ") + path);
  std::vector<char> buf;
  char chunk[1 << 16];
  size_t got;
  while ((got = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
  fclose(f);
  return buf;
}

// Reads, validates, sorts and de-duplicates the boards of `path`.
inline std::vector<go2_board> load_boards(const char* path) {
  const std::vector<char> text = read_file(path);
  const char* s = text.data();
  const size_t len = text.size();
  size_t pos = 0;
  while (pos < len && isspace((unsigned char)s[pos])) ++pos;  // fscanf skips leading blanks
  size_t num_start = pos;
  if (pos < len && (s[pos] == '+' || s[pos] == '-')) ++pos;
  while (pos < len && isdigit((unsigned char)s[pos])) ++pos;
  const std::string num(s + num_start, s + pos);
  char* end = nullptr;
  const long long n = num.empty() ? -1 : strtoll(num.c_str(), &end, 10);
  if (num.empty() || !end || *end || n < 0) die(2, "input does not start with a board count");
  while (pos < len && isspace((unsigned char)s[pos])) ++pos;  // the "\n" of "%lld\n"

  // Line starts, the way fgets(buf, 32) cuts the stream.  Every line holds
  // at least one byte, so a count beyond the file's size fails here, before
  // anything of the count's size is allocated.
  if ((unsigned long long)n > len - pos) {
    fprintf(stderr, "Failed to load boards\n");
    exit(EXIT_FAILURE);
  }
  std::vector<size_t> start((size_t)n), size((size_t)n);
  for (long long i = 0; i < n; ++i) {
    if (pos >= len) {
      fprintf(stderr, "Failed to load boards\n");
      exit(EXIT_FAILURE);
    }
    const size_t room = std::min<size_t>(31, len - pos);
    const void* nl = memchr(s + pos, '\n', room);
    const size_t take = nl ? (size_t)((const char*)nl - (s + pos)) + 1 : room;
    start[(size_t)i] = pos;
    size[(size_t)i] = take;
    pos += take;
  }

  static const DecodeTable tab;
  std::vector<go2_board> v((size_t)n);
  const unsigned t = worker_count((size_t)n);
  std::vector<size_t> first_bad(t, (size_t)-1);
  std::vector<int> bad_kind(t, 0);
  parallel_ranges((size_t)n, t, [&](unsigned k, size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      const char* l = s + start[i];
      // the 16 board characters end before a newline or the line's end
      size_t m = 0;
      while (m < size[i] && l[m] != '\n' && l[m] != '\0') ++m;
      int kind = m < 16 ? 1 : 0;
      uint64_t P = 0, O = 0;
      for (int j = 0; j < 16 && !kind; ++j) {
        const unsigned char c = (unsigned char)l[j];
        if (!tab.ok[c]) kind = 2;
        P |= (uint64_t)tab.p[c] << (4 * j);
        O |= (uint64_t)tab.o[c] << (4 * j);
      }
      if (!kind && (P & O)) kind = 3;
      if (kind && first_bad[k] == (size_t)-1) {
        first_bad[k] = i;
        bad_kind[k] = kind;
      }
      v[i] = go2_board{P, O};
    }
  });
  for (unsigned k = 0; k < t; ++k) {
    if (first_bad[k] == (size_t)-1) continue;
    const std::string where = "line " + std::to_string(first_bad[k] + 2);
    if (bad_kind[k] == 1) die(2, where + " is shorter than 16 characters");
    if (bad_kind[k] == 2) die(2, where + ": board text is not base81");
    die(2, where + ": a square holds both colours");
  }

  // Sort: chunks in parallel, then pairwise merges; then de-duplicate.
  std::vector<size_t> cut(t + 1);
  for (unsigned k = 0; k <= t; ++k) cut[k] = v.size() * k / t;
  parallel_ranges(t, t, [&](unsigned, size_t lo, size_t hi) {
    for (size_t k = lo; k < hi; ++k) std::sort(v.begin() + cut[k], v.begin() + cut[k + 1], board_less);
  });
  for (size_t width = 1; width < t; width *= 2) {
    std::vector<std::thread> th;
    for (size_t k = 0; k + width < t; k += 2 * width) {
      const size_t a = cut[k], m = cut[k + width], b = cut[std::min<size_t>(k + 2 * width, t)];
      th.emplace_back([&v, a, m, b] {
        std::inplace_merge(v.begin() + a, v.begin() + m, v.begin() + b, board_less);
      });
    }
    for (auto& x : th) x.join();
  }
  v.erase(std::unique(v.begin(), v.end(), board_eq), v.end());
  return v;
}

// Writes "<base81> <score>" (or "<base81> <score> <move>") lines, formatted
// in parallel and written in order.
template <class MoveText>
inline void write_results(FILE* out, const std::vector<go2_board>& boards, const int8_t* scores,
                          const uint8_t* moves, MoveText&& move_text) {
  const size_t n = boards.size();
  const unsigned t = worker_count(n);
  std::vector<std::string> part(t);
  parallel_ranges(n, t, [&](unsigned k, size_t lo, size_t hi) {
    std::string& o = part[k];
    o.reserve((hi - lo) * 24);
    char buf[17];
    char num[8];
    for (size_t j = lo; j < hi; ++j) {
      go2_board_to_text(&boards[j], buf);
      o.append(buf, 16);
      o.push_back(' ');
      const int len = snprintf(num, sizeof(num), "%d", (int)scores[j]);
      o.append(num, (size_t)len);
      if (moves) {
        o.push_back(' ');
        o.append(move_text(moves[j]));
      }
      o.push_back('\n');
    }
  });
  for (const auto& o : part)
    if (!o.empty() && fwrite(o.data(), 1, o.size(), out) != o.size()) die(2, "cannot write the output");
}

}  // namespace go2cli
