// textfmt.h -- the reference's text number formats, fast.
//
// History rows print every value as "%10.7E " (HistoryFile.C:301-312) and
// VTK files as C++ streams print doubles, i.e. "%g" (VtkInterface.C:78-278).
// std::to_chars with the matching format and precision yields the same
// correctly rounded digits as printf (checked over 10^7 values of every
// magnitude, subnormals included) at a fraction of its per-call cost; large
// blocks are formatted by several host threads.
#pragma once

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

namespace dnlg {

// "%10.7E": a finite value always takes 13 or 14 characters, so the width
// never pads; printf's INF / NAN spellings are kept for the rest.
inline char *put_e7(char *p, double v) {
  if (!std::isfinite(v)) return p + std::snprintf(p, 32, "%10.7E", v);
  const auto r = std::to_chars(p, p + 32, v, std::chars_format::scientific, 7);
  for (char *q = p; q < r.ptr; q++)
    if (*q == 'e') *q = 'E';
  return r.ptr;
}

// "%g"
inline char *put_g(char *p, double v) { return std::to_chars(p, p + 32, v, std::chars_format::general, 6).ptr; }

// Text of n values v[0..n), each formatted by put and followed by what
// sep(p, k) writes at p (k = 0..n-1, at most two characters; returns the
// new end); split over up to `threads` host threads when n is large.
template <typename Put, typename Sep>
std::string format_values(const double *v, int64_t n, Put put, Sep sep, int threads = 0) {
  constexpr int64_t kChunkMin = 1 << 16;
  if (threads <= 0) threads = (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n / kChunkMin));
  std::vector<std::string> parts(nt);
  auto work = [&](int t) {
    const int64_t a = n * t / nt, b = n * (t + 1) / nt;
    std::string &s = parts[t];
    s.resize((size_t)(b - a) * 34);
    char *p = &s[0];
    for (int64_t k = a; k < b; k++) p = sep(put(p, v[k]), k);
    s.resize(p - &s[0]);
  };
  if (nt == 1) {
    work(0);
  } else {
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
  }
  std::string out;
  size_t total = 0;
  for (auto &s : parts) total += s.size();
  out.reserve(total);
  for (auto &s : parts) out += s;
  return out;
}

// Rows of `width` values: "v v ... v \n" (HistoryFile::save) or
// "v v ... v\n" (space = false: VTK lines).
inline std::string format_rows(const double *v, int64_t n_rows, int64_t width, bool trailing_space, bool e7) {
  auto sep = [=](char *p, int64_t k) {
    const bool last = (k % width) == width - 1;
    if (!last || trailing_space) *p++ = ' ';
    if (last) *p++ = '\n';
    return p;
  };
  return e7 ? format_values(v, n_rows * width, put_e7, sep) : format_values(v, n_rows * width, put_g, sep);
}

}  // namespace dnlg
