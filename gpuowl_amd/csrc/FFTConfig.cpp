// FFT shape selection; see FFTConfig.h. Behaviour follows src/FFTConfig.cpp.
#include "FFTConfig.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace {

struct BpwRow {
  u32 w, m, h;
  double bpw[FFTShape::NUM_BPW];
};

const BpwRow BPW_TABLE[] = {
#include "bpw_table.inc"
};

struct CapRow {
  u32 w, m, h;
  double bpw;
};

const CapRow ENGINE_CAPS[] = {
#include "engine_bpw.inc"
};

// First matching row, as a std::map built from the table keeps the first of duplicate keys.
const BpwRow* findBpw(u32 w, u32 m, u32 h) {
  for (const BpwRow& r : BPW_TABLE) {
    if (r.w == w && r.m == m && r.h == h) { return &r; }
  }
  return nullptr;
}

std::vector<std::string> splitOn(const std::string& s, char delim) {
  std::vector<std::string> out;
  size_t start = 0;
  while (true) {
    size_t p = s.find(delim, start);
    out.push_back(s.substr(start, p == std::string::npos ? std::string::npos : p - start));
    if (p == std::string::npos) { break; }
    start = p + 1;
  }
  return out;
}

// "1K" -> 1024, "1.5M" -> 1572864, "15" -> 15 (src/FFTConfig.cpp:27-34).
u32 parseNum(const std::string& s) {
  if (s.empty()) { throw std::invalid_argument("empty number in FFT spec"); }
  char c = s.back();
  u32 mult = (c == 'k' || c == 'K') ? 1024 : (c == 'm' || c == 'M') ? 1024 * 1024 : 1;
  return u32(strtod(s.c_str(), nullptr) * mult);
}

}  // namespace

std::string numberK(u32 n) {
  const u32 K = 1024, M = K * K;
  if (n % M == 0) { return std::to_string(n / M) + 'M'; }
  char buf[64];
  if (n >= M && (n * u64(100)) % M == 0) {
    snprintf(buf, sizeof(buf), "%.2f", float(n) / M);
    return std::string(buf) + 'M';
  }
  if (n >= K) {
    snprintf(buf, sizeof(buf), "%g", float(n) / K);
    return std::string(buf) + 'K';
  }
  return std::to_string(n);
}

std::vector<FFTShape> FFTShape::allShapes(u32 from, u32 to) {
  std::vector<FFTShape> out;
  for (u32 w : {256u, 512u, 1024u, 4096u}) {
    for (u32 h : {256u, 512u, 1024u}) {
      if (w == 256 && h == 1024) { continue; }
      for (u32 m = 2; m <= 16; ++m) {
        u32 sz = w * h * m * 2;
        if (from <= sz && sz <= to) { out.emplace_back(w, m, h); }
      }
    }
  }
  std::sort(out.begin(), out.end(), [](const FFTShape& a, const FFTShape& b) {
    if (a.size() != b.size()) { return a.size() < b.size(); }
    if (a.width != b.width) {
      if (a.width == 1024 || b.width == 1024) { return a.width == 1024; }
      return a.width < b.width;
    }
    return a.height < b.height;
  });
  return out;
}

std::vector<FFTShape> FFTShape::multiSpec(const std::string& spec) {
  if (spec.empty()) { return allShapes(); }
  std::vector<FFTShape> out;
  for (const std::string& one : splitOn(spec, ',')) {
    auto parts = splitOn(one, ':');
    if (parts.size() == 3) {
      out.emplace_back(parseNum(parts[0]), parseNum(parts[1]), parseNum(parts[2]));
      continue;
    }
    if (parts.size() != 1) { throw std::invalid_argument("FFT spec '" + one + "'"); }
    auto range = splitOn(one, '-');
    if (range.size() > 2) { throw std::invalid_argument("FFT spec '" + one + "'"); }
    u32 from = parseNum(range[0]);
    u32 to = range.size() == 2 ? parseNum(range[1]) : from;
    auto shapes = allShapes(from, to);
    if (shapes.empty()) { throw std::invalid_argument("no FFT shape for '" + one + "'"); }
    out.insert(out.end(), shapes.begin(), shapes.end());
  }
  return out;
}

FFTShape::FFTShape(const std::string& spec) {
  auto v = splitOn(spec, ':');
  if (v.size() != 3) { throw std::invalid_argument("FFT shape '" + spec + "'"); }
  *this = FFTShape(parseNum(v[0]), parseNum(v[1]), parseNum(v[2]));
}

FFTShape::FFTShape(u32 w, u32 m, u32 h) : width{w}, middle{m}, height{h} {
  if (!w || !m || !h) { throw std::invalid_argument("zero FFT dimension"); }
  if (w == 1 && m == 1 && h == 1) { return; }  // uninitialised shape
  if (const BpwRow* r = findBpw(w, m, h)) {
    std::copy(r->bpw, r->bpw + NUM_BPW, bpw.begin());
    return;
  }
  if (height > width) {
    bpw = FFTShape{h, m, w}.bpw;
    return;
  }
  // Below the reference's smallest FFT (256K words) there is nothing to borrow from
  // (the reference's mapping would not terminate); such test shapes get maxBpw 0.
  if (size() < 256u * 2 * 256 * 2) { return; }
  // Not measured: borrow a measured shape of similar geometry, 0.05 bits lower
  // (the same mapping as src/FFTConfig.cpp:141-164).
  while (m < 9) { m *= 2; w /= 2; }
  while (w >= 4 * h) { w /= 2; h *= 2; }
  while (w < h || w < 256 || w == 2048) { w *= 2; h /= 2; }
  while (h < 256) { h *= 2; m /= 2; }
  bpw = FFTShape{w, m, h}.bpw;
  for (double& b : bpw) { b -= 0.05; }
}

double FFTShape::maxBpw() const { return *std::max_element(bpw.begin(), bpw.end()); }

double FFTShape::carry32BPW() const {
  // Empirical CARRY32 limit: 18.35 bits at 6.5M words, -0.5 bit per doubling, capped at 19.
  double b = 18.35 + 0.5 * (std::log2(13.0 * 1024 * 512) - std::log2(double(size())));
  return b > 19.0 ? 19.0 : b;
}

bool FFTShape::isFavoredShape() const {
  return width >= height && ((width == 4096 && height >= 512 && middle >= 9) || (width == 1024 && middle >= 5) ||
                             (width == 512 && middle >= 4) || (width == 256 && middle >= 1));
}

FFTConfig::FFTConfig(FFTShape shape, u32 variant, u32 carry) : shape{shape}, variant{variant}, carry{carry} {
  if (variant / 100 >= 3 || variant % 100 / 10 >= 2 || variant % 10 >= 3) {
    throw std::invalid_argument("FFT variant " + std::to_string(variant));
  }
}

FFTConfig::FFTConfig(const std::string& spec) {
  auto v = splitOn(spec, ':');
  if (v.size() == 1) {
    // A size spec. (The reference intends this but its parser falls through to
    // "FFT spec" for a single-part spec, src/FFTConfig.cpp:188-200.)
    *this = FFTConfig{FFTShape::multiSpec(spec).front(), LAST_VARIANT, CARRY_AUTO};
  } else if (v.size() == 3) {
    *this = FFTConfig{FFTShape{parseNum(v[0]), parseNum(v[1]), parseNum(v[2])}, LAST_VARIANT, CARRY_AUTO};
  } else if (v.size() == 4) {
    *this = FFTConfig{FFTShape{parseNum(v[0]), parseNum(v[1]), parseNum(v[2])}, parseNum(v[3]), CARRY_AUTO};
  } else if (v.size() == 5) {
    u32 c = parseNum(v[4]);
    if (c > 1) { throw std::invalid_argument("FFT spec carry '" + v[4] + "'"); }
    *this = FFTConfig{FFTShape{parseNum(v[0]), parseNum(v[1]), parseNum(v[2])}, parseNum(v[3]),
                      c == 0 ? u32(CARRY_32) : u32(CARRY_64)};
  } else {
    throw std::invalid_argument("FFT spec '" + spec + "'");
  }
}

std::string FFTConfig::spec() const {
  std::string s = shape.spec() + ":" + std::to_string(variant / 100) + std::to_string(variant % 100 / 10) +
                  std::to_string(variant % 10);
  return carry == CARRY_AUTO ? s : s + (carry == CARRY_32 ? ":0" : ":1");
}

double FFTShape::engineCap() const {
  for (const CapRow& r : ENGINE_CAPS) {
    if (r.w == width && r.m == middle && r.h == height) { return r.bpw; }
  }
  return HUGE_VAL;
}

double FFTConfig::maxBpw() const { return std::min(refMaxBpw(), shape.engineCap()); }

double FFTConfig::refMaxBpw() const {
  u32 vW = variant / 100, vM = variant % 100 / 10, vH = variant % 10;
  double b;
  if (vW == vH || (shape.width > 1024 && vW == 1 && vH == 0)) {
    b = shape.bpw[vM * 3 + vH];
  } else {
    b = (shape.bpw[vM * 3 + vW] + shape.bpw[vM * 3 + vH]) / 2.0;
  }
  return carry == CARRY_32 ? std::min(shape.carry32BPW(), b) : b;
}

FFTConfig FFTConfig::bestFit(u32 E, const std::string& spec, double overdrive, bool refLimits) {
  if (!spec.empty()) { return FFTConfig{spec}; }
  // the cheapest tuned FFT that handles E (the list is ordered by cost)
  if (!refLimits) {
    for (const FftTuneEntry& e : FftTune::get()) {
      if (e.fft.maxExp() * overdrive >= E) { return e.fft; }
    }
  }
  for (const FFTShape& shape : FFTShape::allShapes()) {
    for (u32 v : {101u, 202u}) {
      FFTConfig fft{shape, v, CARRY_AUTO};
      if ((refLimits ? fft.refMaxExp() : fft.maxExp()) * overdrive >= E) { return fft; }
    }
  }
  throw std::invalid_argument("no FFT for exponent " + std::to_string(E));
}

// ---------------------------------------------------------------- tune file
namespace {
std::mutex tuneMutex;
std::vector<FftTuneEntry>& tuneList() {
  static std::vector<FftTuneEntry> list;
  return list;
}
}  // namespace

bool FftTune::update(std::vector<FftTuneEntry>& list, const FftTuneEntry& e) {
  const u32 m = e.fft.maxExp();
  for (const FftTuneEntry& r : list) {
    if (r.cost <= e.cost && r.fft.maxExp() >= m) { return false; }  // e is no better anywhere
  }
  list.erase(std::remove_if(list.begin(), list.end(),
                            [&](const FftTuneEntry& r) { return r.cost > e.cost && r.fft.maxExp() <= m; }),
             list.end());
  auto at = std::upper_bound(list.begin(), list.end(), e.cost,
                             [](double c, const FftTuneEntry& r) { return c < r.cost; });
  list.insert(at, e);
  return true;
}

std::vector<FftTuneEntry> FftTune::read(const std::string& path) {
  std::ifstream in(path);
  if (!in) { throw std::invalid_argument("cannot read tune file " + path); }
  std::vector<FftTuneEntry> list;
  std::string line;
  while (std::getline(in, line)) {
    const auto hash = line.find('#');
    std::istringstream is(line.substr(0, hash));
    FftTuneEntry e;
    std::string spec;
    if (!(is >> e.cost)) { continue; }  // blank or comment-only line
    if (!(is >> spec)) { throw std::invalid_argument("bad tune line: " + line); }
    e.fft = FFTConfig{spec};
    if (!list.empty() && (e.cost < list.back().cost || e.fft.maxExp() <= list.back().fft.maxExp())) {
      throw std::invalid_argument("tune file not ordered by cost and maxExp at: " + line);
    }
    list.push_back(e);
  }
  return list;
}

void FftTune::write(const std::string& path, const std::vector<FftTuneEntry>& list) {
  const std::string tmp = path + ".new";
  {
    std::ofstream out(tmp);
    if (!out) { throw std::invalid_argument("cannot write tune file " + path); }
    for (const FftTuneEntry& e : list) {
      char buf[96];
      std::snprintf(buf, sizeof buf, "%6.1f %14s # %u\n", e.cost, e.fft.spec().c_str(), e.fft.maxExp());
      out << buf;
    }
    if (!out.flush()) { throw std::invalid_argument("write failed: " + tmp); }
  }
  if (std::rename(tmp.c_str(), path.c_str())) { throw std::invalid_argument("cannot rename " + tmp); }
}

void FftTune::set(std::vector<FftTuneEntry> list) {
  std::lock_guard<std::mutex> lock(tuneMutex);
  tuneList() = std::move(list);
}

std::vector<FftTuneEntry> FftTune::get() {
  std::lock_guard<std::mutex> lock(tuneMutex);
  return tuneList();
}
