#include "Tune.h"

#include "GpuBatch.h"

#include "Trig.h"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <tuple>

namespace {
std::mutex mtx;
std::map<u32, TuneChoice>& table() {
  static std::map<u32, TuneChoice> t;
  return t;
}
int flagsOf(const TuneChoice& c) { return (c.generic ? 2 : 0) | (c.colOneRegion ? 4 : 0) | (c.colOnePair ? 8 : 0); }
}  // namespace

void TuneTable::set(u32 N, const TuneChoice& c) {
  std::lock_guard<std::mutex> lock(mtx);
  table()[N] = c;
}

bool TuneTable::get(u32 N, TuneChoice* c) {
  std::lock_guard<std::mutex> lock(mtx);
  auto it = table().find(N);
  if (it == table().end()) { return false; }
  *c = it->second;
  return true;
}

void TuneTable::clear() {
  std::lock_guard<std::mutex> lock(mtx);
  table().clear();
}

void TuneTable::load(const std::string& path) {
  std::ifstream in(path);
  if (!in) { throw std::invalid_argument("cannot read tune file " + path); }
  std::string line;
  std::map<u32, TuneChoice> got;
  while (std::getline(in, line)) {
    auto hash = line.find('#');
    if (hash != std::string::npos) { line.resize(hash); }
    std::istringstream is(line);
    u32 N = 0;
    TuneChoice c;
    int flags = 0;
    if (!(is >> N)) { continue; }
    if (!(is >> c.width >> c.lineBlock >> flags >> c.us)) { throw std::invalid_argument("bad tune line: " + line); }
    c.generic = flags & 2;
    c.colOneRegion = flags & 4;
    c.colOnePair = flags & 8;
    got[N] = c;
  }
  std::lock_guard<std::mutex> lock(mtx);
  for (auto& [N, c] : got) { table()[N] = c; }
}

void TuneTable::save(const std::string& path) {
  std::lock_guard<std::mutex> lock(mtx);
  std::ofstream out(path);
  if (!out) { throw std::invalid_argument("cannot write tune file " + path); }
  out << "# N width lineBlock flags us_per_iter\n";
  for (auto& [N, c] : table()) {
    char buf[128];
    std::snprintf(buf, sizeof buf, "%u %u %u %d %.2f\n", N, c.width, c.lineBlock, flagsOf(c), c.us);
    out << buf;
  }
}

std::vector<TuneResult> tuneShape(u32 E, const std::string& spec, u32 iters, u32 warmup, int device) {
  FFTConfig fft = FFTConfig::bestFit(E, spec);
  const u32 ND = fft.size() / 2;
  std::vector<TuneChoice> cands;
  for (u32 w = 256; w <= 4096; w *= 2) {
    if (ND % w) { continue; }
    u32 h = ND / w;
    if (h < 64 || h > 4096 || fftPlan(w).empty() || fftPlan(h).empty()) { continue; }
    for (u32 r = 1; r <= 16 && r * w <= 4096; r *= 2) {
      if (h % r) { continue; }
      for (int v = 0; v < 4; ++v) {
        TuneChoice c;
        c.width = w;
        c.lineBlock = r;
        c.generic = v == 2;
        c.colOneRegion = v == 1;
        c.colOnePair = v == 3;
        cands.push_back(c);
      }
    }
  }
  std::vector<TuneResult> out;
  std::set<std::tuple<std::string, std::string, u32, u32, bool>> seen;
  for (const TuneChoice& c : cands) {
    GpuOptions opt;
    opt.device = device;
    opt.width = c.width;
    opt.lineBlock = c.lineBlock;
    opt.generic = c.generic;
    opt.colOneRegion = c.colOneRegion;
    opt.colOnePair = c.colOnePair;
    try {
      Gpu g(E, fft, opt);
      auto key = std::make_tuple(g.rowKernel, g.colKernel, g.W, g.R, g.useLongCarry);
      if (!seen.insert(key).second) { continue; }  // the same engine as an earlier candidate
      TuneResult r;
      r.choice = c;
      r.choice.us = g.timeSquares(iters, warmup, true);
      r.W = g.W, r.H = g.H, r.R = g.R, r.longCarry = g.useLongCarry;
      r.rowKernel = g.rowKernel, r.colKernel = g.colKernel;
      out.push_back(r);
    } catch (const std::invalid_argument&) {
      // not a valid engine for this exponent (plan, register budget, stability)
    }
  }
  std::sort(out.begin(), out.end(), [](const TuneResult& a, const TuneResult& b) { return a.choice.us < b.choice.us; });
  return out;
}

namespace {
std::mutex batchMtx;
std::map<u32, TuneChoice>& batchTable() {
  static std::map<u32, TuneChoice> t;
  return t;
}
}  // namespace

void BatchTuneTable::set(u32 N, const TuneChoice& c) {
  std::lock_guard<std::mutex> lock(batchMtx);
  batchTable()[N] = c;
}

bool BatchTuneTable::get(u32 N, TuneChoice* c) {
  std::lock_guard<std::mutex> lock(batchMtx);
  auto it = batchTable().find(N);
  if (it == batchTable().end()) { return false; }
  *c = it->second;
  return true;
}

void BatchTuneTable::clear() {
  std::lock_guard<std::mutex> lock(batchMtx);
  batchTable().clear();
}

std::vector<TuneResult> tuneBatch(const std::vector<u32>& exps, const std::string& spec, u32 iters, u32 warmup) {
  std::map<u32, std::vector<u32>> byN;
  for (u32 E : exps) { byN[FFTConfig::bestFit(E, spec).size()].push_back(E); }
  std::vector<TuneResult> all;
  for (auto& [N, es] : byN) {
    const u32 ND = N / 2;
    TuneChoice before;
    const bool had = BatchTuneTable::get(N, &before);
    std::vector<TuneResult> out;
    std::set<std::tuple<std::string, std::string, u32, u32, bool>> seen;
    for (u32 w = 256; w <= 4096; w *= 2) {
      if (ND % w || ND / w < 64 || ND / w > 4096 || fftPlan(w).empty() || fftPlan(ND / w).empty()) { continue; }
      TuneChoice c;
      c.width = w;
      BatchTuneTable::set(N, c);
      try {
        GpuBatch b(es, spec);
        const Gpu& g = *b.jobs.front();
        auto key = std::make_tuple(g.rowKernel, g.colKernel, g.W, g.R, g.useLongCarry);
        if (!seen.insert(key).second) { continue; }
        TuneResult r;
        r.choice = c;
        r.choice.us = b.timeSquares(iters, warmup, true);
        r.W = g.W, r.H = g.H, r.R = g.R, r.longCarry = g.useLongCarry;
        r.rowKernel = g.rowKernel, r.colKernel = g.colKernel;
        out.push_back(r);
      } catch (const std::invalid_argument&) {
        // not a valid engine for these exponents at this split
      } catch (...) {
        // any other failure (device memory, a kernel error): the table keeps what it had
        if (had) {
          BatchTuneTable::set(N, before);
        } else {
          std::lock_guard<std::mutex> lock(batchMtx);
          batchTable().erase(N);
        }
        throw;
      }
    }
    std::sort(out.begin(), out.end(), [](const TuneResult& a, const TuneResult& b) { return a.choice.us < b.choice.us; });
    if (!out.empty()) {
      BatchTuneTable::set(N, out.front().choice);
    } else if (had) {
      BatchTuneTable::set(N, before);
    } else {
      std::lock_guard<std::mutex> lock(batchMtx);
      batchTable().erase(N);
    }
    all.insert(all.end(), out.begin(), out.end());
  }
  return all;
}
