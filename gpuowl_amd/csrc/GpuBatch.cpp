#include "GpuBatch.h"

#include "Tune.h"

#include <chrono>
#include <map>
#include <stdexcept>

namespace {
double now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}
Words wordsOf(u32 E, u32 v) {
  Words w(nWords(E));
  w[0] = v;
  return w;
}
}  // namespace

GpuBatch::GpuBatch(const std::vector<u32>& exponents, const std::string& spec, const GpuOptions& opt) {
  std::map<std::string, size_t> byShape;
  for (u32 E : exponents) {
    FFTConfig fft = FFTConfig::bestFit(E, spec);
    GpuOptions o = opt;
    TuneChoice c;
    // a split tuned for batches of this FFT length, when the caller chose nothing
    if (!o.width && !o.lineBlock && !o.generic && !o.colOneRegion && !o.colOnePair &&
        BatchTuneTable::get(fft.size(), &c)) {
      o.width = c.width;
      o.lineBlock = c.lineBlock;
      o.generic = c.generic;
      o.colOneRegion = c.colOneRegion;
      o.colOnePair = c.colOnePair;
    }
    jobs.push_back(std::make_unique<Gpu>(E, fft, o));
    Gpu& g = *jobs.back();
    auto [it, fresh] = byShape.emplace(g.shapeKey(), groups.size());
    if (fresh) {
      groups.emplace_back();
      groups.back().lead = &g;
    }
    groups[it->second].index.push_back(jobs.size() - 1);
  }
  for (Group& grp : groups) {
    std::vector<JobDesc> d;
    for (size_t i : grp.index) { d.push_back(jobs[i]->job()); }
    grp.descs = Buffer<JobDesc>(grp.lead->queue(), d);
  }
}

const GpuBatch::Group& GpuBatch::groupOf(size_t job) const {
  for (const Group& g : groups) {
    for (size_t i : g.index) {
      if (i == job) { return g; }
    }
  }
  throw std::invalid_argument("job index");
}

u64 GpuBatch::totalWords() const {
  u64 s = 0;
  for (auto& j : jobs) { s += j->N; }
  return s;
}

void GpuBatch::finish() {
  for (Group& g : groups) { g.lead->finish(); }
}

void GpuBatch::squareLoop(Gpu::Buf out, Gpu::Buf in, u32 from, u32 to, bool mul3Last, bool ll) {
  for (Group& g : groups) { g.lead->squareLoop(out, in, from, to, mul3Last, ll, g.launch()); }
}

void GpuBatch::modMul(Gpu::Buf ioA, Gpu::Buf inB, bool mul3) {
  for (Group& g : groups) { g.lead->modMul(ioA, inB, mul3, g.launch()); }
}

std::vector<bool> GpuBatch::doCheck(u32 blockSize) {
  squareLoop(Gpu::AUX, Gpu::CHECK, 0, blockSize, true, false);
  modMul(Gpu::CHECK, Gpu::DATA, false);
  finish();
  std::vector<bool> ok(size());
  for (size_t i = 0; i < size(); ++i) { ok[i] = jobs[i]->isEqual(Gpu::CHECK, Gpu::AUX); }
  return ok;
}

std::vector<u64> GpuBatch::res64(Gpu::Buf b) {
  finish();
  std::vector<u64> r(size());
  for (size_t i = 0; i < size(); ++i) { r[i] = jobs[i]->dataResidue(b); }
  return r;
}

std::vector<PRPResult> GpuBatch::prp(u32 iters, u32 B, u32 checkStep, std::vector<std::vector<u64>>* resAt,
                                     u32 resStep) {
  if (!B) { throw std::invalid_argument("block size"); }
  if (!iters) { throw std::invalid_argument("a batch PRP run needs at least one iteration"); }
  if (!checkStep) { checkStep = 10 * B; }
  if (checkStep % B) { throw std::invalid_argument("check step must be a multiple of the block size"); }
  const size_t n = size();
  std::vector<PRPResult> out(n);
  std::vector<Words> vData(n), vCheck(n);
  for (size_t i = 0; i < n; ++i) {
    vData[i] = wordsOf(jobs[i]->E, 3);
    vCheck[i] = wordsOf(jobs[i]->E, 1);
  }
  u32 vK = 0;
  u32 totalErrors = 0;
  if (resAt) { resAt->assign(n, {}); }
  for (Group& g : groups) {
    u32 cnt = g.lead->roeCount();
    for (size_t i : g.index) { jobs[i]->readROE(cnt); }
    g.lead->restartRoeWindow();
  }
  // A full test of one of the jobs (iters == E) runs on to the next block boundary past E,
  // where it is Gerbicz-checked before x_E (kept on the way) counts, as Gpu::isPrimePRP does
  // (src/Gpu.cpp:1446-1447, 1507); the other jobs report x_iters the same way.
  bool full = false;
  for (auto& j : jobs) { full = full || j->E == iters; }
  const u32 kEndEnd = full ? (iters + B - 1) / B * B : iters;
  std::vector<u64> finalRes(n);
  std::vector<char> finalNine(n, 0);
  double t0 = now();
  // round-off slots advance on each group's lead; every job of the group reads that many
  auto readRoe = [&](std::vector<PRPResult>& res) {
    for (Group& g : groups) {
      u32 cnt = g.lead->roeCount();
      for (size_t i : g.index) { res[i].maxRoe = std::max(res[i].maxRoe, jobs[i]->readROE(cnt).max); }
    }
  };
  while (true) {
    readRoe(out);
    for (size_t i = 0; i < n; ++i) {
      jobs[i]->writeIn(Gpu::DATA, vData[i]);
      jobs[i]->writeIn(Gpu::CHECK, vCheck[i]);
      if (resAt && resStep) { (*resAt)[i].resize(vK / resStep); }
    }
    u32 k = vK;
    // a verified state already holds x_vK in the check (src/Gpu.cpp:1466-1471, 1545)
    bool skip = vK > 0, failed = false;
    std::vector<bool> leadIn(groups.size(), true);
    while (k < kEndEnd) {
      if (skip) {
        skip = false;
      } else if (k % B == 0) {
        modMul(Gpu::CHECK, Gpu::DATA, false);
      }
      ++k;
      bool fault = false;
      for (auto& j : jobs) { fault = fault || j->faultAt() == k; }
      bool leadOut = (k % B == 0) || k == iters || k == kEndEnd || (resStep && k % resStep == 0) || fault;
      for (size_t gi = 0; gi < groups.size(); ++gi) {
        Group& g = groups[gi];
        bool lo = leadOut || g.lead->useLongCarry;
        g.lead->square(Gpu::DATA, Gpu::DATA, leadIn[gi], lo, false, false, g.launch());
        leadIn[gi] = lo;
      }
      if (fault) {
        finish();
        for (auto& j : jobs) {
          if (j->faultAt() == k) { j->injectFault(); }
        }
      }
      if (k == iters) {
        finalRes = res64(Gpu::DATA);
        for (size_t i = 0; i < n; ++i) {
          finalNine[i] = (jobs[i]->E == iters) && equals9(jobs[i]->readData(Gpu::DATA));
        }
      }
      if (resAt && resStep && k % resStep == 0 && k <= iters) {
        std::vector<u64> r = res64(Gpu::DATA);
        for (size_t i = 0; i < n; ++i) { (*resAt)[i].push_back(r[i]); }
      }
      if ((k % checkStep == 0) || (k == kEndEnd && k % B == 0)) {
        std::vector<bool> ok = doCheck(B);
        bool all = true;
        for (size_t i = 0; i < n; ++i) {
          if (ok[i]) {
            ++out[i].nChecks;
          } else {
            ++out[i].nErrors;
            all = false;
          }
        }
        if (!all) {
          if (++totalErrors > 3) { throw GpuError("too many Gerbicz errors in the batch"); }
          failed = true;
          break;
        }
        skip = true;
        readRoe(out);
        for (size_t i = 0; i < n; ++i) {
          vData[i] = jobs[i]->readData(Gpu::DATA);
          vCheck[i] = jobs[i]->readData(Gpu::CHECK);
        }
        vK = k;
      }
    }
    if (failed) { continue; }
    finish();
    readRoe(out);
    double secs = (now() - t0) / std::max<u32>(1, iters);
    for (size_t i = 0; i < n; ++i) {
      out[i].k = iters;
      out[i].res64 = finalRes[i];
      out[i].isPrime = finalNine[i];
      out[i].secsPerIt = secs;
    }
    return out;
  }
}

std::vector<LLResult> GpuBatch::ll(u32 iters, std::vector<std::vector<u64>>* resAt, u32 resStep) {
  const size_t n = size();
  if (resAt) { resAt->assign(n, {}); }
  for (size_t i = 0; i < n; ++i) { jobs[i]->writeIn(Gpu::DATA, wordsOf(jobs[i]->E, 4)); }
  double t0 = now();
  std::vector<bool> leadIn(groups.size(), true);
  for (u32 k = 1; k <= iters; ++k) {
    bool leadOut = (resStep && k % resStep == 0) || k == iters;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      Group& g = groups[gi];
      bool lo = leadOut || g.lead->useLongCarry;
      g.lead->square(Gpu::DATA, Gpu::DATA, leadIn[gi], lo, false, true, g.launch());
      leadIn[gi] = lo;
    }
    if (resAt && resStep && k % resStep == 0) {
      std::vector<u64> r = res64(Gpu::DATA);
      for (size_t i = 0; i < n; ++i) { (*resAt)[i].push_back(r[i]); }
    }
  }
  finish();
  double secs = (now() - t0) / std::max<u32>(1, iters);
  std::vector<LLResult> out(n);
  for (size_t i = 0; i < n; ++i) {
    Words w = jobs[i]->readData(Gpu::DATA);
    out[i].k = iters;
    out[i].res64 = ::res64(w);
    out[i].isPrime = (iters == jobs[i]->E - 2) && w.empty();
    out[i].secsPerIt = secs;
  }
  return out;
}

double GpuBatch::timeSquares(u32 iters, u32 warmup, bool useGraph) {
  for (auto& j : jobs) { j->writeIn(Gpu::DATA, wordsOf(j->E, 3)); }
  for (Group& g : groups) { g.lead->rowLeadInTo(Gpu::DATA, 0, g.launch()); }
  auto step = [&](Group& g) {
    if (g.lead->useLongCarry) {
      g.lead->square(Gpu::DATA, Gpu::DATA, true, true, false, false, g.launch());
    } else {
      g.lead->colSquare(g.launch());
      g.lead->rowFused(false, OWL_ROE_SIZE, g.launch());
    }
  };
  for (u32 i = 0; i < warmup; ++i) {
    for (Group& g : groups) { step(g); }
  }
  const u32 G = 10;
  std::vector<owl_graph> graphs(groups.size(), nullptr);
  if (useGraph) {
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      owlCheck(owl_hip_graph_begin(groups[gi].lead->queue()), "graph begin");
      for (u32 i = 0; i < G; ++i) { step(groups[gi]); }
      owlCheck(owl_hip_graph_end(groups[gi].lead->queue(), &graphs[gi]), "graph end");
    }
  }
  finish();
  double t0 = now();
  u32 done = 0;
  while (done < iters) {
    for (size_t gi = 0; gi < groups.size(); ++gi) {
      if (useGraph) {
        owlCheck(owl_hip_graph_launch(graphs[gi], groups[gi].lead->queue()), "graph launch");
      } else {
        step(groups[gi]);
      }
    }
    done += useGraph ? G : 1;
  }
  finish();
  double us = (now() - t0) * 1e6 / done;
  for (owl_graph g : graphs) { owl_hip_graph_destroy(g); }
  // leave every state consistent: one more squaring ending in a lead-out
  for (Group& g : groups) {
    if (!g.lead->useLongCarry) {
      g.lead->colSquare(g.launch());
      g.lead->rowCarryOut(Gpu::DATA, false, false, OWL_ROE_SIZE, g.launch());
      g.lead->carryAll(Gpu::DATA, g.launch());
    }
  }
  finish();
  return us;
}
