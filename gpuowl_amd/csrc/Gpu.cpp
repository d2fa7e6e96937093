// Engine for one exponent; see Gpu.h. Host-side behaviour follows src/Gpu.cpp.
#include "Gpu.h"

#include "Trig.h"
#include "TrigBufCache.h"
#include "Tune.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace {

u32 pow2AtLeast(u32 x, u32 lo, u32 hi) {
  u32 t = lo;
  while (t < x && t < hi) { t *= 2; }
  return t;
}

}  // namespace

Kernel::Kernel(const std::string& name) : name_{name} { owlCheck(owl_hip_kernel_get(name.c_str(), &k_), "kernel lookup"); }

void Kernel::operator()(owl_queue q, u32 gridX, u32 gridY, u32 block, u32 lds, std::initializer_list<void*> args) const {
  std::vector<void*> a(args);
  int rc = owl_hip_launch(q, k_, gridX, gridY, block, lds, a.data());
  if (rc) { throw GpuError("launch " + name_ + ": " + owl_last_error()); }
}

template<typename T> T* Gpu::alloc(size_t n) {
  owned_.push_back(std::make_unique<Buffer<char>>(std::max<size_t>(n * sizeof(T), 4)));
  return reinterpret_cast<T*>(owned_.back()->get());
}

template<typename T> T* Gpu::upload(const std::vector<T>& v) {
  T* p = alloc<T>(v.size());
  owlCheck(owl_hip_write(queue_, p, v.data(), v.size() * sizeof(T), 1), "upload");
  return p;
}

bool equals9(const Words& w) {
  if (w.empty() || w[0] != 9) { return false; }
  return std::all_of(w.begin() + 1, w.end(), [](u32 x) { return x == 0; });
}

Gpu::Gpu(u32 E, const FFTConfig& fftIn, const GpuOptions& optIn) : E{E}, fft{fftIn} {
  GpuOptions opt = optIn;
  owlCheck(owl_hip_set_device(opt.device), "set device");
  device = opt.device;
  W = fft.shape.width;
  H = fft.shape.middle * fft.shape.height;
  N = fft.size();
  ND = N / 2;
  if (W * H != ND) { throw std::invalid_argument("bad FFT geometry"); }
  // a tuned choice for this FFT length applies when the caller chose nothing
  TuneChoice tuned;
  if (!opt.width && !opt.lineBlock && !opt.generic && !opt.colOneRegion && !opt.colOnePair && TuneTable::get(N, &tuned)) {
    opt.width = tuned.width;
    opt.lineBlock = tuned.lineBlock;
    opt.generic = tuned.generic;
    opt.colOneRegion = tuned.colOneRegion;
    opt.colOnePair = tuned.colOnePair;
    tunedLineBlock_ = true;
  }
  if (opt.width) {
    if (ND % opt.width) { throw std::invalid_argument("width does not divide the FFT length"); }
    W = opt.width;
    H = ND / W;
  } else if (H > 4096) {
    // The column pass holds a column pair in LDS (H <= 4096). Taller reference shapes
    // (e.g. 256:9:512, H = 4608) are factored again as W' x H' with W' a larger power
    // of two: the FFT length N, which fixes the words, weights and results, stays the
    // reference's; only the split of the ND-point FFT between the passes changes.
    while (W < 1024 && H % 2 == 0 && H / 2 >= 256) { W *= 2, H /= 2; }
    while (H > 4096 && H % 2 == 0 && W < 4096) { W *= 2, H /= 2; }
  }
  if (H > 4096 || W > 4096 || H < 64) { throw std::invalid_argument("FFT shape " + fft.spec() + " not supported"); }
  double bpw = double(E) / N;
  if (bpw < FFTShape::MIN_BPW) { throw std::invalid_argument("FFT size too large for exponent"); }
  if (E % 32 == 0 || E < 64) { throw std::invalid_argument("exponent not supported"); }
  useLongCarry = opt.longCarry;

  // Line block R: the largest one with a shape-specialized row pass that leaves >= 128
  // groups; otherwise as many lines per group as LDS allows, keeping >= 256 groups, then
  // longer if needed for the fused carry to be stable. Stable means a block's carry-in dies
  // within its 2R-word chain: (2R - 1) * floor(E/N) >= log2(N)/2 + 4. Below that the
  // iteration runs with a full carry (lead-out + lead-in) every time, as the reference
  // does below 12 bits per word (src/Gpu.cpp:522). A tuned choice (Tune.h) overrides this.
  const u32 rMax = std::min<u32>(4096 / W, 16);
  auto fusedStable = [&](u32 r) { return (2.0 * r - 1) * (E / N) >= 0.5 * std::log2(double(N)) + 4; };
  auto haveKernel = [](const std::string& name) {
    owl_kernel k = nullptr;
    return owl_hip_kernel_get(name.c_str(), &k) == 0;
  };
  R = opt.lineBlock;
  if (R && tunedLineBlock_ && (!fusedStable(R) || H % R || R * W > 4096)) { R = 0; }  // tuned for another exponent
  if (!R && !opt.generic) {
    // the largest line block with a shape-specialized row pass that keeps >= 128 groups
    for (u32 r = rMax; r >= 1 && !R; r /= 2) {
      if (H % r == 0 && H / r >= 128 && fusedStable(r) && haveKernel("rowx" + std::to_string(W) + "r" + std::to_string(r))) {
        R = r;
      }
    }
  }
  if (!R) {
    R = rMax;
    while (R > 1 && H / R < 256) { R /= 2; }
    while (R < rMax && !fusedStable(R) && H % (2 * R) == 0) { R *= 2; }
  }
  if (!R || H % R || R * W > 4096) { throw std::invalid_argument("bad line block"); }
  nb = H / R;
  useLongCarry = useLongCarry || !fusedStable(R);

  // shape-specialized passes when compiled for this shape (and R fits their lane mapping)
  rowKernel = "rowx" + std::to_string(W) + "r" + std::to_string(R);
  bool rowFixed = !opt.generic && haveKernel(rowKernel);
  rowChunked = rowFixed && nb >= 8;
  // column pass: k_colx (both columns in LDS) unless the one-region k_coly is asked for
  // or k_colx's lane mapping does not fit ((H/16) % R == 0); then k_coly if compiled
  const std::string colY = "coly" + std::to_string(H), colX = "colx" + std::to_string(H);
  bool useX = !opt.generic && !opt.colOneRegion && haveKernel(colX) && (H / 16) % R == 0;
  bool useY = !opt.generic && !useX && haveKernel(colY);
  colKernel = useY ? colY : colX;
  bool colFixed = useY || useX;

  rowPlan = fftPlan(W);
  colPlan = fftPlan(H);
  if (rowPlan.size() > OWL_MAX_STAGES || colPlan.size() > OWL_MAX_STAGES || rowPlan.empty() || colPlan.empty()) {
    throw std::invalid_argument("no FFT plan for shape " + fft.spec());
  }
  rowT = pow2AtLeast(R * W / 16, 64, 512);
  while (rowT < 512 && !planFits(rowPlan, W, R, rowT, 16)) { rowT *= 2; }
  colT = pow2AtLeast(2 * H / 16, 64, 512);
  while (colT < 512 && !planFits(colPlan, H, 2, colT, 16)) { colT *= 2; }
  if (!planFits(rowPlan, W, R, rowT, 16) || !planFits(colPlan, H, 2, colT, 16)) {
    throw std::invalid_argument("FFT plan does not fit the register budget for " + fft.spec());
  }
  rowLds = R * W * 16;
  colLds = 2 * H * 16;
  if (rowFixed) {
    owl_kernel k = nullptr;
    owlCheck(owl_hip_kernel_get(rowKernel.c_str(), &k), "kernel");
    owlCheck(owl_hip_kernel_max_threads(k, &rowT), "kernel");
    // R regions of rowRegion<W, R>() (kfixed.h) + per-wave round-off + the rowTwSlots<W, rowT>()-entry
    // twiddle tables (W >= 512)
    rowLds = R * (W + W / 16 + ((R == 4 || R == 8) ? 6 : 4)) * 16 + 64 +
             (W >= 1024 ? (rowT + 240 + W / 16) * 16 : W >= 512 ? (132 + rowT) * 16 : 0);
  } else {
    rowKernel = "row" + std::to_string(rowT);
  }
  if (colFixed) {
    owl_kernel k = nullptr;
    owlCheck(owl_hip_kernel_get(colKernel.c_str(), &k), "kernel");
    owlCheck(owl_hip_kernel_max_threads(k, &colT), "kernel");
    // the column regions + a colTwSize<H>()-entry twiddle table (kfixed.h)
    const u32 twSize = 68 + H / 64 + 2 * (64 + H / 64 + 16);
    // the two-region pass adds exact root tables for H >= 2304 (colRootSlots<H>())
    const u32 rootSlots = (!useY && H >= 2304) ? 16 * 15 + (H == 4096 ? 0 : 16 * ((H / 256 - 1) | 1) + 240 + H / 16) : 0;
    // two colRegionX<H>() regions (k_colx) or one colRegion<H>() (k_coly)
    colLds = ((useY ? H + H / 16 + 4 : 2 * (H + 4)) + twSize + rootSlots) * 16;
    if (useX && !opt.colOnePair && haveKernel(colX + "m")) {
      int cus = 0, lds = 0;
      owlCheck(owl_hip_device_info(opt.device, nullptr, 0, nullptr, nullptr, &cus, &lds), "device info");
      colMulti = true;
      colLdsM = colLds + twSize * 16;  // the next pair's twiddle table
      slotsM = u32(std::max(cus, 1)) * std::max<u32>(1, u32(lds) / colLdsM);
    }
  } else {
    colKernel = "col" + std::to_string(colT);
  }

  owlCheck(owl_hip_queue_create(&queue_), "queue");

  // ---- tables: twiddles shared per shape (TrigBufCache), weights per exponent (long double, rounded once)
  const u32 step = stepOf(N, E);
  const u32 nHi = (H + 63) / 64;
  trig_ = TrigBufCache::get(W, H, queue_);
  std::vector<LineW> lineW(H);
  for (u32 l = 0; l < H; ++l) {
    for (u32 rep = 0; rep < 2; ++rep) {
      u32 e = u32(u64(step) * (2 * l + rep) % N);
      lineW[l].e[rep] = e;
      lineW[l].w[rep] = double(exp2l((long double)e / N));
      lineW[l].iw[rep] = double(exp2l(-(long double)e / N));
    }
  }
  std::vector<ColW> colW(W);
  for (u32 c = 0; c < W; ++c) {
    u32 e = u32(u64(step) * 2 * H * c % N);
    colW[c].e = e;
    colW[c].w = double(exp2l((long double)e / N));
    colW[c].iw = double(exp2l(-(long double)e / N) / ND);
    colW[c].pad = 0;
  }

  JobDesc& J = job_;
  J.E = E;
  J.N = N;
  J.ND = ND;
  J.W = W;
  J.H = H;
  J.R = R;
  J.nb = nb;
  J.bitsLo = E / N;
  J.rE = E % N;
  J.nHi = nHi;
  J.lgR = 0;
  while ((1u << J.lgR) < R) { ++J.lgR; }
  J.nRow = u32(rowPlan.size());
  J.nCol = u32(colPlan.size());
  for (u32 i = 0; i < OWL_MAX_STAGES; ++i) {
    J.rowPlan[i] = i < rowPlan.size() ? rowPlan[i] : 0;
    J.colPlan[i] = i < colPlan.size() ? colPlan[i] : 0;
  }
  J.twRow = trig_->row.get();
  J.twCol = trig_->col.get();
  J.twMidA = trig_->midA.get();
  J.twMidB = trig_->midB.get();
  J.twPair = trig_->pair.get();
  J.twMidS = trig_->midS.get();
  J.lineW = upload(lineW);
  J.colW = upload(colW);
  J.rowMeta = nullptr;
  if (rowFixed) {
    // Per row-pass thread and block: which of its words are big and which weight exponents wrap
    // (the tests of wordWeight, src/Gpu.cpp:40-127's bitsCF), so the pass does not recompute them.
    const u32 T = rowT, F0 = W / T;
    // thread tid's columns are colBase(tid) + k*T (RowPass::colBase, kfixed.h)
    auto colBase = [F0](u32 tid) -> u32 {
      const u32 b = tid & 15, hi = tid & ~15u;
      if (F0 == 2) { return hi | (b & 3) | ((b >> 3) << 2) | (((b >> 2) & 1) << 3); }
      if (F0 == 4) { return hi | (b & 1) | (((b >> 3) & 1) << 1) | (((b >> 1) & 3) << 2); }
      if (F0 == 8) { return hi | ((b >> 3) & 1) | ((b & 7) << 1); }
      return tid;
    };
    std::vector<u64> meta(size_t(nb) * T, 0);
    for (u32 lb = 0; lb < nb; ++lb) {
      for (u32 tid = 0; tid < T; ++tid) {
        u64 m = 0;
        for (u32 k = 0; k < F0; ++k) {
          for (u32 r = 0; r < R; ++r) {
            for (u32 rep = 0; rep < 2; ++rep) {
              const u32 w = (k * R + r) * 2 + rep;
              const u32 t = lineW[lb * R + r].e[rep] + colW[colBase(tid) + k * T].e;
              const bool wrap = t >= N;
              const bool big = (wrap ? t - N : t) < J.rE;
              m |= (u64(big) << w) | (u64(wrap) << (32 + w));
            }
          }
        }
        meta[size_t(lb) * T + tid] = m;
      }
    }
    J.rowMeta = upload(meta);
  }
  for (int i = 0; i < 3; ++i) { J.freq[i] = alloc<double2>(ND); }
  for (int i = 0; i < 4; ++i) { J.words[i] = alloc<int2>(ND); }
  for (int i = 0; i < 2; ++i) { J.carry[i] = alloc<int64_t>(size_t(nb) * W); }
  J.stair = alloc<uint64_t>(size_t(nb) * W);
  J.ticket = alloc<uint64_t>(1);
  J.roe = alloc<uint32_t>(OWL_ROE_SIZE);
  J.flags = alloc<uint32_t>(8);
  const size_t nMap = (size_t(nb) * W + 63) / 64, nSum = (nMap + 63) / 64;
  J.fixMap = alloc<uint64_t>(nMap + nSum);
  J.small = alloc<int32_t>(64);
  J.sum = alloc<uint64_t>(1);
  owlCheck(owl_hip_fill32(queue_, J.stair, 0, 2 * nb * W), "fill");
  owlCheck(owl_hip_fill32(queue_, J.ticket, 0, 2), "fill");
  owlCheck(owl_hip_fill32(queue_, J.roe, 0, OWL_ROE_SIZE), "fill");
  owlCheck(owl_hip_fill32(queue_, J.flags, 0, 8), "fill");
  owlCheck(owl_hip_fill32(queue_, J.fixMap, 0, 2 * (nMap + nSum)), "fill");
  for (int i = 0; i < 4; ++i) { owlCheck(owl_hip_fill32(queue_, J.words[i], 0, 2 * ND), "fill"); }
  jobs_ = Buffer<JobDesc>(1);
  jobs_.write(queue_, {J});

  kRow = Kernel(rowKernel);
  colOne_ = colKernel;
  kCol = Kernel(colOne_);
  if (colMulti) { kColM = Kernel(colOne_ + "m"); }
  colKernel = colKernelFor(1);
  kCarryB = Kernel("carryB");
  kCarryFix = Kernel("carryFix");
  kWordsIn = Kernel("wordsIn");
  kWordsOut = Kernel("wordsOut");
  kResidue = Kernel("readResidue");
  kIsEqual = Kernel("isEqual");
  kSum64 = Kernel("sum64");
  kCopy = Kernel("copyWords");
  finish();
}

Gpu::~Gpu() {
  if (queue_) {
    owl_hip_finish(queue_);
    owned_.clear();
    trig_.reset();
    jobs_ = Buffer<JobDesc>();
    owl_hip_queue_destroy(queue_);
  }
}

void Gpu::finish() { owlCheck(owl_hip_finish(queue_), "finish"); }

u32 Gpu::nextRoeSlot() {
  if (roeTaken_ >= roeWindow_) { return OWL_ROE_SIZE; }  // past the measured window
  ++roeTaken_;
  u32 s = std::min<u32>(roePos, OWL_ROE_SIZE - 1);
  ++roePos;
  return s;
}

u32 Gpu::nextStairTag() {
  do {
    stairTag_ = stairTag_ % 0xFFFE + 1;
  } while (stairTag_ == lastTag_ || (stairTag_ >= resLo_ && stairTag_ <= resHi_));
  return lastTag_ = stairTag_;
}

owl_graph Gpu::captureFused(u32 G, bool ll, u32 roeSlot, Launch L, u32* lastTagOut) {
  if (G < 2) { throw std::invalid_argument("captureFused: at least 2 squarings"); }
  const u32 before = lastTag_;
  resLo_ = 1, resHi_ = 0;
  if (stairTag_ + G + 2 > 0xFFFE) { stairTag_ = 0; }  // keep the captured range unwrapped
  owl_graph g = nullptr;
  owlCheck(owl_hip_graph_begin(L.q), "graph begin");
  u32 first = 0;
  for (u32 i = 0; i < G; ++i) {
    colSquare(L);
    rowFused(ll, roeSlot, L);
    if (!i) { first = lastTag_; }
  }
  owlCheck(owl_hip_graph_end(L.q, &g), "graph end");
  // the graph's tags are contiguous (up to one step over `before`): no later direct tag
  // may fall in [first, last]
  u32 last = lastTag_;
  resLo_ = first, resHi_ = last;
  lastTag_ = before;
  *lastTagOut = last;
  return g;
}

void Gpu::residueOn(owl_queue q, Buf b) {
  const JobDesc* j = jobs_.get();
  int wi = b;
  kResidue(q, 1, 1, 64, 0, {&j, &wi});
}

u64 Gpu::residueTaken() {
  int w[64];
  owlCheck(owl_hip_read(queue_, w, job_.small, sizeof(w), 1), "read residue");
  return residueFromEnds(w, N, E);
}

void Gpu::copyOn(owl_queue q, Buf dst, Buf src) {
  const JobDesc* j = jobs_.get();
  int d = dst, s = src;
  kCopy(q, 256, 1, 256, 0, {&j, &d, &s});
}

u64 Gpu::deviceBytes() const {
  u64 s = sizeof(JobDesc);
  for (auto& b : owned_) { s += b->size(); }
  return s;
}

u64 Gpu::footprintBound(u32 N, u32 W, u32 H) {
  const u64 ND = N / 2;
  // 3 spectra (16 B) + 4 word buffers (8 B) per complex point; carries and stairs
  // (3 x 8 B per column and block, at most ND); the row pass's word metadata (8 B per thread and
  // block, at most ND bytes); the carry fix map (one bit per column and block, plus a bit per map
  // word); weights; round-off, flags, residue
  return ND * (3 * 16 + 4 * 8) + 3 * 8 * ND + ND + ND / 8 + ND / 256 + 16 + u64(H) * sizeof(LineW) + u64(W) * sizeof(ColW) +
         4 * OWL_ROE_SIZE + 4096;
}

// ---------------------------------------------------------------- launches
std::string Gpu::shapeKey() const {
  return rowKernel + "/" + colKernel + "/" + std::to_string(W) + "x" + std::to_string(H) + "/R" + std::to_string(R) +
         (useLongCarry ? "/long" : "");
}

void Gpu::rowFused(bool ll, u32 roeSlot, Launch L) {
  // k_rowx: eight chunks of blocks, one per XCD, ceil(nb/8) + 1 groups each (kfixed.h, fusedItem);
  // the runtime-plan k_row: nb + 1 groups in start-ticket order
  int mode = ROW_FUSED, fi = 0, wi = 0, opts = (ll ? 1 : 0) | (rowChunked ? 4 : 0);
  u32 tag = nextStairTag();
  kRow(L.q, rowChunked ? 8 * ((nb + 7) / 8 + 1) : nb + 1, L.n, rowT, rowLds,
       {&L.jobs, &mode, &fi, &wi, &opts, &roeSlot, &tag});
}

void Gpu::rowLeadInTo(Buf in, int freqIdx, Launch L) {
  int mode = ROW_LEADIN, fi = freqIdx, wi = in, opts = 0;
  u32 slot = OWL_ROE_SIZE, tag = 0;
  kRow(L.q, nb, L.n, rowT, rowLds, {&L.jobs, &mode, &fi, &wi, &opts, &slot, &tag});
}

void Gpu::rowCarryOut(Buf out, bool ll, bool mul3, u32 roeSlot, Launch L) {
  int mode = ROW_CARRYOUT, fi = 0, wi = out, opts = (ll ? 1 : 0) | (mul3 ? 2 : 0);
  u32 tag = 0;
  kRow(L.q, nb, L.n, rowT, rowLds, {&L.jobs, &mode, &fi, &wi, &opts, &roeSlot, &tag});
}

void Gpu::carryAll(Buf b, Launch L) {
  int wi = b, ci = 0, co = 1;
  kCarryB(L.q, (nb * W + 255) / 256, L.n, 256, 0, {&L.jobs, &wi, &ci, &co});
  kCarryFix(L.q, 1, L.n, 256, 0, {&L.jobs, &wi, &co});
}

// Pairs per column group for a launch of nJobs jobs: the largest power of two up to
// kMaxPairs whose grid (W/2 / pairs groups per job) still gives each of the device's
// slotsM group slots a group; 1 when the multi-pair pass is not available. (At 7.5M words
// two pairs per group take the column pass from 60.2 to 55.3 us; in 64-job batches of
// 1.5M words, 2, 4, 8 or 16 pairs were within 2% of one pair and not kept for H < 2304.)
u32 Gpu::colPairsPerGroup(u32 nJobs) const {
  constexpr u32 kMaxPairs = 2;
  u32 best = 1;
  if (!colMulti) { return best; }
  for (u32 p = 2; p <= kMaxPairs && p <= W / 2; p *= 2) {
    if (u64(W / 2 / p) * nJobs < slotsM) { break; }
    best = p;
  }
  return best;
}

std::string Gpu::colKernelFor(u32 nJobs) const { return colPairsPerGroup(nJobs) > 1 ? colOne_ + "m" : colOne_; }

void Gpu::colPass(int mode, int fi, int oi, Launch L) {
  const u32 p = colPairsPerGroup(L.n);
  if (p > 1) {
    kColM(L.q, W / 2 / p, L.n, colT, colLdsM, {&L.jobs, &mode, &fi, &oi});
  } else {
    kCol(L.q, W / 2, L.n, colT, colLds, {&L.jobs, &mode, &fi, &oi});
  }
}

void Gpu::colSquare(Launch L) { colPass(COL_SQUARE, 0, 0, L); }
void Gpu::colMul(Launch L) { colPass(COL_MUL, 0, 2, L); }
void Gpu::colFwdSpec(Launch L) { colPass(COL_FWD, 2, 2, L); }

// ---------------------------------------------------------------- squaring
void Gpu::square(Buf out, Buf in, bool leadIn, bool leadOut, bool doMul3, bool doLL) {
  square(out, in, leadIn, leadOut, doMul3, doLL, self());
}

void Gpu::square(Buf out, Buf in, bool leadIn, bool leadOut, bool doMul3, bool doLL, Launch L) {
  if (leadIn) { rowLeadInTo(in, 0, L); }
  colSquare(L);
  if (leadOut) {
    rowCarryOut(out, doLL, doMul3, nextRoeSlot(), L);
    carryAll(out, L);
  } else {
    rowFused(doLL, nextRoeSlot(), L);
  }
}

u32 Gpu::squareLoop(Buf out, Buf in, u32 from, u32 to, bool doTailMul3, bool doLL) {
  return squareLoop(out, in, from, to, doTailMul3, doLL, self());
}

u32 Gpu::squareLoop(Buf out, Buf in, u32 from, u32 to, bool doTailMul3, bool doLL, Launch L) {
  bool leadIn = true;
  for (u32 k = from; k < to; ++k) {
    bool leadOut = useLongCarry || k == to - 1;
    square(out, k == from ? in : out, leadIn, leadOut, doTailMul3 && k == to - 1, doLL, L);
    leadIn = leadOut;
  }
  return to;
}

void Gpu::modMul(Buf ioA, Buf inB, bool mul3) { modMul(ioA, inB, mul3, self()); }

void Gpu::modMul(Buf ioA, Buf inB, bool mul3, Launch L) {
  rowLeadInTo(inB, 2, L);
  colFwdSpec(L);
  rowLeadInTo(ioA, 0, L);
  colMul(L);
  rowCarryOut(ioA, false, mul3, nextRoeSlot(), L);
  carryAll(ioA, L);
}

void Gpu::injectFault() {
  faultAt_ = 0;
  Words w = readData(DATA);
  w[w.size() / 2] ^= 1u << 7;
  writeIn(DATA, w);
}

bool Gpu::doCheck(u32 blockSize) {
  squareLoop(AUX, CHECK, 0, blockSize, true, false);
  modMul(CHECK, DATA);
  return isEqual(CHECK, AUX);
}

u32 Gpu::flagsRead() {
  std::vector<u32> f(8);
  owlCheck(owl_hip_read(queue_, f.data(), job_.flags, 8 * 4, 1), "read flags");
  return (f[0] ? 1 : 0) | (f[1] ? 2 : 0) | (f[2] ? 4 : 0);
}

bool Gpu::isEqual(Buf a, Buf b) {
  owlCheck(owl_hip_fill32(queue_, job_.flags + 1, 0, 1), "fill");
  const JobDesc* j = jobs_.get();
  int wa = a, wb = b;
  kIsEqual(queue_, 256, 1, 256, 0, {&j, &wa, &wb});
  u32 f = flagsRead();
  if (f & 4) { throw GpuError("carry hand-off timed out"); }
  return !(f & 2);
}

void Gpu::copy(Buf dst, Buf src) {
  const JobDesc* j = jobs_.get();
  int d = dst, s = src;
  kCopy(queue_, 256, 1, 256, 0, {&j, &d, &s});
}

// ---------------------------------------------------------------- state I/O
void Gpu::writeDigits(Buf b, const std::vector<int>& digits) {
  if (digits.size() != N) { throw std::invalid_argument("writeDigits: size"); }
  Buffer<int32_t> seq(N);
  owlCheck(owl_hip_write(queue_, seq.get(), digits.data(), N * 4, 0), "write");
  const JobDesc* j = jobs_.get();
  int wi = b;
  int32_t* p = seq.get();
  kWordsIn(queue_, (ND + 255) / 256, 1, 256, 0, {&j, &wi, &p});
  finish();
}

void Gpu::writeIn(Buf b, const Words& words) { writeDigits(b, expandBits(words, N, E)); }

std::vector<int> Gpu::readDigits(Buf b) {
  Buffer<int32_t> seq(N);
  std::vector<int> out(N);
  readDigitsVia(b, seq.get(), out.data());
  return out;
}

void Gpu::readDigitsNoAlloc(Buf b, int32_t* out) {
  int32_t* staging = nullptr;
  if (N > OWL_NOALLOC_MAX_WORDS) {
    throw std::invalid_argument("read_digits_noalloc: N = " + std::to_string(N) + " words exceeds the static staging "
                                "area of " + std::to_string(OWL_NOALLOC_MAX_WORDS) + " words");
  }
  if (owlNoAllocScratch(device, &staging)) { throw GpuError("read_digits_noalloc: no address for the staging area"); }
  readDigitsVia(b, staging, out);
}

// k_words_out stages the N digits at `staging` (device), k_sum64 checksums the words buffer; the
// digits are copied to `out` (host) and accepted when their sum matches (src/Gpu.cpp:677-705).
void Gpu::readDigitsVia(Buf b, int32_t* staging, int32_t* out) {
  const JobDesc* j = jobs_.get();
  int wi = b;
  for (int attempt = 0; attempt < 3; ++attempt) {
    owlCheck(owl_hip_fill32(queue_, job_.sum, 0, 2), "fill");
    kSum64(queue_, 256, 1, 256, 0, {&j, &wi});
    kWordsOut(queue_, (ND + 255) / 256, 1, 256, 0, {&j, &wi, &staging});
    u64 gpuSum = 0;
    owlCheck(owl_hip_read(queue_, &gpuSum, job_.sum, 8, 0), "read");
    owlCheck(owl_hip_read(queue_, out, staging, size_t(N) * 4, 1), "read");
    u64 hostSum = 0;
    for (u32 k = 0; k < N; k += 2) { hostSum += u32(out[k]) | (u64(u32(out[k + 1])) << 32); }
    if (hostSum == gpuSum) { return; }
  }
  throw GpuError("persistent GPU read errors");
}

Words Gpu::readData(Buf b) {
  u32 f = flagsRead();
  if (f & 4) { throw GpuError("carry hand-off timed out"); }
  std::vector<int> d = readDigits(b);
  if (std::all_of(d.begin(), d.end(), [](int x) { return x == 0; })) { return {}; }
  return compactBits(d, E);
}

u64 Gpu::dataResidue(Buf b) {
  const JobDesc* j = jobs_.get();
  int wi = b;
  kResidue(queue_, 1, 1, 64, 0, {&j, &wi});
  int w[64];
  owlCheck(owl_hip_read(queue_, w, job_.small, sizeof(w), 1), "read residue");
  u32 f = flagsRead();
  if (f & 4) { throw GpuError("carry hand-off timed out"); }
  return residueFromEnds(w, N, E);
}

RoeInfo Gpu::readROE(u32 nSlots) {
  u32 n = std::min<u32>(nSlots, OWL_ROE_SIZE);
  RoeInfo r;
  if (!n) { return r; }
  std::vector<uint32_t> bits(n);
  owlCheck(owl_hip_read(queue_, bits.data(), job_.roe, n * 4, 1), "read roe");
  owlCheck(owl_hip_fill32(queue_, job_.roe, 0, n), "fill");
  roePos = 0;
  double sum = 0, sum2 = 0;
  for (uint32_t b : bits) {
    float x;
    memcpy(&x, &b, 4);
    r.max = std::max<double>(r.max, x);
    sum += x;
    sum2 += double(x) * x;
  }
  r.N = n;
  r.mean = sum / n;
  r.sd = std::sqrt(std::max(0.0, n * sum2 - sum * sum)) / n;
  return r;
}

// ---------------------------------------------------------------- workflows
namespace {
double now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}
Words makeWords(u32 E, u32 v) {
  Words w(nWords(E));
  w[0] = v;
  return w;
}
}  // namespace

void Gpu::writeState(const Words& check, u32 B) {
  if (B < 2) { throw std::invalid_argument("block size"); }
  // c^(2^n - 1) for n = 1, 2, 4, ... while 2n divides B with B/(2n) >= 2, then B/n - 1
  // steps of y <- y^(2^n) * c^(2^n - 1), the last one times 3
  writeIn(CHECK, check);
  copy(DATA, CHECK);
  copy(AUX, CHECK);
  u32 n = 1;
  for (; B % (2 * n) == 0 && B / (2 * n) >= 2; n *= 2) {
    squareLoop(DATA, 0, n);
    modMul(DATA, AUX);
    copy(AUX, DATA);
  }
  const u32 q = B / n;
  for (u32 i = 1; i < q; ++i) {
    squareLoop(DATA, 0, n);
    modMul(DATA, AUX, i == q - 1);
  }
}

PRPResult Gpu::isPrimePRP(u32 iters, u32 B, u32 checkStep, std::vector<u64>* resAt, u32 resStep,
                          Saver<PRPState>* saver, const ProofStore* proof) {
  if (!B) { throw std::invalid_argument("block size"); }
  PRPResult out;
  Words vData = makeWords(E, 3), vCheck = makeWords(E, 1);
  u32 vK = 0;
  // a state verified in this run holds x_vK in its check already (src/Gpu.cpp:1466-1471,
  // 1545); a savefile's check does not
  bool vCheckHasX = false;
  double elapsedBefore = 0;
  if (saver) {
    for (int attempt = 0;; ++attempt) {
      PRPState s = saver->load();
      if (s.k == 0) { break; }  // no savefile: the start state
      if (s.blockSize != B && (!checkStep || checkStep % s.blockSize)) { checkStep = 0; }
      B = s.blockSize;
      writeState(s.check, B);
      if (dataResidue() == s.res64) {
        vData = readData(DATA);
        vCheck = s.check;
        vK = s.k;
        out.nErrors = s.nErrors;
        elapsedBefore = s.elapsed;
        break;
      }
      if (attempt) { throw GpuError("savefile res64 does not match its check"); }
      saver->dropMostRecent();  // try an earlier savefile (src/Gpu.cpp:1169-1191)
    }
  }
  if (!checkStep) { checkStep = 10 * B; }
  if (checkStep % B) { throw std::invalid_argument("check step must be a multiple of the block size"); }
  const u32 nErrorsBefore = out.nErrors;
  // a full test runs on to the next block boundary past E, where it is checked before its
  // result (x_E, kept on the way) counts (src/Gpu.cpp:1446-1447, 1507)
  const u32 kEndEnd = (iters == E) ? (E + B - 1) / B * B : iters;
  u64 finalRes = 0;
  Words finalWords;
  bool haveFinal = false;  // x_iters was reached in this run (not when a savefile is past it)
  double t0 = now();
  readROE();
  restartRoeWindow();
  double maxRoe = 0;
  while (true) {
    writeIn(DATA, vData);
    writeIn(CHECK, vCheck);
    u32 k = vK;
    if (resAt && resStep) { resAt->resize(vK / resStep); }
    bool skip = vK > 0 && vCheckHasX, leadIn = true, failed = false;
    u32 proofK = proof ? proof->next(k) : ~0u;  // the next iteration whose residue the proof keeps
    while (k < kEndEnd) {
      if (skip) {
        skip = false;
      } else if (k % B == 0) {
        modMul(CHECK, DATA);
      }
      ++k;
      bool fault = (k == faultAt_);
      bool leadOut = (k % B == 0) || k == iters || (resStep && k % resStep == 0) || useLongCarry || fault ||
                     k == proofK;
      square(DATA, DATA, leadIn, leadOut);
      leadIn = leadOut;
      if (fault) { injectFault(); }
      if (k == proofK) {  // kept at once, as the reference does; a reload past it writes it again
        proof->save(k, readData(DATA));
        proofK = proof->next(k);
      }
      if (k == iters) {
        haveFinal = true;
        finalRes = dataResidue();
        if (iters == E) { finalWords = readData(DATA); }
      }
      if (resAt && resStep && k % resStep == 0 && k <= iters) { resAt->push_back(dataResidue()); }
      bool check = (k % checkStep == 0) || (k == kEndEnd && k % B == 0);
      if (check) {
        maxRoe = std::max(maxRoe, readROE().max);
        Words saved;
        if (saver) { saved = readData(CHECK); }  // the check before it takes x_k
        if (this->doCheck(B)) {
          ++out.nChecks;
          skip = true;
          vData = readData(DATA);
          vCheck = readData(CHECK);
          vCheckHasX = true;
          vK = k;
          if (saver && k < E) {
            saved.resize(nWords(E), 0);
            saver->save({E, k, B, res64(vData), saved, out.nErrors, elapsedBefore + (now() - t0)});
          }
        } else {
          ++out.nErrors;
          if (out.nErrors - nErrorsBefore > 3) { throw GpuError("too many Gerbicz errors"); }
          failed = true;
          break;
        }
      }
    }
    if (failed) { continue; }
    out.k = haveFinal ? iters : k;
    break;
  }
  maxRoe = std::max(maxRoe, readROE().max);
  out.res64 = haveFinal ? finalRes : dataResidue();
  if (out.k == E) { out.isPrime = equals9(finalWords); }
  out.maxRoe = maxRoe;
  out.secsPerIt = (now() - t0) / std::max<u32>(1, iters);
  return out;
}

Words Gpu::expMul(const Words& A, u64 h, const Words& Bw, bool squareB) {
  writeIn(CHECK, Bw);
  if (squareB) { squareLoop(CHECK, 0, 1); }
  if (h == 0) {
    writeIn(DATA, makeWords(E, 1));
  } else {
    // left-to-right binary: square for every bit below the top one, multiply by A where it is set
    writeIn(AUX, A);
    copy(DATA, AUX);
    u32 run = 0;
    for (int b = 62 - __builtin_clzll(h) + 1; b-- > 0;) {
      ++run;
      if ((h >> b) & 1) {
        squareLoop(DATA, 0, run);
        modMul(DATA, AUX);
        run = 0;
      }
    }
    if (run) { squareLoop(DATA, 0, run); }
  }
  modMul(DATA, CHECK);
  Words r = readData(DATA);
  r.resize(nWords(E), 0);
  return r;
}

Words Gpu::expExp2(const Words& A, u32 n) {
  writeIn(DATA, A);
  if (n) { squareLoop(DATA, 0, n); }
  Words r = readData(DATA);
  r.resize(nWords(E), 0);
  return r;
}

PrpProof Gpu::makeProof(const ProofStore& st) {
  const std::vector<u32>& pts = st.points();
  const u32 power = st.power();
  PrpProof pr;
  pr.E = E;
  pr.B = st.load(E);
  auto hash = proofHash(E, pr.B);
  std::vector<u64> hashes;
  for (u32 p = 0; p < power; ++p) {
    // M_p = prod over the 2^p points at odd multiples of the level's span, each raised to
    // the hashes of the levels that split it (ProofSet::computeProof, src/Proof.cpp:245-297)
    std::vector<Words> stack;
    const u32 s = 1u << (power - p - 1);
    for (u32 i = 0; i < (1u << p); ++i) {
      stack.push_back(st.load(pts[s * (2 * i + 1) - 1]));
      for (u32 k = 0; i & (1u << k); ++k) {
        Words top = std::move(stack.back());
        stack.pop_back();
        stack.back() = expMul(stack.back(), hashes[p - 1 - k], top);
      }
    }
    if (stack.size() != 1) { throw std::logic_error("proof: level did not reduce to one middle"); }
    if (std::all_of(stack[0].begin(), stack[0].end(), [](u32 w) { return w == 0; })) {
      throw GpuError("proof: zero middle");
    }
    pr.middles.push_back(std::move(stack[0]));
    hash = proofHash(E, hash, pr.middles.back());
    hashes.push_back(hash[0]);
  }
  return pr;
}

bool Gpu::verifyProof(const PrpProof& pr) {
  if (pr.E != E || pr.middles.empty()) { throw std::invalid_argument("proof of another exponent"); }
  Words A = makeWords(E, 3), Bw = pr.B;
  Bw.resize(nWords(E), 0);
  auto hash = proofHash(E, Bw);
  u32 span = E;
  for (const Words& M : pr.middles) {
    hash = proofHash(E, hash, M);
    const u64 h = hash[0];
    Bw = expMul(M, h, Bw, span % 2);
    A = expMul(A, h, M);
    span = (span + 1) / 2;
  }
  return expExp2(A, span) == Bw;
}

LLResult Gpu::isPrimeLL(u32 iters, std::vector<u64>* resAt, u32 resStep, Saver<LLState>* saver, u32 saveStep) {
  LLResult out;
  readROE();
  restartRoeWindow();
  u32 k0 = 0;
  double elapsedBefore = 0;
  if (saver) {
    LLState s = saver->load();
    writeIn(DATA, s.data);
    if (dataResidue() != res64(s.data)) { throw GpuError("invalid LL savefile (res64)"); }
    k0 = s.k;
    elapsedBefore = s.elapsed;
  } else {
    writeIn(DATA, makeWords(E, 4));
  }
  if (resAt && resStep) { resAt->assign(std::min(k0, iters) / resStep, 0); }
  double t0 = now();
  bool leadIn = true;
  for (u32 k = k0 + 1; k <= iters; ++k) {
    const bool save = saver && saveStep && (k % saveStep == 0 || k == iters);
    bool leadOut = (resStep && k % resStep == 0) || k == iters || save || useLongCarry;
    square(DATA, DATA, leadIn, leadOut, false, true);
    leadIn = leadOut;
    if (resAt && resStep && k % resStep == 0) { resAt->push_back(dataResidue()); }
    if (save) {
      Words w = readData(DATA);
      if (!w.empty()) {  // an all-zero residue is only saved by finishing (src/Gpu.cpp:1626-1642)
        w.resize(nWords(E), 0);
        saver->save({E, k, w, elapsedBefore + (now() - t0)});
      }
    }
  }
  // a savefile already past iters (left by a longer run) is reported as what it holds: x_k0
  out.k = std::max(k0, iters);
  Words w = readData();
  out.res64 = res64(w);
  out.isPrime = (out.k == E - 2) && w.empty();
  out.secsPerIt = (now() - t0) / std::max<u32>(1, iters);
  return out;
}

Words Gpu::isCERT(const Words& start, u32 squarings) {
  if (!squarings) { throw std::invalid_argument("CERT: no squarings"); }
  Words x = start;
  x.resize(nWords(E), 0);
  // only the E low bits are the start value (the file's last byte may carry more)
  if (E % 32) { x.back() &= (1u << (E % 32)) - 1; }
  readROE();
  restartRoeWindow(0);  // the reference measures no round-off for CERT (wantROE = 0)
  writeIn(DATA, x);
  squareLoop(DATA, 0, squarings);
  Words r = readData(DATA);
  r.resize(nWords(E), 0);
  return r;
}

RoeInfo Gpu::measureROE(u32 iters, u32 warmup) {
  if (!iters || iters > OWL_ROE_SIZE) { throw std::invalid_argument("measureROE: 1 .. 100,000 iterations"); }
  writeIn(DATA, makeWords(E, 3));
  if (warmup) { squareLoop(DATA, 0, warmup); }
  readROE();  // the warm-up iterations are not counted
  restartRoeWindow(iters);
  squareLoop(DATA, warmup, warmup + iters);
  RoeInfo r = readROE();
  restartRoeWindow();
  return r;
}

double Gpu::timePRP(u32 iters, u32 warmup, u32 B) {
  // Like the reference timePRP (src/Gpu.cpp:1337-1398): includes one Gerbicz update per block.
  writeIn(DATA, makeWords(E, 3));
  writeIn(CHECK, makeWords(E, 1));
  readROE();
  restartRoeWindow();
  u32 k = 0;
  modMul(CHECK, DATA);
  square(DATA, DATA, true, useLongCarry);
  ++k;
  while (k < warmup) {
    square(DATA, DATA, useLongCarry, useLongCarry);
    ++k;
  }
  finish();
  double t0 = now();
  bool leadIn = useLongCarry;
  while (true) {
    while (k % B < B - 1) {
      square(DATA, DATA, leadIn, useLongCarry);
      ++k;
      leadIn = useLongCarry;
    }
    square(DATA, DATA, useLongCarry, true);
    leadIn = true;
    ++k;
    if (k >= iters) { break; }
    modMul(CHECK, DATA);
  }
  finish();
  double secsPerIt = (now() - t0) / (iters - warmup);
  if (!doCheck(B)) { return 0.1 * 1e6; }
  return secsPerIt * 1e6;
}

double Gpu::timeSquares(u32 iters, u32 warmup, bool useGraph) {
  writeIn(DATA, makeWords(E, 3));
  rowLeadIn(DATA);
  for (u32 i = 0; i < warmup; ++i) {
    colSquare();
    rowFused(false, OWL_ROE_SIZE);
  }
  owl_graph graph = nullptr;
  const u32 G = 10;
  if (useGraph) {
    owlCheck(owl_hip_graph_begin(queue_), "graph begin");
    for (u32 i = 0; i < G; ++i) {
      colSquare();
      rowFused(false, OWL_ROE_SIZE);
    }
    owlCheck(owl_hip_graph_end(queue_, &graph), "graph end");
    owlCheck(owl_hip_graph_launch(graph, queue_), "graph launch");
  }
  finish();
  owl_event a, b;
  owlCheck(owl_hip_event_create(&a), "event");
  owlCheck(owl_hip_event_create(&b), "event");
  owlCheck(owl_hip_event_record(a, queue_), "event");
  u32 done = 0;
  while (done < iters) {
    if (useGraph) {
      owlCheck(owl_hip_graph_launch(graph, queue_), "graph launch");
      done += G;
    } else {
      colSquare();
      rowFused(false, OWL_ROE_SIZE);
      ++done;
    }
  }
  owlCheck(owl_hip_event_record(b, queue_), "event");
  float ms = 0;
  owlCheck(owl_hip_event_elapsed_ms(a, b, &ms), "elapsed");
  owl_hip_event_destroy(a);
  owl_hip_event_destroy(b);
  if (graph) { owl_hip_graph_destroy(graph); }
  // leave the state consistent: one more squaring ending in a lead-out
  colSquare();
  rowCarryOut(DATA, false, false, OWL_ROE_SIZE);
  carryAll(DATA);
  finish();
  return ms * 1e3 / done;
}
