// The modular-squaring engine for one exponent (the reference's class Gpu,
// src/Gpu.h:213-300, restricted to the PRP/LL squaring path).
#pragma once

#include "Buffer.h"
#include "FFTConfig.h"
#include "owl_hip.h"
#include "Proof.h"
#include "Saver.h"
#include "state.h"

#include <hip/hip_vector_types.h>

#include "job.h"
#include "TrigBufCache.h"

#include <initializer_list>
#include <memory>
#include <string>
#include <vector>

// Device address of the static staging array of the no-allocation digit readout on `device`
// (resolved once per device, kernels.hip); 0 on success.
extern "C" int owlNoAllocScratch(int device, int32_t** ptr);

// A named kernel of the offline-compiled library (src/Kernel.h:17-79).
class Kernel {
  owl_kernel k_ = nullptr;
  std::string name_;

public:
  Kernel() = default;
  explicit Kernel(const std::string& name);
  void operator()(owl_queue q, u32 gridX, u32 gridY, u32 block, u32 lds, std::initializer_list<void*> args) const;
  const std::string& name() const { return name_; }
};

// Round-off statistics over a set of iterations (src/Gpu.h:62-81).
struct RoeInfo {
  u32 N = 0;
  double max = 0, mean = 0, sd = 0;
};

struct GpuOptions {
  bool longCarry = false;  // lead-out + lead-in around every iteration (the reference's -carry long)
  bool generic = false;    // runtime-plan kernels even where a shape-specialized pass exists
  bool colOneRegion = false;  // the one-region column pass (k_coly) where k_colx exists
  bool colOnePair = false;    // k_colx with one column pair per group where two would be chosen
  u32 width = 0;           // W of the W x H split of the FFT; 0 = the reference's (refactored when too tall)
  u32 lineBlock = 0;       // lines per row-pass block; 0 = automatic
  int device = 0;
};

struct PRPResult {
  u64 res64 = 0;          // res64 of x_k at the last iteration run
  u32 k = 0;              // iterations done
  u32 nErrors = 0;        // Gerbicz check failures (each one was recovered by reloading)
  u32 nChecks = 0;        // Gerbicz checks passed
  bool isPrime = false;   // k == E and the residue is 9
  double maxRoe = 0;      // max round-off over the measured iterations
  double secsPerIt = 0;
};

struct LLResult {
  u64 res64 = 0;
  u32 k = 0;
  bool isPrime = false;   // k == E-2 and the residue is 0
  double secsPerIt = 0;
};

class Gpu {
public:
  enum Buf : int { DATA = 0, AUX = 1, CHECK = 2, BASE = 3 };

  Gpu(u32 E, const FFTConfig& fft, const GpuOptions& opt = {});
  ~Gpu();
  Gpu(const Gpu&) = delete;
  Gpu& operator=(const Gpu&) = delete;

  // geometry
  const u32 E;
  const FFTConfig fft;
  int device = 0;  // the HIP device the engine's queue and buffers are on
  u32 N, ND, W, H, R, nb, rowT, colT, rowLds, colLds;
  std::string rowKernel, colKernel;  // colKernel: what a launch of this job alone runs
  bool rowChunked = false;           // fused row launches in chunked per-XCD block order (k_rowx, nb >= 8)
  // Column pass with several column pairs per group ("colx<H>m", compiled for every k_colx
  // shape): colLdsM bytes of LDS (one more twiddle table), slotsM = groups the device holds
  // at once. colPairsPerGroup(n) is 1 (the one-pair "colx<H>") unless a launch of n jobs
  // still gives every slot a group with fewer groups per job.
  bool colMulti = false;
  u32 colLdsM = 0, slotsM = 0;
  u32 colPairsPerGroup(u32 nJobs) const;
  std::string colKernelFor(u32 nJobs) const;  // the column kernel a launch of nJobs jobs runs
  bool useLongCarry;
  std::vector<u32> rowPlan, colPlan;

  // state I/O (src/Gpu.cpp:801-811, 677-705)
  void writeIn(Buf b, const Words& words);
  void writeDigits(Buf b, const std::vector<int>& digits);  // sequential balanced digits
  std::vector<int> readDigits(Buf b);                       // sequential digits, checksum-verified
  // readDigits into out (N int32, host) through the library's static staging array instead of
  // a staging buffer allocated per call; rejects N > OWL_NOALLOC_MAX_WORDS
  void readDigitsNoAlloc(Buf b, int32_t* out);
  Words readData(Buf b = DATA);                             // canonical residue words
  u64 dataResidue(Buf b = DATA);                            // res64 via readResidue (src/Gpu.cpp:955-975)

  // the squaring path (src/Gpu.cpp:851-945)
  void square(Buf out, Buf in, bool leadIn, bool leadOut, bool doMul3 = false, bool doLL = false);
  u32 squareLoop(Buf out, Buf in, u32 from, u32 to, bool doTailMul3 = false, bool doLL = false);
  u32 squareLoop(Buf io, u32 from, u32 to, bool doLL = false) { return squareLoop(io, io, from, to, false, doLL); }

  // Gerbicz support (src/Gpu.cpp:711-771, 947-953)
  void modMul(Buf ioA, Buf inB, bool mul3 = false);
  bool doCheck(u32 blockSize);
  bool isEqual(Buf a, Buf b);
  void copy(Buf dst, Buf src);

  RoeInfo readROE() { return readROE(roePos); }
  // Round-off is measured on the next `n` squarings and multiplies (then not, like the
  // reference's 2,000 iterations per run, src/Gpu.cpp:1459-1464); the run loops restart it.
  void restartRoeWindow(u32 n = 2000) { roeWindow_ = n, roeTaken_ = 0; }
  RoeInfo readROE(u32 nSlots);   // the first nSlots round-off slots (then cleared)
  u32 roeCount() const { return roePos; }
  void finish();

  // workflows (src/Gpu.cpp:1337-1663), without proofs. With a saver, a run starts from the
  // most recent savefile (src/Gpu.cpp:1169-1191, 1595-1609) and writes one at every verified
  // Gerbicz check point (PRP, src/Gpu.cpp:1545-1551) or every saveStep iterations (LL,
  // src/Gpu.cpp:1619-1646); a resumed PRP run keeps the savefile's block size. resAt holds
  // 0 for the check points before the starting k.
  PRPResult isPrimePRP(u32 iters, u32 blockSize = 1000, u32 checkStep = 0, std::vector<u64>* resAt = nullptr,
                       u32 resStep = 0, Saver<PRPState>* saver = nullptr, const ProofStore* proof = nullptr);
  // proofs (src/Proof.cpp:84-126, 245-297; src/Gpu.cpp:813-889): A^h * B (B squared first
  // when squareB) and A^(2^n) on the device; a proof from the residues a full run kept;
  // the check of a proof (true when A^(2^span) meets B at the end of the hash chain)
  Words expMul(const Words& A, u64 h, const Words& B, bool squareB = false);
  Words expExp2(const Words& A, u32 n);
  PrpProof makeProof(const ProofStore& store);
  bool verifyProof(const PrpProof& proof);
  LLResult isPrimeLL(u32 iters, std::vector<u64>* resAt = nullptr, u32 resStep = 0, Saver<LLState>* saver = nullptr,
                     u32 saveStep = 10'000);
  // CERT (src/Gpu.cpp:1665-1725): `squarings` plain squarings of the start value (the M<E>.cert
  // file's (E-1)/8+1 little-endian bytes as words); the final residue, (E-1)/32+1 words.
  Words isCERT(const Words& start, u32 squarings);
  // DATA = 3 * check^(2^blockSize - 1) and CHECK = check: x_k from a PRP savefile's check
  // (src/Gpu.cpp:738-765)
  void writeState(const Words& check, u32 blockSize);
  double timePRP(u32 iters = 1000, u32 warmup = 30, u32 blockSize = 200);
  // Gpu::measureROE (src/Gpu.cpp:1268-1335): the round-off of `iters` squarings of a PRP-3 run
  // after `warmup` unmeasured ones, one sample (the max over the words) per squaring.
  RoeInfo measureROE(u32 iters = 2000, u32 warmup = 50);
  double timeSquares(u32 iters, u32 warmup, bool useGraph);

  // raw access for tests and the C-ABI
  owl_queue queue() const { return queue_; }
  const JobDesc& job() const { return job_; }
  const JobDesc* jobDev() const { return jobs_.get(); }
  u32 flagsRead();  // bit0 carry leftover, bit1 mismatch, bit2 hand-off timeout

  // Where a pass runs: n jobs of this shape (descriptors in device memory, this Gpu's
  // geometry and kernels) on queue q. Gpu's own calls use self().
  struct Launch {
    const JobDesc* jobs;
    u32 n;
    owl_queue q;
  };
  Launch self() const { return {jobs_.get(), 1, queue_}; }

  // kernel-level launches (used by square/modMul; exposed for tests, timing and batches)
  void rowFused(bool ll, u32 roeSlot) { rowFused(ll, roeSlot, self()); }
  void rowLeadIn(Buf in) { rowLeadInTo(in, 0, self()); }
  void rowCarryOut(Buf out, bool ll, bool mul3, u32 roeSlot) { rowCarryOut(out, ll, mul3, roeSlot, self()); }
  void colSquare() { colSquare(self()); }
  void colMul() { colMul(self()); }                // state *= spectrum in freq[2]
  void colFwdSpec() { colFwdSpec(self()); }        // freq[2] <- spectrum of freq[2]
  void rowLeadInTo(Buf in, int freqIdx) { rowLeadInTo(in, freqIdx, self()); }
  void carryAll(Buf b) { carryAll(b, self()); }

  void rowFused(bool ll, u32 roeSlot, Launch L);
  void rowLeadInTo(Buf in, int freqIdx, Launch L);
  void rowCarryOut(Buf out, bool ll, bool mul3, u32 roeSlot, Launch L);
  void colSquare(Launch L);
  void colMul(Launch L);
  void colFwdSpec(Launch L);
  void colPass(int mode, int fi, int oi, Launch L);
  void carryAll(Buf b, Launch L);
  void square(Buf out, Buf in, bool leadIn, bool leadOut, bool doMul3, bool doLL, Launch L);
  u32 squareLoop(Buf out, Buf in, u32 from, u32 to, bool doTailMul3, bool doLL, Launch L);
  void modMul(Buf ioA, Buf inB, bool mul3, Launch L);
  u32 nextRoeSlot();
  void resetRoeCount() { roePos = 0; }
  // A tag for the next fused launch: never the tag of the fused launch before it (noteTag)
  // nor one of a captured graph's tags (reserveTags), so that consecutive launches on a
  // job's carry stairs always differ, also around graph replays and re-captures.
  u32 nextStairTag();
  void noteTag(u32 t) { lastTag_ = t; }
  u32 lastTag() const { return lastTag_; }
  void reserveTags(u32 lo, u32 hi) { resLo_ = lo, resHi_ = hi; }
  // take over another engine's tag walk (a group's launches moving to a new lead)
  void inheritTags(const Gpu& o) {
    stairTag_ = o.stairTag_, lastTag_ = o.lastTag_, resLo_ = o.resLo_, resHi_ = o.resHi_;
  }
  // G fused squarings (column pass + fused row pass) captured as one graph on L.q;
  // its tags are reserved until the next capture. Returns the graph and its last tag.
  owl_graph captureFused(u32 G, bool ll, u32 roeSlot, Launch L, u32* lastTagOut);
  // res64 of b read on queue q into this job's small buffer; residueTaken() converts it
  // after q has finished (the two halves of dataResidue, for launches shared by a group)
  void residueOn(owl_queue q, Buf b);
  u64 residueTaken();
  void copyOn(owl_queue q, Buf dst, Buf src);
  // device bytes this engine holds (state, Gerbicz buffers, carries, tables of its own)
  u64 deviceBytes() const;
  // an upper bound of deviceBytes() for an exponent at FFT size N, before allocating
  static u64 footprintBound(u32 N, u32 W, u32 H);
  // the shape key of a batch group: jobs with equal keys can share launches
  std::string shapeKey() const;
  // Failure-handling tests: the PRP loops flip one bit of the data right after iteration k
  // (once), as a hardware error would; the next Gerbicz check must catch it.
  void armFault(u32 k) { faultAt_ = k; }
  u32 faultAt() const { return faultAt_; }
  void injectFault();

private:
  owl_queue queue_ = nullptr;
  JobDesc job_{};
  Buffer<JobDesc> jobs_;
  std::vector<std::unique_ptr<Buffer<char>>> owned_;
  std::shared_ptr<const TrigTables> trig_;
  u32 roePos = 0, roeWindow_ = 2000, roeTaken_ = 0;
  u32 stairTag_ = 0, lastTag_ = 0, resLo_ = 1, resHi_ = 0;
  u32 faultAt_ = 0;
  bool tunedLineBlock_ = false;
  std::string colOne_;  // the one-pair column kernel
  Kernel kRow, kCol, kColM, kCarryB, kCarryFix, kWordsIn, kWordsOut, kResidue, kIsEqual, kSum64, kCopy;

  template<typename T> T* alloc(size_t n);
  template<typename T> T* upload(const std::vector<T>& v);
  void readDigitsVia(Buf b, int32_t* staging, int32_t* out);
};

// Mersenne-prime type-1 PRP residue is 9 (src/Gpu.cpp:1036-1040).
bool equals9(const Words& w);
