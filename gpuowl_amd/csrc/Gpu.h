// The modular-squaring engine for one exponent (the reference's class Gpu,
// src/Gpu.h:213-300, restricted to the PRP/LL squaring path).
#pragma once

#include "Buffer.h"
#include "FFTConfig.h"
#include "owl_hip.h"
#include "state.h"

#include <hip/hip_vector_types.h>

#include "job.h"

#include <initializer_list>
#include <memory>
#include <string>
#include <vector>

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
  u32 N, ND, W, H, R, nb, rowT, colT;
  bool useLongCarry;
  std::vector<u32> rowPlan, colPlan;

  // state I/O (src/Gpu.cpp:801-811, 677-705)
  void writeIn(Buf b, const Words& words);
  void writeDigits(Buf b, const std::vector<int>& digits);  // sequential balanced digits
  std::vector<int> readDigits(Buf b);                       // sequential digits, checksum-verified
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

  RoeInfo readROE();
  void finish();

  // workflows (src/Gpu.cpp:1337-1663), without savefiles or proofs
  PRPResult isPrimePRP(u32 iters, u32 blockSize = 1000, u32 checkStep = 0, std::vector<u64>* resAt = nullptr,
                       u32 resStep = 0);
  LLResult isPrimeLL(u32 iters, std::vector<u64>* resAt = nullptr, u32 resStep = 0);
  double timePRP(u32 iters = 1000, u32 warmup = 30, u32 blockSize = 200);
  double timeSquares(u32 iters, u32 warmup, bool useGraph);

  // raw access for tests and the C-ABI
  owl_queue queue() const { return queue_; }
  const JobDesc& job() const { return job_; }
  const JobDesc* jobDev() const { return jobs_.get(); }
  u32 flagsRead();  // bit0 carry leftover, bit1 mismatch, bit2 hand-off timeout

  // kernel-level launches (used by square/modMul; exposed for tests and timing)
  void rowFused(bool ll, u32 roeSlot);
  void rowLeadIn(Buf in);
  void rowCarryOut(Buf out, bool ll, bool mul3, u32 roeSlot);
  void colSquare();
  void colMul();        // state *= spectrum in freq[2]
  void colFwdSpec();    // freq[2] <- spectrum of freq[2]
  void rowLeadInTo(Buf in, int freqIdx);
  void carryAll(Buf b);

private:
  owl_queue queue_ = nullptr;
  JobDesc job_{};
  Buffer<JobDesc> jobs_;
  std::vector<std::unique_ptr<Buffer<char>>> owned_;
  u32 roePos = 0;
  u32 stairTag_ = 0;
  Kernel kRow, kCol, kCarryB, kCarryFix, kWordsIn, kWordsOut, kResidue, kIsEqual, kSum64, kCopy;

  template<typename T> T* alloc(size_t n);
  template<typename T> T* upload(const std::vector<T>& v);
  u32 nextRoeSlot();
};

// Mersenne-prime type-1 PRP residue is 9 (src/Gpu.cpp:1036-1040).
bool equals9(const Words& w);
