// Many exponents advanced together. Jobs of the same FFT geometry form a group
// that shares every launch (the kernels index their job descriptor by blockIdx.y),
// and each group runs on its own queue, so groups of different shapes overlap.
// The reference runs one exponent per Gpu (at most 4 worker threads, src/main.cpp:110-117).
#pragma once

#include "Gpu.h"

#include <memory>
#include <vector>

class GpuBatch {
public:
  struct Group {
    std::vector<size_t> index;  // positions in the batch
    Buffer<JobDesc> descs;      // the group's job descriptors, in device memory
    Gpu* lead = nullptr;        // geometry, kernels, queue and tags of the group
    Gpu::Launch launch() const { return {descs.get(), u32(index.size()), lead->queue()}; }
  };

  // Every exponent gets FFTConfig::bestFit(E, spec) (spec empty: the reference's choice).
  GpuBatch(const std::vector<u32>& exponents, const std::string& spec, const GpuOptions& opt = {});

  std::vector<std::unique_ptr<Gpu>> jobs;
  std::vector<Group> groups;

  size_t size() const { return jobs.size(); }
  const Group& groupOf(size_t job) const;
  u64 totalWords() const;
  void finish();

  // All jobs: out = in^(2^(to-from)) (x3 on the last if mul3Last; LL if ll).
  void squareLoop(Gpu::Buf out, Gpu::Buf in, u32 from, u32 to, bool mul3Last, bool ll);
  void modMul(Gpu::Buf ioA, Gpu::Buf inB, bool mul3);
  std::vector<bool> doCheck(u32 blockSize);
  std::vector<u64> res64(Gpu::Buf b);

  // PRP (x0 = 3) or LL (x0 = 4) for iters iterations on every job in lockstep; PRP with
  // Gerbicz checks every checkStep (a failed job reloads the whole batch from the last
  // verified state). res64 of every job after every resStep iterations -> resAt[job].
  std::vector<PRPResult> prp(u32 iters, u32 blockSize, u32 checkStep, std::vector<std::vector<u64>>* resAt,
                             u32 resStep);
  std::vector<LLResult> ll(u32 iters, std::vector<std::vector<u64>>* resAt, u32 resStep);

  // Steady-state squarings of every job: microseconds per batch iteration (all jobs
  // advanced once), optionally replayed from captured graphs (one per group).
  double timeSquares(u32 iters, u32 warmup, bool useGraph);
};
