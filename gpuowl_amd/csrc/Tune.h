// Autotuning of the engine's FFT split, line block and pass variants per FFT length
// (in place of the reference's tune.cpp / tune.txt, src/tune.cpp:328-738,
// src/TuneEntry.cpp:7-84, which tune the choice of FFT *length* per exponent; here the
// length stays FFTConfig::bestFit's and the tuner picks how the engine runs it).
#pragma once

#include "Gpu.h"

#include <string>
#include <vector>

struct TuneChoice {
  u32 width = 0;      // W of the W x H split (0 = the engine's rule)
  u32 lineBlock = 0;  // R (0 = the engine's rule)
  bool generic = false, colOneRegion = false, colOnePair = false;
  double us = 0;      // measured microseconds per iteration
};

// Process-wide table N -> choice, consulted by Gpu when no option overrides it.
namespace TuneTable {
void set(u32 N, const TuneChoice& c);
bool get(u32 N, TuneChoice* c);
void clear();
// text lines "N width lineBlock flags us" (flags: 2 = generic, 4 = one-region column pass,
// 8 = one column pair per column group)
void load(const std::string& path);
void save(const std::string& path);
}  // namespace TuneTable

// Process-wide table N -> choice for jobs run in batches (GpuBatch), filled by tuneBatch: a
// batch of many jobs can prefer another split than one job alone (more row groups per CU,
// e.g. 512 x 1536 instead of 1024 x 768 at N = 1,572,864).
namespace BatchTuneTable {
void set(u32 N, const TuneChoice& c);
bool get(u32 N, TuneChoice* c);
void clear();
}  // namespace BatchTuneTable

struct TuneResult {
  TuneChoice choice;
  u32 W = 0, H = 0, R = 0;
  bool longCarry = false;
  std::string rowKernel, colKernel;
};

// Time every candidate (split, line block, pass variant) for exponent E at the FFT
// length bestFit(E, spec) gives: steady squarings replayed from graphs, iters of them.
// Fastest first. Candidates that the engine rejects are skipped.
std::vector<TuneResult> tuneShape(u32 E, const std::string& spec, u32 iters, u32 warmup, int device = 0);

// For every FFT length among the exponents (bestFit(E, spec)), time a batch of those
// exponents under each W x H split (line block and pass variants by the engine's rules) over
// iters graph-replayed batch iterations, record the fastest in BatchTuneTable and return all
// results (us = microseconds per batch iteration), grouped by length, fastest first.
std::vector<TuneResult> tuneBatch(const std::vector<u32>& exponents, const std::string& spec, u32 iters, u32 warmup);
