// A request loop over one device: PRP and LL jobs arrive at any time, are admitted while
// their device footprint fits the memory budget and a slot is free, and advance together
// in rounds of Gerbicz blocks (blockSize iterations; groups of smaller FFTs run more blocks
// per round, so that a round takes about as long for every group). Jobs of one FFT
// geometry and kind share every launch
// of a round (one group, one queue, one captured graph of fused squarings); groups of
// different geometries run on their own queues and overlap on the device. Admission,
// retirement and Gerbicz verification happen between rounds, so a request that arrives
// while others are running starts at the next round boundary (continuous batching).
//
// The reference runs one exponent per worker (src/main.cpp:110-117, src/Task.cpp) and has no
// request loop; each job here follows its isPrimePRP / isPrimeLL (src/Gpu.cpp:1400-1663):
// Gerbicz update every block, verification every checkStep and past the end, the result
// residue taken at the requested iteration and the run continued to the next block boundary
// for the final check (src/Gpu.cpp:1441-1445, 1496-1506).
#pragma once

#include "Gpu.h"

#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

struct ServeResult {
  u32 id = 0, E = 0, iters = 0;
  bool ll = false;
  u64 res64 = 0;          // res64 of x_iters
  bool isPrime = false;   // iters == E (PRP, x == 9) or E - 2 (LL, x == 0)
  u32 nErrors = 0, nChecks = 0;
  double maxRoe = 0;
  double tSubmit = 0, tAdmit = 0, tDone = 0;  // seconds on the server's clock
  int status = 0;         // 0 done, 1 failed (repeated Gerbicz errors)
  u64 res64Type1 = 0;     // a PRP request run to E: the type-1 residue x_E / 9
  std::string res2048;    // of the same, 512 hex digits
};

struct ServeInfo {
  u32 waiting = 0, active = 0, groups = 0, done = 0;
  u64 bytesInUse = 0, budget = 0;
  u64 iterations = 0;      // squarings completed over all jobs (verification work excluded)
  u64 wordIterations = 0;  // the same, weighted by FFT size N
  u32 rounds = 0;
};

class Server {
public:
  struct Options {
    u64 memBudget = 0;       // device bytes for job state; 0 = 3/4 of the free memory
    u32 maxJobs = 256;       // active jobs at most
    u32 blockSize = 1000;    // iterations per round and Gerbicz block
    u32 checkStep = 50000;   // verification cadence (a multiple of blockSize)
    std::string spec;        // FFT spec for every job; empty = FFTConfig::bestFit's choice
    GpuOptions gpu;
  };

  explicit Server(const Options& opt);
  ~Server();

  // Queue a request: iters iterations of PRP (x0 = 3) or LL (x0 = 4). Returns its id.
  u32 submit(u32 E, bool ll, u32 iters);
  // Admit what fits, run one round on every active job, verify and retire. Returns the number
  // of jobs that finished in this round.
  u32 step();
  // Finished results since the last call.
  std::vector<ServeResult> take();
  // Failure-handling tests: flip one data bit of request `id` (a PRP request not admitted yet)
  // at the end of the round in which it passes iteration k, once, as a hardware error would;
  // its next Gerbicz check must catch it and reload its last verified state.
  void armFault(u32 id, u32 k);
  ServeInfo info() const;
  double now() const;

private:
  struct Request {
    u32 id, E, iters;
    bool ll;
    double tSubmit;
  };
  struct Job {
    Request req;
    std::unique_ptr<Gpu> gpu;
    u64 bytes = 0;
    u32 k = 0, kEndEnd = 0, vK = 0;
    u32 faultAt = 0;  // armFault: inject after the round that passes this iteration (0 = none)
    bool skip = false, captured = false, resRead = false, wantCheck = false;
    Words vData, vCheck;
    ServeResult res;
  };
  struct Group {
    std::string key;
    bool ll = false;
    std::vector<Job*> members;
    Buffer<JobDesc> all, sub;
    u32 cap = 0;
    Gpu* lead = nullptr;
    owl_graph graph = nullptr;
    u32 graphLast = 0;
    bool dirty = true;
    Gpu::Launch launch() const { return {all.get(), u32(members.size()), lead->queue()}; }
  };

  Options opt_;
  double t0_;
  u32 nextId_ = 1;
  std::deque<Request> waiting_;
  std::map<u32, u32> faults_;  // armFault requests by id, taken at admission
  std::vector<std::unique_ptr<Job>> active_;
  std::map<std::string, std::unique_ptr<Group>> groups_;
  std::vector<ServeResult> done_;
  u64 bytesInUse_ = 0, iterations_ = 0, wordIterations_ = 0;
  u32 rounds_ = 0, nDone_ = 0;

  static constexpr u32 GRAPH_LEN = 32;

  void admit();
  void prepare(Group& g);
  Gpu::Launch subset(Group& g, const std::vector<Job*>& js);
  void runRound(Group& g, u32 blocks);
  void fused(Group& g, u32 count, u32 roeSlot);
  void verify(Group& g);
  void retire();
  Group& groupOf(Job& j);
};
