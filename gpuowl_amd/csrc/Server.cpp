// Request loop with continuous batching; see Server.h.
#include "Server.h"

#include "Worktodo.h"

#include <algorithm>
#include <chrono>
#include <stdexcept>

namespace {
double steadySecs() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}
Words small(u32 E, u32 v) {
  Words w(nWords(E));
  w[0] = v;
  return w;
}
}  // namespace

Server::Server(const Options& opt) : opt_{opt}, t0_{steadySecs()} {
  if (opt_.blockSize < 2) { throw std::invalid_argument("block size must be at least 2"); }
  if (!opt_.checkStep) { opt_.checkStep = 50 * opt_.blockSize; }
  if (opt_.checkStep % opt_.blockSize) { throw std::invalid_argument("check step must be a multiple of the block size"); }
  if (!opt_.maxJobs) { throw std::invalid_argument("maxJobs"); }
  if (!opt_.memBudget) {
    size_t total = 0, freeMem = 0;
    int cu = 0, lds = 0;
    owlCheck(owl_hip_device_info(opt_.gpu.device, nullptr, 0, &total, &freeMem, &cu, &lds), "device info");
    opt_.memBudget = freeMem / 4 * 3;
  }
}

Server::~Server() {
  for (auto& [key, g] : groups_) {
    if (g->lead) { owl_hip_finish(g->lead->queue()); }
    if (g->graph) { owl_hip_graph_destroy(g->graph); }
  }
  groups_.clear();
  active_.clear();
}

double Server::now() const { return steadySecs() - t0_; }

u32 Server::submit(u32 E, bool ll, u32 iters) {
  if (!iters) { throw std::invalid_argument("a request needs at least one iteration"); }
  if (ll && iters > E - 2) { throw std::invalid_argument("LL runs at most E - 2 iterations"); }
  if (!ll && iters > E) { throw std::invalid_argument("PRP runs at most E iterations"); }
  FFTConfig fft = FFTConfig::bestFit(E, opt_.spec);  // throws for an exponent no FFT covers
  if (double(E) / fft.size() < FFTShape::MIN_BPW || E % 32 == 0 || E < 64) {
    throw std::invalid_argument("exponent " + std::to_string(E) + " is too small for FFT " + fft.spec());
  }
  // a request that could never be admitted would wait forever: reject it now
  const u64 need = Gpu::footprintBound(fft.size(), fft.shape.width, fft.shape.middle * fft.shape.height);
  if (need > opt_.memBudget) {
    throw std::invalid_argument("exponent " + std::to_string(E) + " needs " + std::to_string(need) +
                                " device bytes, more than the server's memory budget of " +
                                std::to_string(opt_.memBudget));
  }
  u32 id = nextId_++;
  waiting_.push_back({id, E, iters, ll, now()});
  return id;
}

Server::Group& Server::groupOf(Job& j) {
  std::string key = j.gpu->shapeKey() + (j.req.ll ? "/ll" : "/prp");
  auto& slot = groups_[key];
  if (!slot) {
    slot = std::make_unique<Group>();
    slot->key = key;
    slot->ll = j.req.ll;
  }
  return *slot;
}

void Server::admit() {
  for (auto it = waiting_.begin(); it != waiting_.end() && active_.size() < opt_.maxJobs;) {
    const Request& r = *it;
    FFTConfig fft = FFTConfig::bestFit(r.E, opt_.spec);
    u32 N = fft.size(), W = fft.shape.width, H = fft.shape.middle * fft.shape.height;
    if (bytesInUse_ + Gpu::footprintBound(N, W, H) > opt_.memBudget) {
      ++it;  // first fit: a smaller request behind may still fit
      continue;
    }
    auto job = std::make_unique<Job>();
    job->req = r;
    job->res.id = r.id;
    job->res.E = r.E;
    job->res.iters = r.iters;
    job->res.ll = r.ll;
    job->res.tSubmit = r.tSubmit;
    job->res.tAdmit = now();
    try {
      job->gpu = std::make_unique<Gpu>(r.E, fft, opt_.gpu);
    } catch (const std::invalid_argument&) {
      job->res.status = 2;  // rejected: no engine for this exponent
      job->res.tDone = now();
      done_.push_back(job->res);
      ++nDone_;
      it = waiting_.erase(it);
      continue;
    }
    Gpu& g = *job->gpu;
    if (auto f = faults_.find(r.id); f != faults_.end()) {
      job->faultAt = f->second;
      faults_.erase(f);
    }
    job->bytes = g.deviceBytes();
    job->kEndEnd = (r.iters + opt_.blockSize - 1) / opt_.blockSize * opt_.blockSize;
    if (r.ll) {
      g.writeIn(Gpu::DATA, small(r.E, 4));
    } else {
      job->vData = small(r.E, 3);
      job->vCheck = small(r.E, 1);
      g.writeIn(Gpu::DATA, job->vData);
      g.writeIn(Gpu::CHECK, job->vCheck);
    }
    bytesInUse_ += job->bytes;
    Group& grp = groupOf(*job);
    if (!grp.lead) { grp.lead = &g; }
    grp.members.push_back(job.get());
    grp.dirty = true;
    active_.push_back(std::move(job));
    it = waiting_.erase(it);
  }
}

void Server::prepare(Group& g) {
  const u32 n = u32(g.members.size());
  if (n > g.cap) {
    g.cap = std::max<u32>({n, 2 * g.cap, 8});
    g.all = Buffer<JobDesc>(g.cap);
    g.sub = Buffer<JobDesc>(g.cap);
  }
  std::vector<JobDesc> d;
  for (Job* j : g.members) { d.push_back(j->gpu->job()); }
  g.all.write(g.lead->queue(), d);
  if (g.graph) {
    owlCheck(owl_hip_graph_destroy(g.graph), "graph destroy");
    g.graph = nullptr;
  }
  if (!g.lead->useLongCarry) { g.graph = g.lead->captureFused(GRAPH_LEN, g.ll, OWL_ROE_SIZE, g.launch(), &g.graphLast); }
  g.dirty = false;
}

Gpu::Launch Server::subset(Group& g, const std::vector<Job*>& js) {
  if (js.size() == g.members.size()) { return g.launch(); }
  std::vector<JobDesc> d;
  for (Job* j : js) { d.push_back(j->gpu->job()); }
  g.sub.write(g.lead->queue(), d);  // blocking: the queue is idle between rounds
  return {g.sub.get(), u32(js.size()), g.lead->queue()};
}

// count fused squarings (column pass + fused row pass), from captured graphs where possible;
// a round that measures round-off (slot 0) launches them directly
void Server::fused(Group& g, u32 count, u32 roeSlot) {
  Gpu& L = *g.lead;
  Gpu::Launch ln = g.launch();
  u32 i = 0;
  for (; g.graph && roeSlot >= OWL_ROE_SIZE && i + GRAPH_LEN <= count; i += GRAPH_LEN) {
    owlCheck(owl_hip_graph_launch(g.graph, ln.q), "graph launch");
    L.noteTag(g.graphLast);
  }
  for (; i < count; ++i) {
    L.colSquare(ln);
    L.rowFused(g.ll, roeSlot, ln);
  }
}

// `blocks` Gerbicz blocks of blockSize iterations for every member of the group
void Server::runRound(Group& g, u32 blocks) {
  Gpu& L = *g.lead;
  const u32 B = opt_.blockSize;
  Gpu::Launch ln = g.launch();
  std::vector<u32> kStart;
  for (Job* j : g.members) { kStart.push_back(j->k); }

  for (u32 blk = 0; blk < blocks; ++blk) {
    // Gerbicz block update: check *= data (src/Gpu.cpp:1466-1471); after the first block
    // every member takes it, so only the first may need a subset launch
    if (!g.ll) {
      std::vector<Job*> mul;
      for (Job* j : g.members) {
        if (j->skip) {
          j->skip = false;
        } else {
          mul.push_back(j);
        }
      }
      if (!mul.empty()) { L.modMul(Gpu::CHECK, Gpu::DATA, false, subset(g, mul)); }
    }

    // segment ends: the requested iteration of every job that reaches it in this block
    std::vector<u32> cuts;
    for (Job* j : g.members) {
      if (!j->captured && j->req.iters > j->k && j->req.iters <= j->k + B) { cuts.push_back(j->req.iters - j->k); }
    }
    cuts.push_back(B);
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());

    // round-off is measured (slot 0) in the first two blocks of every job, then not
    bool measure = false;
    for (Job* j : g.members) { measure = measure || j->k < 2 * B; }
    const u32 roeSlot = measure ? 0 : OWL_ROE_SIZE;
    u32 pos = 0;
    for (u32 cut : cuts) {
      const u32 n = cut - pos;
      if (L.useLongCarry) {
        for (u32 i = 0; i < n; ++i) { L.square(Gpu::DATA, Gpu::DATA, true, true, false, g.ll, ln); }
      } else {
        L.rowLeadInTo(Gpu::DATA, 0, ln);
        fused(g, n - 1, roeSlot);
        L.colSquare(ln);
        L.rowCarryOut(Gpu::DATA, g.ll, false, roeSlot, ln);
        L.carryAll(Gpu::DATA, ln);
      }
      for (Job* j : g.members) {
        if (!j->captured && j->req.iters == j->k + cut) {
          j->gpu->residueOn(ln.q, Gpu::DATA);
          bool final = j->req.iters == (g.ll ? j->req.E - 2 : j->req.E);
          if (final) { j->gpu->copyOn(ln.q, Gpu::BASE, Gpu::DATA); }
          j->captured = true;
        }
      }
      pos = cut;
    }
    for (Job* j : g.members) {
      u32 useful = j->k < j->req.iters ? std::min(B, j->req.iters - j->k) : 0;
      iterations_ += useful;
      wordIterations_ += u64(useful) * j->gpu->N;
      j->k += B;
    }
  }
  // a PRP job is verified when the round crossed a check point, and past its end
  for (size_t i = 0; i < g.members.size(); ++i) {
    Job* j = g.members[i];
    if (!g.ll) { j->wantCheck = (j->k / opt_.checkStep != kStart[i] / opt_.checkStep) || j->k >= j->kEndEnd; }
  }
}

// Gerbicz verification of the jobs due for it (src/Gpu.cpp:767-771)
void Server::verify(Group& g) {
  std::vector<Job*> due;
  for (Job* j : g.members) {
    if (j->wantCheck) { due.push_back(j); }
  }
  if (due.empty()) { return; }
  Gpu::Launch ls = subset(g, due);
  g.lead->squareLoop(Gpu::AUX, Gpu::CHECK, 0, opt_.blockSize, true, false, ls);
  g.lead->modMul(Gpu::CHECK, Gpu::DATA, false, ls);
}

void Server::retire() {
  std::vector<std::unique_ptr<Job>> keep;
  for (auto& jp : active_) {
    Job& j = *jp;
    bool finished = j.k >= j.kEndEnd && j.captured && j.res.status == 0;
    if (!finished && j.res.status == 0) {
      keep.push_back(std::move(jp));
      continue;
    }
    j.res.tDone = now();
    done_.push_back(j.res);
    ++nDone_;
    Group& g = groupOf(j);
    g.members.erase(std::find(g.members.begin(), g.members.end(), &j));
    g.dirty = true;
    if (g.lead == j.gpu.get()) {
      Gpu* next = g.members.empty() ? nullptr : g.members.front()->gpu.get();
      if (next) { next->inheritTags(*g.lead); }
      if (g.graph) {
        owlCheck(owl_hip_graph_destroy(g.graph), "graph destroy");
        g.graph = nullptr;
      }
      g.lead = next;
    }
    bytesInUse_ -= j.bytes;
    if (g.members.empty()) {
      std::string key = g.key;
      groups_.erase(key);
    }
  }
  active_ = std::move(keep);
}

u32 Server::step() {
  admit();
  std::vector<Group*> gs;
  for (auto& [key, g] : groups_) {
    if (g->dirty) { prepare(*g); }
    gs.push_back(g.get());
  }
  // balance the round: a group of a smaller FFT runs proportionally more blocks (up to 16),
  // never past the end of its shortest job
  u64 nMax = 0;
  for (Group* g : gs) { nMax = std::max<u64>(nMax, g->lead->N); }
  for (Group* g : gs) {
    u32 blocks = u32(std::min<u64>(16, std::max<u64>(1, nMax / g->lead->N)));
    for (Job* j : g->members) {
      u32 left = j->k < j->kEndEnd ? (j->kEndEnd - j->k) / opt_.blockSize : 0;
      blocks = std::min(blocks, std::max<u32>(1, left));
    }
    runRound(*g, blocks);
  }
  for (Group* g : gs) { g->lead->finish(); }

  // armed faults: one bit of the data flipped after the round that passed iteration faultAt
  for (Group* g : gs) {
    for (Job* j : g->members) {
      if (j->faultAt && j->k >= j->faultAt) {
        j->faultAt = 0;
        j->gpu->injectFault();
      }
    }
  }

  // residues taken in this round
  for (Group* g : gs) {
    for (Job* j : g->members) {
      if (j->captured && !j->resRead) {
        j->resRead = true;
        j->res.res64 = j->gpu->residueTaken();
        if (j->req.iters == (g->ll ? j->req.E - 2 : j->req.E)) {
          Words w = j->gpu->readData(Gpu::BASE);
          j->res.isPrime = g->ll ? w.empty() : equals9(w);
          if (!g->ll) {  // the residue results lines report (src/Gpu.cpp:1495-1504)
            w.resize(nWords(j->req.E), 0);
            Words q = div9(w, j->req.E);
            j->res.res64Type1 = res64(q);
            j->res.res2048 = res2048Hex(q);
          }
        }
      }
    }
  }

  for (Group* g : gs) {
    if (!g->ll) { verify(*g); }
  }
  for (Group* g : gs) { g->lead->finish(); }
  for (Group* g : gs) {
    const u32 nSlots = std::max<u32>(1, std::min<u32>(g->lead->roeCount(), OWL_ROE_SIZE));
    for (Job* j : g->members) { j->res.maxRoe = std::max(j->res.maxRoe, j->gpu->readROE(nSlots).max); }
    g->lead->resetRoeCount();
    for (Job* j : g->members) {
      if (!j->wantCheck) { continue; }
      j->wantCheck = false;
      if (j->gpu->isEqual(Gpu::CHECK, Gpu::AUX)) {
        ++j->res.nChecks;
        j->skip = true;
        j->vK = j->k;
        if (j->k < j->kEndEnd) {
          j->vData = j->gpu->readData(Gpu::DATA);
          j->vCheck = j->gpu->readData(Gpu::CHECK);
        }
      } else {
        // reload the last verified state (src/Gpu.cpp:1561-1575)
        if (++j->res.nErrors > 3) {
          j->res.status = 1;
          continue;
        }
        j->gpu->writeIn(Gpu::DATA, j->vData);
        j->gpu->writeIn(Gpu::CHECK, j->vCheck);
        if (j->vK < j->req.iters) {
          j->captured = j->resRead = false;
          j->res.res64 = 0;
          j->res.isPrime = false;
        }
        j->k = j->vK;
        j->skip = j->vK > 0;
      }
    }
  }
  u32 before = nDone_;
  retire();
  ++rounds_;
  return nDone_ - before;
}

void Server::armFault(u32 id, u32 k) {
  auto it = std::find_if(waiting_.begin(), waiting_.end(), [&](const Request& r) { return r.id == id; });
  if (it == waiting_.end()) { throw std::invalid_argument("armFault: no waiting request with that id"); }
  if (it->ll) { throw std::invalid_argument("armFault: LL requests have no error check"); }
  if (!k || k >= it->iters) { throw std::invalid_argument("armFault: the iteration must lie inside the request"); }
  faults_[id] = k;
}

std::vector<ServeResult> Server::take() {
  std::vector<ServeResult> out;
  out.swap(done_);
  return out;
}

ServeInfo Server::info() const {
  ServeInfo s;
  s.waiting = u32(waiting_.size());
  s.active = u32(active_.size());
  s.groups = u32(groups_.size());
  s.done = nDone_;
  s.bytesInUse = bytesInUse_;
  s.budget = opt_.memBudget;
  s.iterations = iterations_;
  s.wordIterations = wordIterations_;
  s.rounds = rounds_;
  return s;
}
