// extern "C" entry points of the engine (declared in include/owl.h).
#include "owl.h"

#include "Gpu.h"
#include "GpuBatch.h"
#include "Server.h"
#include "Trig.h"
#include "Tune.h"
#include "Worktodo.h"

#include <deque>
#include <memory>
#include <array>

#include <cstdio>
#include <cstring>
#include <exception>
#include <string>

struct owl_gpu_s {
  Gpu* gpu;
};
struct owl_batch_s {
  GpuBatch* batch;
};
struct owl_server_s {
  Server* server;
  std::deque<ServeResult> ready;
};

// Shares the thread-local error slot of the device layer.
extern "C" int owl_set_error(const char* msg);

namespace {
template<typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    owl_set_error(e.what());
    return -1;
  } catch (const char* s) {
    owl_set_error(s);
    return -1;
  } catch (...) {
    owl_set_error("unknown error");
    return -1;
  }
}

Gpu::Buf buf(int b) {
  if (b < 0 || b > 3) { throw std::invalid_argument("buffer index must be 0..3"); }
  return Gpu::Buf(b);
}
}  // namespace

extern "C" {

int owl_fft_best_fit(uint32_t E, const char* spec, char* outSpec, size_t outLen, uint32_t* N, uint32_t* maxExp) {
  return guarded([&] {
    FFTConfig fft = FFTConfig::bestFit(E, spec ? spec : "");
    if (outSpec && outLen) { snprintf(outSpec, outLen, "%s", fft.spec().c_str()); }
    if (N) { *N = fft.size(); }
    if (maxExp) { *maxExp = fft.maxExp(); }
  });
}

int owl_fft_best_fit_ref(uint32_t E, const char* spec, char* outSpec, size_t outLen, uint32_t* N, uint32_t* maxExp) {
  return guarded([&] {
    FFTConfig fft = FFTConfig::bestFit(E, spec ? spec : "", 1.0, true);
    if (outSpec && outLen) { snprintf(outSpec, outLen, "%s", fft.spec().c_str()); }
    if (N) { *N = fft.size(); }
    if (maxExp) { *maxExp = fft.refMaxExp(); }
  });
}

int owl_fft_tune_load(const char* path) {
  return guarded([&] {
    if (!path) { throw std::invalid_argument("tune file path"); }
    FftTune::set(FftTune::read(path));
  });
}

int owl_fft_tune_save(const char* path) {
  return guarded([&] {
    if (!path) { throw std::invalid_argument("tune file path"); }
    FftTune::write(path, FftTune::get());
  });
}

int owl_fft_tune_clear(void) {
  return guarded([&] { FftTune::set({}); });
}

int owl_fft_tune_add(double cost, const char* spec, int* updated) {
  return guarded([&] {
    if (!spec) { throw std::invalid_argument("spec"); }
    std::vector<FftTuneEntry> list = FftTune::get();
    const bool changed = FftTune::update(list, {cost, FFTConfig{spec}});
    if (changed) { FftTune::set(std::move(list)); }
    if (updated) { *updated = changed; }
  });
}

int owl_fft_tune_entry(uint32_t i, double* cost, char* spec, size_t specLen, uint32_t* maxExp, int* found) {
  return guarded([&] {
    std::vector<FftTuneEntry> list = FftTune::get();
    *found = i < list.size();
    if (!*found) { return; }
    if (cost) { *cost = list[i].cost; }
    if (spec && specLen) { snprintf(spec, specLen, "%s", list[i].fft.spec().c_str()); }
    if (maxExp) { *maxExp = list[i].fft.maxExp(); }
  });
}

int owl_gpu_create(uint32_t E, const char* spec, int flags, uint32_t lineBlock, owl_gpu* out) {
  *out = nullptr;
  return guarded([&] {
    FFTConfig fft = FFTConfig::bestFit(E, spec ? spec : "");
    GpuOptions opt;
    opt.longCarry = flags & OWL_LONG_CARRY;
    opt.generic = flags & OWL_GENERIC;
    opt.colOneRegion = flags & OWL_COL_ONE_REGION;
    opt.colOnePair = flags & OWL_COL_ONE_PAIR;
    opt.lineBlock = lineBlock;
    *out = new owl_gpu_s{new Gpu(E, fft, opt)};
  });
}

int owl_gpu_destroy(owl_gpu g) {
  return guarded([&] {
    if (g) {
      delete g->gpu;
      delete g;
    }
  });
}

void fillInfo(const Gpu& G, owl_gpu_info_t* info);

int owl_gpu_info(owl_gpu g, owl_gpu_info_t* info) {
  return guarded([&] { fillInfo(*g->gpu, info); });
}

void fillInfo(const Gpu& G, owl_gpu_info_t* info) {
  {
    memset(info, 0, sizeof(*info));
    info->E = G.E;
    info->N = G.N;
    info->W = G.W;
    info->H = G.H;
    info->R = G.R;
    info->nb = G.nb;
    info->rowThreads = G.rowT;
    info->colThreads = G.colT;
    info->maxExp = G.fft.maxExp();
    info->longCarry = G.useLongCarry;
    snprintf(info->spec, sizeof(info->spec), "%s", G.fft.spec().c_str());
    snprintf(info->rowKernel, sizeof(info->rowKernel), "%s", G.rowKernel.c_str());
    snprintf(info->colKernel, sizeof(info->colKernel), "%s", G.colKernel.c_str());
  }
}

int owl_gpu_write_words(owl_gpu g, int b, const uint32_t* words, uint32_t n) {
  return guarded([&] { g->gpu->writeIn(buf(b), Words(words, words + n)); });
}

int owl_gpu_read_words(owl_gpu g, int b, uint32_t* words, uint32_t n, int* isZero) {
  return guarded([&] {
    Words w = g->gpu->readData(buf(b));
    *isZero = w.empty();
    if (!w.empty()) {
      if (w.size() != n) { throw std::invalid_argument("word count mismatch"); }
      memcpy(words, w.data(), n * 4);
    }
  });
}

int owl_gpu_write_digits(owl_gpu g, int b, const int32_t* digits, uint32_t n) {
  return guarded([&] { g->gpu->writeDigits(buf(b), std::vector<int>(digits, digits + n)); });
}

int owl_gpu_read_digits(owl_gpu g, int b, int32_t* digits, uint32_t n) {
  return guarded([&] {
    std::vector<int> d = g->gpu->readDigits(buf(b));
    if (d.size() != n) { throw std::invalid_argument("digit count mismatch"); }
    memcpy(digits, d.data(), n * 4);
  });
}

int owl_gpu_read_digits_noalloc(owl_gpu g, int b, int32_t* digits, uint32_t n) {
  return guarded([&] {
    if (n != g->gpu->N) { throw std::invalid_argument("digit count mismatch"); }
    g->gpu->readDigitsNoAlloc(buf(b), digits);
  });
}

int owl_gpu_res64(owl_gpu g, int b, uint64_t* res) {
  return guarded([&] { *res = g->gpu->dataResidue(buf(b)); });
}

int owl_gpu_square_loop(owl_gpu g, int out, int in, uint32_t from, uint32_t to, int mul3Last, int ll) {
  return guarded([&] {
    if (from >= to) { throw std::invalid_argument("empty iteration range"); }
    g->gpu->squareLoop(buf(out), buf(in), from, to, mul3Last != 0, ll != 0);
    g->gpu->finish();
  });
}

int owl_gpu_mod_mul(owl_gpu g, int ioA, int inB, int mul3) {
  return guarded([&] {
    g->gpu->modMul(buf(ioA), buf(inB), mul3 != 0);
    g->gpu->finish();
  });
}

int owl_gpu_do_check(owl_gpu g, uint32_t blockSize, int* ok) {
  return guarded([&] { *ok = g->gpu->doCheck(blockSize); });
}

int owl_gpu_is_equal(owl_gpu g, int a, int b, int* equal) {
  return guarded([&] { *equal = g->gpu->isEqual(buf(a), buf(b)); });
}

int owl_gpu_copy(owl_gpu g, int dst, int src) {
  return guarded([&] {
    g->gpu->copy(buf(dst), buf(src));
    g->gpu->finish();
  });
}

int owl_gpu_read_roe(owl_gpu g, uint32_t* n, double* maxRoe, double* mean, double* sd) {
  return guarded([&] {
    RoeInfo r = g->gpu->readROE();
    *n = r.N;
    *maxRoe = r.max;
    *mean = r.mean;
    *sd = r.sd;
  });
}

int owl_gpu_measure_roe(owl_gpu g, uint32_t iters, uint32_t warmup, uint32_t* n, double* maxRoe, double* mean,
                        double* sd, double* z) {
  return guarded([&] {
    RoeInfo r = g->gpu->measureROE(iters, warmup);
    *n = r.N;
    *maxRoe = r.max;
    *mean = r.mean;
    *sd = r.sd;
    // Gumbel fit by moments (RoeInfo, src/Gpu.h:62-81): beta = sd*sqrt(6)/pi, mu = mean - gamma*beta,
    // Z = (0.5 - mu) / beta
    const double beta = r.sd * 0.779696801233676, mu = r.mean - beta * 0.577215664901533;
    *z = (r.N && beta > 0) ? (0.5 - mu) / beta : 0.0;
  });
}

int owl_gpu_prp(owl_gpu g, uint32_t iters, uint32_t blockSize, uint32_t checkStep, uint64_t* res, uint32_t resStep,
                uint32_t resCap, owl_prp_result_t* out) {
  return guarded([&] {
    std::vector<u64> at;
    PRPResult r = g->gpu->isPrimePRP(iters, blockSize, checkStep, res ? &at : nullptr, res ? resStep : 0);
    for (size_t i = 0; i < at.size() && i < resCap; ++i) { res[i] = at[i]; }
    out->res64 = r.res64;
    out->k = r.k;
    out->nErrors = r.nErrors;
    out->nChecks = r.nChecks;
    out->isPrime = r.isPrime;
    out->maxRoe = r.maxRoe;
    out->secsPerIt = r.secsPerIt;
  });
}

int owl_gpu_ll(owl_gpu g, uint32_t iters, uint64_t* res, uint32_t resStep, uint32_t resCap, owl_ll_result_t* out) {
  return guarded([&] {
    std::vector<u64> at;
    LLResult r = g->gpu->isPrimeLL(iters, res ? &at : nullptr, res ? resStep : 0);
    for (size_t i = 0; i < at.size() && i < resCap; ++i) { res[i] = at[i]; }
    out->res64 = r.res64;
    out->k = r.k;
    out->isPrime = r.isPrime;
    out->secsPerIt = r.secsPerIt;
  });
}

int owl_gpu_prp_saved(owl_gpu g, const char* dir, uint32_t nSavefiles, uint32_t iters, uint32_t blockSize,
                      uint32_t checkStep, uint64_t* res, uint32_t resStep, uint32_t resCap, owl_prp_result_t* out) {
  return guarded([&] {
    if (!dir) { throw std::invalid_argument("savefile directory"); }
    Saver<PRPState> saver(dir, g->gpu->E, blockSize, nSavefiles);
    std::vector<u64> at;
    PRPResult r = g->gpu->isPrimePRP(iters, blockSize, checkStep, res ? &at : nullptr, res ? resStep : 0, &saver);
    for (size_t i = 0; i < at.size() && i < resCap; ++i) { res[i] = at[i]; }
    out->res64 = r.res64;
    out->k = r.k;
    out->nErrors = r.nErrors;
    out->nChecks = r.nChecks;
    out->isPrime = r.isPrime;
    out->maxRoe = r.maxRoe;
    out->secsPerIt = r.secsPerIt;
  });
}

int owl_gpu_ll_saved(owl_gpu g, const char* dir, uint32_t nSavefiles, uint32_t saveStep, uint32_t iters, uint64_t* res,
                     uint32_t resStep, uint32_t resCap, owl_ll_result_t* out) {
  return guarded([&] {
    if (!dir) { throw std::invalid_argument("savefile directory"); }
    Saver<LLState> saver(dir, g->gpu->E, 1000, nSavefiles);
    std::vector<u64> at;
    LLResult r = g->gpu->isPrimeLL(iters, res ? &at : nullptr, res ? resStep : 0, &saver, saveStep);
    for (size_t i = 0; i < at.size() && i < resCap; ++i) { res[i] = at[i]; }
    out->res64 = r.res64;
    out->k = r.k;
    out->isPrime = r.isPrime;
    out->secsPerIt = r.secsPerIt;
  });
}

int owl_savefile_write_prp(const char* path, uint32_t E, uint32_t k, uint32_t blockSize, uint64_t res64,
                           uint32_t nErrors, double elapsed, const uint32_t* check, uint32_t n) {
  return guarded([&] {
    if (!path || !check || !E || n != nWords(E)) { throw std::invalid_argument("savefile arguments"); }
    writePRPState(path, {E, k, blockSize, res64, Words(check, check + n), nErrors, elapsed});
  });
}

int owl_savefile_read_prp(const char* path, uint32_t* E, uint32_t* k, uint32_t* blockSize, uint64_t* res64,
                          uint32_t* nErrors, double* elapsed, uint32_t* check, uint32_t cap) {
  return guarded([&] {
    if (!path) { throw std::invalid_argument("savefile path"); }
    PRPState s = readPRPState(path);
    if (check && cap < s.check.size()) { throw std::invalid_argument("check buffer too small"); }
    *E = s.exponent, *k = s.k, *blockSize = s.blockSize, *res64 = s.res64, *nErrors = s.nErrors, *elapsed = s.elapsed;
    if (check) { std::copy(s.check.begin(), s.check.end(), check); }
  });
}

int owl_savefile_write_ll(const char* path, uint32_t E, uint32_t k, double elapsed, const uint32_t* data, uint32_t n) {
  return guarded([&] {
    if (!path || !data || !E || n != nWords(E)) { throw std::invalid_argument("savefile arguments"); }
    writeLLState(path, {E, k, Words(data, data + n), elapsed});
  });
}

int owl_savefile_read_ll(const char* path, uint32_t* E, uint32_t* k, double* elapsed, uint32_t* data, uint32_t cap) {
  return guarded([&] {
    if (!path) { throw std::invalid_argument("savefile path"); }
    LLState s = readLLState(path);
    if (data && cap < s.data.size()) { throw std::invalid_argument("data buffer too small"); }
    *E = s.exponent, *k = s.k, *elapsed = s.elapsed;
    if (data) { std::copy(s.data.begin(), s.data.end(), data); }
  });
}

int owl_worktodo_parse(const char* line, int* kind, uint32_t* exponent, char* aid, size_t aidLen, uint32_t* squarings) {
  return guarded([&] {
    WorkTask t = parseWorktodoLine(line ? line : "");
    *kind = t.kind;
    if (exponent) { *exponent = t.exponent; }
    if (aid && aidLen) { snprintf(aid, aidLen, "%s", t.aid.c_str()); }
    if (squarings) { *squarings = t.squarings; }
  });
}

int owl_worktodo_best(const char* const* lines, uint32_t n, int32_t* index) {
  return guarded([&] {
    std::vector<std::string> v;
    for (uint32_t i = 0; i < n; ++i) { v.push_back(lines[i] ? lines[i] : ""); }
    *index = bestWorktodoLine(v);
  });
}

int owl_prp_type1(const uint32_t* words, uint32_t n, uint32_t E, uint64_t* r64, char* res2048, size_t len) {
  return guarded([&] {
    if (!words || n != nWords(E)) { throw std::invalid_argument("type-1 residue: words"); }
    Words q = div9(Words(words, words + n), E);
    if (r64) { *r64 = res64(q); }
    if (res2048) {
      if (len < 513) { throw std::invalid_argument("res2048 buffer too small"); }
      snprintf(res2048, len, "%s", res2048Hex(q).c_str());
    }
  });
}

int owl_result_json(int kind, uint32_t E, int isPrime, uint64_t r64, const char* res2048, uint32_t fftLength,
                    uint32_t nErrors, const char* aid, const char* user, const char* uid, const char* programName,
                    const char* programVersion, char* out, size_t outLen) {
  return guarded([&] {
    ResultInfo r;
    r.kind = WorkTask::Kind(kind);
    r.exponent = E;
    r.isPrime = isPrime != 0;
    r.res64 = r64;
    r.res2048 = res2048 ? res2048 : "";
    r.fftLength = fftLength;
    r.nErrors = nErrors;
    r.aid = aid ? aid : "";
    r.user = user ? user : "";
    r.uid = uid ? uid : "";
    if (programName) { r.programName = programName; }
    if (programVersion) { r.programVersion = programVersion; }
    const std::string s = resultJson(r);
    if (!out || outLen <= s.size()) { throw std::invalid_argument("result buffer too small"); }
    snprintf(out, outLen, "%s", s.c_str());
  });
}

int owl_result_json_ex(const owl_result_t* in, char* out, size_t outLen) {
  return guarded([&] {
    if (!in) { throw std::invalid_argument("result"); }
    auto sv = [](const char* p) { return std::string(p ? p : ""); };
    ResultInfo r;
    r.kind = WorkTask::Kind(in->kind);
    r.exponent = in->exponent;
    r.isPrime = in->isPrime != 0;
    r.res64 = in->res64;
    r.res2048 = sv(in->res2048);
    r.fftLength = in->fftLength;
    r.nErrors = in->nErrors;
    r.proofPower = in->proofPower;
    r.proofMd5 = sv(in->proofMd5);
    r.sha3Hash = sv(in->sha3Hash);
    r.squarings = in->squarings;
    r.aid = sv(in->aid);
    r.user = sv(in->user);
    r.uid = sv(in->uid);
    if (in->programName) { r.programName = in->programName; }
    if (in->programVersion) { r.programVersion = in->programVersion; }
    const std::string s = resultJson(r);
    if (!out || outLen <= s.size()) { throw std::invalid_argument("result buffer too small"); }
    snprintf(out, outLen, "%s", s.c_str());
  });
}

int owl_proof_points(uint32_t E, uint32_t power, uint32_t* points, uint32_t cap, uint32_t* n) {
  return guarded([&] {
    std::vector<u32> p = proofPoints(E, power);
    if (points && cap < p.size()) { throw std::invalid_argument("points buffer too small"); }
    if (points) { std::copy(p.begin(), p.end(), points); }
    *n = u32(p.size());
  });
}

int owl_gpu_prp_proof(owl_gpu g, const char* residueDir, uint32_t power, uint32_t blockSize, uint32_t checkStep,
                      const char* proofPath, owl_prp_result_t* out) {
  return guarded([&] {
    if (!residueDir || !proofPath) { throw std::invalid_argument("proof paths"); }
    Gpu& G = *g->gpu;
    ProofStore store(residueDir, G.E, power);
    PRPResult r = G.isPrimePRP(G.E, blockSize, checkStep, nullptr, 0, nullptr, &store);
    PrpProof proof = G.makeProof(store);
    proof.save(proofPath);
    if (!G.verifyProof(PrpProof::load(proofPath))) { throw GpuError("proof verification failed"); }
    out->res64 = r.res64;
    out->k = r.k;
    out->nErrors = r.nErrors;
    out->nChecks = r.nChecks;
    out->isPrime = r.isPrime;
    out->maxRoe = r.maxRoe;
    out->secsPerIt = r.secsPerIt;
  });
}

int owl_gpu_proof_verify(owl_gpu g, const char* proofPath, int* ok, int* isPrime) {
  return guarded([&] {
    if (!proofPath) { throw std::invalid_argument("proof path"); }
    PrpProof p = PrpProof::load(proofPath);
    *ok = g->gpu->verifyProof(p);
    if (isPrime) { *isPrime = equals9(p.B); }
  });
}

int owl_gpu_cert(owl_gpu g, const char* certPath, uint32_t squarings, int removeFile, char* sha3Hex, size_t len,
                 uint64_t* r64) {
  return guarded([&] {
    if (!certPath || !sha3Hex || len < 65) { throw std::invalid_argument("cert: path or hash buffer"); }
    const u32 E = g->gpu->E;
    const size_t nBytes = (E - 1) / 8 + 1;
    Words start(nWords(E), 0);
    {
      std::unique_ptr<FILE, int (*)(FILE*)> f{fopen(certPath, "rb"), fclose};
      if (!f) { throw std::invalid_argument(std::string("cannot read CERT start file ") + certPath); }
      if (fread(start.data(), 1, nBytes, f.get()) != nBytes) {
        throw std::invalid_argument(std::string("short CERT start file ") + certPath);
      }
    }
    Words x = g->gpu->isCERT(start, squarings);
    // the reference prints the digest's four little-endian u64 from the last to the first
    const std::array<u64, 4> h = proofHash(E, x);
    snprintf(sha3Hex, len, "%016llx%016llx%016llx%016llx", (unsigned long long)h[3], (unsigned long long)h[2],
             (unsigned long long)h[1], (unsigned long long)h[0]);
    if (r64) { *r64 = res64(x); }
    if (removeFile) { std::remove(certPath); }
  });
}

int owl_proof_md5(const char* path, char* md5Hex32, size_t len) {
  return guarded([&] {
    if (!path || !md5Hex32 || len < 33) { throw std::invalid_argument("md5: path or buffer"); }
    snprintf(md5Hex32, len, "%s", md5File(path).c_str());
  });
}

int owl_gpu_exp_mul(owl_gpu g, const uint32_t* A, uint64_t h, const uint32_t* B, int squareB, uint32_t* out,
                    uint32_t n) {
  return guarded([&] {
    const u32 E = g->gpu->E;
    if (!A || !B || !out || n != nWords(E)) { throw std::invalid_argument("exp_mul words"); }
    Words r = g->gpu->expMul(Words(A, A + n), h, Words(B, B + n), squareB != 0);
    std::copy(r.begin(), r.end(), out);
  });
}

int owl_sha3_256(const void* data, size_t size, uint8_t* out) {
  return guarded([&] {
    if ((!data && size) || !out) { throw std::invalid_argument("sha3 arguments"); }
    std::array<u64, 4> d = Sha3_256{}.update(data, size).finish();
    std::memcpy(out, d.data(), 32);
  });
}

int owl_trig_root1(uint32_t n, uint64_t k, double* c, double* s) {
  return guarded([&] {
    if (!n) { throw std::invalid_argument("root1: n = 0"); }
    auto [cc, ss] = root1(n, k);
    *c = cc;
    *s = ss;
  });
}

int owl_crc32(const void* data, size_t size, uint32_t* crc) {
  return guarded([&] {
    if (!data && size) { throw std::invalid_argument("crc32 data"); }
    *crc = crc32(data, size);
  });
}

int owl_gpu_time_prp(owl_gpu g, uint32_t iters, uint32_t warmup, uint32_t blockSize, double* us) {
  return guarded([&] { *us = g->gpu->timePRP(iters, warmup, blockSize); });
}

int owl_gpu_time_squares(owl_gpu g, uint32_t iters, uint32_t warmup, int useGraph, double* us) {
  return guarded([&] { *us = g->gpu->timeSquares(iters, warmup, useGraph != 0); });
}

int owl_gpu_time_kernels(owl_gpu g, uint32_t iters, double* usCol, double* usRow) {
  return guarded([&] {
    Gpu& G = *g->gpu;
    G.rowLeadIn(Gpu::DATA);
    owl_queue q = G.queue();
    std::vector<owl_event> ev(2 * iters + 1);
    for (auto& e : ev) { owlCheck(owl_hip_event_create(&e), "event"); }
    owlCheck(owl_hip_event_record(ev[0], q), "event");
    for (uint32_t i = 0; i < iters; ++i) {
      G.colSquare();
      owlCheck(owl_hip_event_record(ev[2 * i + 1], q), "event");
      G.rowFused(false, OWL_ROE_SIZE);
      owlCheck(owl_hip_event_record(ev[2 * i + 2], q), "event");
    }
    double col = 0, row = 0;
    for (uint32_t i = 0; i < iters; ++i) {
      float a = 0, b = 0;
      owlCheck(owl_hip_event_elapsed_ms(ev[2 * i], ev[2 * i + 1], &a), "elapsed");
      owlCheck(owl_hip_event_elapsed_ms(ev[2 * i + 1], ev[2 * i + 2], &b), "elapsed");
      col += a;
      row += b;
    }
    for (auto& e : ev) { owl_hip_event_destroy(e); }
    *usCol = col * 1e3 / iters;
    *usRow = row * 1e3 / iters;
    G.colSquare();
    G.rowCarryOut(Gpu::DATA, false, false, OWL_ROE_SIZE);
    G.carryAll(Gpu::DATA);
    G.finish();
  });
}

int owl_gpu_time_copy(owl_gpu g, uint32_t iters, double* us) {
  return guarded([&] {
    Gpu& G = *g->gpu;
    owl_queue q = G.queue();
    size_t bytes = size_t(G.ND) * 16;
    owlCheck(owl_hip_copy(q, G.job().freq[1], G.job().freq[0], bytes), "copy");
    owl_event a, b;
    owlCheck(owl_hip_event_create(&a), "event");
    owlCheck(owl_hip_event_create(&b), "event");
    owlCheck(owl_hip_event_record(a, q), "event");
    for (uint32_t i = 0; i < iters; ++i) {
      owlCheck(owl_hip_copy(q, G.job().freq[1 + (i & 1)], G.job().freq[0], bytes), "copy");
    }
    owlCheck(owl_hip_event_record(b, q), "event");
    float ms = 0;
    owlCheck(owl_hip_event_elapsed_ms(a, b, &ms), "elapsed");
    owl_hip_event_destroy(a);
    owl_hip_event_destroy(b);
    *us = ms * 1e3 / iters;
  });
}

int owl_tune_run(uint32_t E, const char* spec, uint32_t iters, owl_tune_entry_t* out, uint32_t cap, uint32_t* n) {
  return guarded([&] {
    std::vector<TuneResult> rs = tuneShape(E, spec ? spec : "", iters, std::max<u32>(10, iters / 10));
    u32 i = 0;
    for (; i < cap && i < rs.size(); ++i) {
      const TuneResult& r = rs[i];
      owl_tune_entry_t& o = out[i];
      o.width = r.choice.width;
      o.lineBlock = r.choice.lineBlock;
      o.flags = (r.choice.generic ? OWL_GENERIC : 0) | (r.choice.colOneRegion ? OWL_COL_ONE_REGION : 0) |
                (r.choice.colOnePair ? OWL_COL_ONE_PAIR : 0);
      o.usPerIter = r.choice.us;
      o.W = r.W, o.H = r.H, o.R = r.R;
      o.longCarry = r.longCarry;
      std::snprintf(o.rowKernel, sizeof o.rowKernel, "%s", r.rowKernel.c_str());
      std::snprintf(o.colKernel, sizeof o.colKernel, "%s", r.colKernel.c_str());
    }
    *n = i;
  });
}

int owl_tune_batch(const uint32_t* exponents, uint32_t n, const char* spec, uint32_t iters, owl_tune_entry_t* out,
                   uint32_t cap, uint32_t* nOut) {
  return guarded([&] {
    if (!exponents || !n || !iters) { throw std::invalid_argument("batch tuning arguments"); }
    std::vector<TuneResult> rs = tuneBatch(std::vector<u32>(exponents, exponents + n), spec ? spec : "", iters,
                                           std::max<u32>(5, iters / 4));
    u32 i = 0;
    for (; i < cap && i < rs.size(); ++i) {
      const TuneResult& r = rs[i];
      owl_tune_entry_t& o = out[i];
      o.width = r.choice.width;
      o.lineBlock = r.choice.lineBlock;
      o.flags = 0;
      o.usPerIter = r.choice.us;
      o.W = r.W, o.H = r.H, o.R = r.R;
      o.longCarry = r.longCarry;
      std::snprintf(o.rowKernel, sizeof o.rowKernel, "%s", r.rowKernel.c_str());
      std::snprintf(o.colKernel, sizeof o.colKernel, "%s", r.colKernel.c_str());
    }
    *nOut = i;
  });
}

int owl_tune_batch_clear(void) {
  return guarded([&] { BatchTuneTable::clear(); });
}

int owl_tune_set(uint32_t N, uint32_t width, uint32_t lineBlock, int flags, double us) {
  return guarded([&] {
    TuneChoice c;
    c.width = width;
    c.lineBlock = lineBlock;
    c.generic = flags & OWL_GENERIC;
    c.colOneRegion = flags & OWL_COL_ONE_REGION;
    c.colOnePair = flags & OWL_COL_ONE_PAIR;
    c.us = us;
    TuneTable::set(N, c);
  });
}

int owl_tune_get(uint32_t N, uint32_t* width, uint32_t* lineBlock, int* flags, double* us, int* found) {
  return guarded([&] {
    TuneChoice c;
    *found = TuneTable::get(N, &c);
    *width = c.width;
    *lineBlock = c.lineBlock;
    *flags = (c.generic ? OWL_GENERIC : 0) | (c.colOneRegion ? OWL_COL_ONE_REGION : 0) | (c.colOnePair ? OWL_COL_ONE_PAIR : 0);
    *us = c.us;
  });
}

int owl_tune_clear(void) {
  return guarded([&] { TuneTable::clear(); });
}

int owl_tune_load(const char* path) {
  return guarded([&] { TuneTable::load(path); });
}

int owl_tune_save(const char* path) {
  return guarded([&] { TuneTable::save(path); });
}

int owl_server_create(uint64_t memBudget, uint32_t maxJobs, uint32_t blockSize, uint32_t checkStep, const char* spec,
                      int flags, owl_server* out) {
  *out = nullptr;
  return guarded([&] {
    Server::Options o;
    o.memBudget = memBudget;
    o.maxJobs = maxJobs;
    o.blockSize = blockSize;
    o.checkStep = checkStep;
    o.spec = spec ? spec : "";
    o.gpu.longCarry = flags & OWL_LONG_CARRY;
    o.gpu.generic = flags & OWL_GENERIC;
    o.gpu.colOneRegion = flags & OWL_COL_ONE_REGION;
    o.gpu.colOnePair = flags & OWL_COL_ONE_PAIR;
    *out = new owl_server_s{new Server(o), {}};
  });
}

int owl_server_destroy(owl_server s) {
  return guarded([&] {
    if (s) {
      delete s->server;
      delete s;
    }
  });
}

int owl_server_submit(owl_server s, uint32_t E, int ll, uint32_t iters, uint32_t* id) {
  return guarded([&] { *id = s->server->submit(E, ll != 0, iters); });
}

int owl_server_arm_fault(owl_server s, uint32_t id, uint32_t k) {
  return guarded([&] { s->server->armFault(id, k); });
}

int owl_server_step(owl_server s, uint32_t* nDone) {
  return guarded([&] {
    u32 n = s->server->step();
    for (ServeResult& r : s->server->take()) { s->ready.push_back(r); }
    if (nDone) { *nDone = n; }
  });
}

int owl_server_take(owl_server s, owl_serve_result_t* out, uint32_t cap, uint32_t* n) {
  return guarded([&] {
    u32 i = 0;
    for (; i < cap && !s->ready.empty(); ++i) {
      const ServeResult& r = s->ready.front();
      owl_serve_result_t& o = out[i];
      o.id = r.id;
      o.E = r.E;
      o.iters = r.iters;
      o.ll = r.ll;
      o.res64 = r.res64;
      o.isPrime = r.isPrime;
      o.nErrors = r.nErrors;
      o.nChecks = r.nChecks;
      o.maxRoe = r.maxRoe;
      o.tSubmit = r.tSubmit;
      o.tAdmit = r.tAdmit;
      o.tDone = r.tDone;
      o.status = r.status;
      o.res64Type1 = r.res64Type1;
      snprintf(o.res2048, sizeof o.res2048, "%s", r.res2048.c_str());
      s->ready.pop_front();
    }
    *n = i;
  });
}

int owl_server_info(owl_server s, owl_serve_info_t* info) {
  return guarded([&] {
    ServeInfo x = s->server->info();
    info->waiting = x.waiting;
    info->active = x.active;
    info->groups = x.groups;
    info->done = x.done;
    info->rounds = x.rounds;
    info->bytesInUse = x.bytesInUse;
    info->budget = x.budget;
    info->iterations = x.iterations;
    info->wordIterations = x.wordIterations;
  });
}

int owl_server_now(owl_server s, double* secs) {
  return guarded([&] { *secs = s->server->now(); });
}

int owl_batch_create(const uint32_t* exponents, uint32_t n, const char* spec, int flags, owl_batch* out) {
  *out = nullptr;
  return guarded([&] {
    GpuOptions opt;
    opt.longCarry = flags & OWL_LONG_CARRY;
    opt.generic = flags & OWL_GENERIC;
    opt.colOneRegion = flags & OWL_COL_ONE_REGION;
    opt.colOnePair = flags & OWL_COL_ONE_PAIR;
    *out = new owl_batch_s{new GpuBatch(std::vector<u32>(exponents, exponents + n), spec ? spec : "", opt)};
  });
}

int owl_batch_destroy(owl_batch b) {
  return guarded([&] {
    if (b) {
      delete b->batch;
      delete b;
    }
  });
}

int owl_batch_info(owl_batch b, uint32_t* nJobs, uint32_t* nGroups, uint64_t* totalWords) {
  return guarded([&] {
    *nJobs = u32(b->batch->size());
    *nGroups = u32(b->batch->groups.size());
    *totalWords = b->batch->totalWords();
  });
}

int owl_gpu_arm_fault(owl_gpu g, uint32_t k) {
  return guarded([&] { g->gpu->armFault(k); });
}

int owl_batch_arm_fault(owl_batch b, uint32_t job, uint32_t k) {
  return guarded([&] {
    if (job >= b->batch->size()) { throw std::invalid_argument("job index"); }
    b->batch->jobs[job]->armFault(k);
  });
}

int owl_batch_job_info(owl_batch b, uint32_t job, owl_gpu_info_t* info) {
  return guarded([&] {
    if (job >= b->batch->size()) { throw std::invalid_argument("job index"); }
    fillInfo(*b->batch->jobs[job], info);
    // the column kernel of the job's group launches (several pairs per group in a large group)
    const GpuBatch::Group& g = b->batch->groupOf(job);
    snprintf(info->colKernel, sizeof(info->colKernel), "%s", g.lead->colKernelFor(u32(g.index.size())).c_str());
  });
}

int owl_batch_prp(owl_batch b, uint32_t iters, uint32_t blockSize, uint32_t checkStep, uint64_t* res, uint32_t resStep,
                  uint32_t resCap, owl_prp_result_t* out) {
  return guarded([&] {
    std::vector<std::vector<u64>> at;
    auto r = b->batch->prp(iters, blockSize, checkStep, res ? &at : nullptr, res ? resStep : 0);
    for (size_t i = 0; i < r.size(); ++i) {
      if (res) {
        for (size_t k = 0; k < at[i].size() && k < resCap; ++k) { res[i * resCap + k] = at[i][k]; }
      }
      out[i].res64 = r[i].res64;
      out[i].k = r[i].k;
      out[i].nErrors = r[i].nErrors;
      out[i].nChecks = r[i].nChecks;
      out[i].isPrime = r[i].isPrime;
      out[i].maxRoe = r[i].maxRoe;
      out[i].secsPerIt = r[i].secsPerIt;
    }
  });
}

int owl_batch_ll(owl_batch b, uint32_t iters, uint64_t* res, uint32_t resStep, uint32_t resCap, owl_ll_result_t* out) {
  return guarded([&] {
    std::vector<std::vector<u64>> at;
    auto r = b->batch->ll(iters, res ? &at : nullptr, res ? resStep : 0);
    for (size_t i = 0; i < r.size(); ++i) {
      if (res) {
        for (size_t k = 0; k < at[i].size() && k < resCap; ++k) { res[i * resCap + k] = at[i][k]; }
      }
      out[i].res64 = r[i].res64;
      out[i].k = r[i].k;
      out[i].isPrime = r[i].isPrime;
      out[i].secsPerIt = r[i].secsPerIt;
    }
  });
}

int owl_batch_time_squares(owl_batch b, uint32_t iters, uint32_t warmup, int useGraph, double* us) {
  return guarded([&] { *us = b->batch->timeSquares(iters, warmup, useGraph != 0); });
}

int owl_gpu_freq_io(owl_gpu g, int idx, double* host, uint32_t n, int write) {
  return guarded([&] {
    Gpu& G = *g->gpu;
    if (idx < 0 || idx > 2 || n != 2 * G.ND) { throw std::invalid_argument("freq_io: bad buffer or size"); }
    double2* p = G.job().freq[idx];
    if (write) {
      owlCheck(owl_hip_write(G.queue(), p, host, size_t(n) * 8, 1), "write");
    } else {
      owlCheck(owl_hip_read(G.queue(), host, p, size_t(n) * 8, 1), "read");
    }
  });
}

int owl_gpu_pass(owl_gpu g, int pass, int arg) {
  return guarded([&] {
    Gpu& G = *g->gpu;
    switch (pass) {
      case OWL_PASS_LEADIN: G.rowLeadIn(buf(arg)); break;
      case OWL_PASS_SQUARE: G.colSquare(); break;
      case OWL_PASS_FUSED: G.rowFused(arg != 0, G.nextRoeSlot()); break;
      case OWL_PASS_CARRYOUT:
        G.rowCarryOut(buf(arg), false, false, G.nextRoeSlot());
        G.carryAll(buf(arg));
        break;
      case OWL_PASS_SPECTRUM: G.colFwdSpec(); break;
      case OWL_PASS_MUL: G.colMul(); break;
      default: throw std::invalid_argument("unknown pass");
    }
    G.finish();
    if (G.flagsRead() & 4) { throw GpuError("carry hand-off timed out"); }
  });
}

}  // extern "C"
