// extern "C" entry points of the engine (declared in include/owl.h).
#include "owl.h"

#include "Gpu.h"

#include <cstring>
#include <exception>
#include <string>

struct owl_gpu_s {
  Gpu* gpu;
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

int owl_gpu_create(uint32_t E, const char* spec, int longCarry, uint32_t lineBlock, owl_gpu* out) {
  *out = nullptr;
  return guarded([&] {
    FFTConfig fft = FFTConfig::bestFit(E, spec ? spec : "");
    GpuOptions opt;
    opt.longCarry = longCarry != 0;
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

int owl_gpu_info(owl_gpu g, owl_gpu_info_t* info) {
  return guarded([&] {
    Gpu& G = *g->gpu;
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
  });
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
      G.rowFused(false, OWL_ROE_SIZE - 1);
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
    G.rowCarryOut(Gpu::DATA, false, false, OWL_ROE_SIZE - 1);
    G.carryAll(Gpu::DATA);
    G.finish();
  });
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
      case OWL_PASS_FUSED: G.rowFused(arg != 0, OWL_ROE_SIZE - 1); break;
      case OWL_PASS_CARRYOUT:
        G.rowCarryOut(buf(arg), false, false, OWL_ROE_SIZE - 1);
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
