/* ORACLE -- test infrastructure only (golden fixtures and the CPU baseline leg of bench.py).
 * Usage: gmp_loop E iters [ll] [step] [full]
 *   step = 0: one JSON line with the elapsed seconds and the final res64;
 *   step > 0: "k res64" after every step-th iteration (flushed), then the JSON line.
 *   full = 1: start from the full-width value 3^(2^k0) mod M with k0 = ceil(log2 E) + 1
 *             instead of 3 (the squarings from 3 are cheap until the value fills E bits);
 *             k0 is not timed and k counts from k0.
 * Single thread; the same loop as oracle_run in gmp_loop.c. */
#include <gmp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

int oracle_run(uint32_t E, int ll, uint32_t iters, uint32_t step, uint64_t* out, uint32_t cap, uint32_t* finalWords);

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s E iters [ll] [step]\n", argv[0]);
    return 2;
  }
  uint32_t E = (uint32_t)strtoul(argv[1], NULL, 10);
  uint32_t iters = (uint32_t)strtoul(argv[2], NULL, 10);
  int ll = argc > 3 && atoi(argv[3]);
  uint32_t step = argc > 4 ? (uint32_t)strtoul(argv[4], NULL, 10) : 0;
  int full = argc > 5 && atoi(argv[5]);
  uint64_t res = 0;
  struct timespec a, b;
  if (full) {
    /* timed squarings of a full-width residue: skip the cheap start, then time iters */
    uint32_t k0 = 1;
    while ((1u << (k0 - 1)) < E) { ++k0; }
    mpz_t x, M, t;
    mpz_inits(x, M, t, NULL);
    mpz_set_ui(M, 1);
    mpz_mul_2exp(M, M, E);
    mpz_sub_ui(M, M, 1);
    mpz_set_ui(x, 3);
    for (uint32_t pass = 0; pass < 2; ++pass) {
      if (pass == 1) { clock_gettime(CLOCK_MONOTONIC, &a); }
      for (uint32_t k = 0; k < (pass ? iters : k0); ++k) {
        mpz_mul(x, x, x);
        while (mpz_sizeinbase(x, 2) > E) {
          mpz_tdiv_q_2exp(t, x, E);
          mpz_tdiv_r_2exp(x, x, E);
          mpz_add(x, x, t);
        }
        if (mpz_cmp(x, M) >= 0) { mpz_sub(x, x, M); }
      }
    }
    clock_gettime(CLOCK_MONOTONIC, &b);
    res = mpz_size(x) ? mpz_getlimbn(x, 0) : 0;
    mpz_clears(x, M, t, NULL);
    double secs = (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
    printf("{\"E\": %u, \"iters\": %u, \"k0\": %u, \"secs\": %.6f, \"res64\": \"%016llx\"}\n", E, iters, k0, secs,
           (unsigned long long)res);
    return 0;
  }
  clock_gettime(CLOCK_MONOTONIC, &a);
  if (step) {
    /* run in chunks of step so that progress appears as it happens */
    mpz_t x, M, t;
    mpz_inits(x, M, t, NULL);
    mpz_set_ui(M, 1);
    mpz_mul_2exp(M, M, E);
    mpz_sub_ui(M, M, 1);
    mpz_set_ui(x, ll ? 4 : 3);
    for (uint32_t k = 1; k <= iters; ++k) {
      mpz_mul(x, x, x);
      if (ll) {
        if (mpz_cmp_ui(x, 2) < 0) { mpz_add(x, x, M); }
        mpz_sub_ui(x, x, 2);
      }
      while (mpz_sizeinbase(x, 2) > E) {
        mpz_tdiv_q_2exp(t, x, E);
        mpz_tdiv_r_2exp(x, x, E);
        mpz_add(x, x, t);
      }
      if (mpz_cmp(x, M) >= 0) { mpz_sub(x, x, M); }
      if (k % step == 0) {
        uint64_t r = mpz_size(x) ? mpz_getlimbn(x, 0) : 0;
        printf("%u %016llx\n", k, (unsigned long long)r);
        fflush(stdout);
        res = r;
      }
    }
    mpz_clears(x, M, t, NULL);
  } else {
    oracle_run(E, ll, iters, iters, &res, 1, NULL);
  }
  clock_gettime(CLOCK_MONOTONIC, &b);
  double secs = (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
  printf("{\"E\": %u, \"iters\": %u, \"ll\": %d, \"secs\": %.6f, \"res64\": \"%016llx\"}\n", E, iters, ll, secs,
         (unsigned long long)res);
  return 0;
}
