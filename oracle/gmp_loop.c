/* ORACLE -- test infrastructure only. Used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker; never part of the product path.
 *
 * Exact PRP / LL iteration mod M = 2^E - 1 with GMP (libgmp 6.2.1, /opt/conda):
 *   PRP: x0 = 3,  x <- x^2 mod M      (src/Gpu.cpp:1461-1583, x0 from src/Saver.cpp:160-162)
 *   LL:  x0 = 4,  x <- x^2 - 2 mod M  (src/Gpu.cpp:1586-1663, x0 from src/Saver.cpp:164-166)
 * res64 is the low 64 bits of the canonical residue x_k in [0, M) (src/Gpu.cpp:955-975),
 * the intermediate residue printed by the reference every log step.
 * The reduction x = (x mod 2^E) + (x >> E) is the textbook Mersenne fold.
 * Pinned by the Mersenne-prime known answers (PRP residue 9 at k = E, LL residue 0 at
 * k = E - 2) and by brute force with Python integers (tests/test_oracle.py).
 */
#include <gmp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static void fold(mpz_t x, const mpz_t M, uint32_t E, mpz_t t) {
  /* x < 2^(2E+1) here; two folds bring it below 2M, one subtraction below M */
  while (mpz_sizeinbase(x, 2) > E) {
    mpz_tdiv_q_2exp(t, x, E);
    mpz_tdiv_r_2exp(x, x, E);
    mpz_add(x, x, t);
  }
  if (mpz_cmp(x, M) >= 0) { mpz_sub(x, x, M); }
}

static uint64_t low64(const mpz_t x) {
  uint64_t r = 0;
  size_t n = mpz_size(x);
  if (n > 0) { r = mpz_getlimbn(x, 0); }
  if (sizeof(mp_limb_t) == 4 && n > 1) { r |= (uint64_t)mpz_getlimbn(x, 1) << 32; }
  return r;
}

static void toWords(const mpz_t x, uint32_t E, uint32_t* words) {
  uint32_t nw = (E - 1) / 32 + 1;
  memset(words, 0, nw * 4);
  size_t count = 0;
  mpz_export(words, &count, -1, 4, 0, 0, x);
}

static void fromWords(mpz_t x, uint32_t E, const uint32_t* words) {
  uint32_t nw = (E - 1) / 32 + 1;
  mpz_import(x, nw, -1, 4, 0, 0, words);
}

/* Runs iters iterations; res64 after every step-th iteration goes to out[0..cap).
 * Final residue words ((E-1)/32+1 u32) to finalWords if not NULL. Returns the number of res64 written. */
int oracle_run(uint32_t E, int ll, uint32_t iters, uint32_t step, uint64_t* out, uint32_t cap, uint32_t* finalWords) {
  mpz_t x, M, t;
  mpz_inits(x, M, t, NULL);
  mpz_set_ui(M, 1);
  mpz_mul_2exp(M, M, E);
  mpz_sub_ui(M, M, 1);
  mpz_set_ui(x, ll ? 4 : 3);
  uint32_t n = 0;
  for (uint32_t k = 1; k <= iters; ++k) {
    mpz_mul(x, x, x);
    if (ll) {
      if (mpz_cmp_ui(x, 2) < 0) { mpz_add(x, x, M); }
      mpz_sub_ui(x, x, 2);
    }
    fold(x, M, E, t);
    if (step && k % step == 0 && n < cap) { out[n++] = low64(x); }
  }
  if (finalWords) { toWords(x, E, finalWords); }
  mpz_clears(x, M, t, NULL);
  return (int)n;
}

/* out = a * b (* 3 if mul3) mod 2^E - 1 on packed words. */
void oracle_mulmod(uint32_t E, const uint32_t* a, const uint32_t* b, int mul3, uint32_t* out) {
  mpz_t x, y, M, t;
  mpz_inits(x, y, M, t, NULL);
  mpz_set_ui(M, 1);
  mpz_mul_2exp(M, M, E);
  mpz_sub_ui(M, M, 1);
  fromWords(x, E, a);
  fromWords(y, E, b);
  mpz_mul(x, x, y);
  if (mul3) { mpz_mul_ui(x, x, 3); }
  fold(x, M, E, t);
  toWords(x, E, out);
  mpz_clears(x, y, M, t, NULL);
}

/* out = a^(2^n) mod 2^E - 1 on packed words. */
void oracle_square_n(uint32_t E, const uint32_t* a, uint32_t n, uint32_t* out) {
  mpz_t x, M, t;
  mpz_inits(x, M, t, NULL);
  mpz_set_ui(M, 1);
  mpz_mul_2exp(M, M, E);
  mpz_sub_ui(M, M, 1);
  fromWords(x, E, a);
  for (uint32_t i = 0; i < n; ++i) {
    mpz_mul(x, x, x);
    fold(x, M, E, t);
  }
  toWords(x, E, out);
  mpz_clears(x, M, t, NULL);
}
