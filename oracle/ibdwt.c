/* ORACLE -- test infrastructure only. Used by tests/ as the checker; never part of
 * the product path.
 *
 * A plain-C restatement of the reference's IBDWT squaring path, function by function:
 *   word layout        step/extra/isBigWord/bitlen   src/state.h:14-17
 *   packing            compactBits / expandBits       src/state.cpp:26-115
 *   res64 readout      bufResidue                     src/Gpu.cpp:955-975 (+ readResidue, src/cl/etc.cl:9-15)
 *   weights            weight / invWeight             src/Gpu.cpp:43-49   (w_k = 2^(extra(k)/N))
 *   rounding           RNDVAL round-to-nearest        src/cl/carryutil.cl:56-106
 *   carry              carryStep / full propagation   src/cl/carryutil.cl:108-149, src/cl/carry.cl:9-41,
 *                                                     src/cl/carryb.cl:5-32 (lead-out path: fully normalized)
 *   MUL3 / LL -2       carry.cl MUL3, LL carry-in      src/cl/carryutil.cl:93-104, src/cl/carry.cl:17
 * The transform itself is a straightforward mixed-radix complex FFT of length N in
 * long double over the real weighted digits (the reference's FFT factorization is
 * an implementation detail that does not change the integer result while the
 * round-off stays below 0.5). Pinned against GMP (res64 and full residues), the
 * reference's own state.cpp built in oracle/_ref, and closed forms
 * (tests/test_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef long double ld;
typedef struct { ld re, im; } cld;

static uint32_t stepOf(uint32_t N, uint32_t E) { return N - (E % N); }
uint32_t ibdwt_extra(uint32_t N, uint32_t E, uint32_t k) { return (uint32_t)((uint64_t)stepOf(N, E) * k % N); }
int ibdwt_is_big(uint32_t N, uint32_t E, uint32_t k) { return ibdwt_extra(N, E, k) + stepOf(N, E) < N; }
uint32_t ibdwt_bitlen(uint32_t N, uint32_t E, uint32_t k) { return E / N + ibdwt_is_big(N, E, k); }

/* Balanced digits -> canonical words ((E-1)/32+1). Returns 0, or -1 on a digit out of range. */
int ibdwt_compact(const int32_t* d, uint32_t N, uint32_t E, uint32_t* out) {
  uint32_t nw = (E - 1) / 32 + 1, pos = 0, have = 0;
  uint64_t acc = 0;
  int borrow = 0;
  for (uint32_t k = 0; k < N; ++k) {
    uint32_t b = ibdwt_bitlen(N, E, k);
    int64_t v = (int64_t)d[k] + borrow;
    borrow = 0;
    if (v < 0) { v += (int64_t)1 << b; borrow = -1; }
    if (v < 0 || v >= ((int64_t)1 << b)) { return -1; }
    acc |= (uint64_t)v << have;
    have += b;
    while (have >= 32) { out[pos++] = (uint32_t)acc; acc >>= 32; have -= 32; }
  }
  if (have) { out[pos++] = (uint32_t)acc; }
  if (pos != nw) { return -1; }
  for (uint32_t p = 0; borrow && p < nw; ++p) {
    int64_t v = (int64_t)out[p] + borrow;
    out[p] = (uint32_t)v;
    borrow = v < 0 ? -1 : 0;
  }
  return 0;
}

/* Canonical words -> balanced digits. */
void ibdwt_expand(const uint32_t* words, uint32_t N, uint32_t E, int32_t* d) {
  uint64_t acc = 0;
  uint32_t have = 0, next = 0;
  for (uint32_t k = 0; k < N; ++k) {
    uint32_t b = ibdwt_bitlen(N, E, k);
    if (have < b) { acc += (uint64_t)words[next++] << have; have += 32; }
    uint64_t low = acc & (((uint64_t)1 << b) - 1);
    int32_t v = (int32_t)((int64_t)(low << (64 - b)) >> (64 - b));
    acc >>= b;
    have -= b;
    if (v < 0) { acc += 1; }
    d[k] = v;
  }
  d[0] += (int32_t)acc;
}

/* res64 from the top 32 and bottom 32 digits (w[0..31] = digits N-32..N-1, w[32..63] = digits 0..31). */
uint64_t ibdwt_residue_from_ends(const int32_t* w, uint32_t N, uint32_t E) {
  int carry = 0;
  for (int i = 0; i < 32; ++i) { carry = (w[i] + carry < 0) ? -1 : 0; }
  uint64_t res = 0;
  uint32_t hasBits = 0;
  for (uint32_t k = 0; k < 32 && hasBits < 64; ++k) {
    uint32_t len = ibdwt_bitlen(N, E, k);
    int64_t v = (int64_t)w[32 + k] + carry;
    carry = v < 0 ? -1 : 0;
    if (v < 0) { v += (int64_t)1 << len; }
    res |= (uint64_t)v << hasBits;
    hasBits += len;
  }
  return res;
}

/* ---- long double mixed-radix FFT (recursive decimation in time, any factorization) */
static cld cmulx(cld a, cld b) { cld r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }

static void fftRec(cld* x, uint32_t n, uint32_t stride, cld* out, int sign, cld* tmp) {
  if (n == 1) { out[0] = x[0]; return; }
  uint32_t p = 2;
  while (n % p) { ++p; }
  uint32_t m = n / p;
  /* sub-transforms of x[r], x[r+p], ... for r < p */
  for (uint32_t r = 0; r < p; ++r) { fftRec(x + r * stride, m, stride * p, out + r * m, sign, tmp); }
  const ld twoPi = 6.283185307179586476925286766559005768L;
  for (uint32_t k = 0; k < m; ++k) {
    for (uint32_t q = 0; q < p; ++q) {
      cld s = {0, 0};
      for (uint32_t r = 0; r < p; ++r) {
        uint64_t e = (uint64_t)r * (k + (uint64_t)q * m) % n;
        ld ang = sign * twoPi * (ld)e / (ld)n;
        cld w = {cosl(ang), sinl(ang)};
        cld t = cmulx(out[r * m + k], w);
        s.re += t.re;
        s.im += t.im;
      }
      tmp[k + q * m] = s;
    }
  }
  memcpy(out, tmp, sizeof(cld) * n);
}

static void fft(cld* x, uint32_t n, int sign) {
  cld* out = (cld*)malloc(sizeof(cld) * n);
  cld* tmp = (cld*)malloc(sizeof(cld) * n);
  fftRec(x, n, 1, out, sign, tmp);
  memcpy(x, out, sizeof(cld) * n);
  free(out);
  free(tmp);
}

/* One squaring of the balanced digits d (length N), in place, fully carried.
 * ll: subtract 2 (as the carry into word 0); mul3: multiply by 3 before carrying.
 * Returns the max round-off |v - rint(v)|. */
double ibdwt_square(int32_t* d, uint32_t N, uint32_t E, int ll, int mul3) {
  cld* x = (cld*)malloc(sizeof(cld) * N);
  for (uint32_t k = 0; k < N; ++k) {
    ld w = exp2l((ld)ibdwt_extra(N, E, k) / N);
    x[k].re = d[k] * w;
    x[k].im = 0;
  }
  fft(x, N, -1);
  for (uint32_t k = 0; k < N; ++k) {
    cld t = {x[k].re * x[k].re - x[k].im * x[k].im, 2 * x[k].re * x[k].im};
    x[k] = t;
  }
  fft(x, N, +1);
  double roe = 0;
  int64_t* u = (int64_t*)malloc(sizeof(int64_t) * N);
  for (uint32_t k = 0; k < N; ++k) {
    ld v = x[k].re / N * exp2l(-(ld)ibdwt_extra(N, E, k) / N);
    ld r = rintl(v);
    double e = (double)fabsl(v - r);
    if (e > roe) { roe = e; }
    u[k] = (int64_t)r * (mul3 ? 3 : 1);
  }
  if (ll) { u[0] -= 2; }
  /* carry propagation around the cyclic word order until every digit is balanced */
  int64_t carry = 0;
  for (int pass = 0; pass < 4; ++pass) {
    for (uint32_t k = 0; k < N; ++k) {
      uint32_t b = ibdwt_bitlen(N, E, k);
      int64_t t = u[k] + carry;
      int64_t lo = (int64_t)((uint64_t)t << (64 - b)) >> (64 - b);
      carry = (t - lo) >> b;
      u[k] = lo;
    }
    if (!carry) { break; }
  }
  for (uint32_t k = 0; k < N; ++k) { d[k] = (int32_t)u[k]; }
  free(u);
  free(x);
  return roe;
}

/* PRP (ll = 0, x0 = 3) or LL (ll = 1, x0 = 4) for iters iterations at FFT length N.
 * res64 after every step-th iteration to out[0..cap). Returns the max round-off. */
double ibdwt_run(uint32_t E, uint32_t N, int ll, uint32_t iters, uint32_t step, uint64_t* out, uint32_t cap) {
  uint32_t nw = (E - 1) / 32 + 1;
  uint32_t* words = (uint32_t*)calloc(nw, 4);
  int32_t* d = (int32_t*)malloc(4 * N);
  words[0] = ll ? 4 : 3;
  ibdwt_expand(words, N, E, d);
  double maxRoe = 0;
  uint32_t n = 0;
  for (uint32_t k = 1; k <= iters; ++k) {
    double r = ibdwt_square(d, N, E, ll, 0);
    if (r > maxRoe) { maxRoe = r; }
    if (step && k % step == 0 && n < cap) {
      int32_t ends[64];
      for (int i = 0; i < 32; ++i) { ends[i] = d[N - 32 + i]; ends[32 + i] = d[i]; }
      out[n++] = ibdwt_residue_from_ends(ends, N, E);
    }
  }
  free(words);
  free(d);
  return maxRoe;
}
