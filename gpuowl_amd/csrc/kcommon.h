// Device helpers shared by the generic (runtime-plan) kernels in kernels.hip and
// the shape-specialized kernels in kfixed.h: global-address-space access,
// complex arithmetic, in-register DFTs, IBDWT word weights, rounding, round-off
// reporting, tagged stairway carries, mid twiddles and the real-signal pairing.
#pragma once
#include <hip/hip_runtime.h>

#include "job.h"

template<int R> struct Roots;
#include "roots_small.h"

typedef double2 T2;

// Pointers read from a JobDesc are generic; these helpers access them through the
// global address space so that the compiler emits global_load/store (not flat_*).
#define GLOBAL __attribute__((address_space(1)))
typedef double v2d __attribute__((ext_vector_type(2)));
typedef int v2i __attribute__((ext_vector_type(2)));
__device__ __forceinline__ T2 gld(const T2* p) {
  v2d v = *(const GLOBAL v2d*)p;
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void gst(T2* p, T2 x) {
  v2d v = {x.x, x.y};
  *(GLOBAL v2d*)p = v;
}
// Streaming accesses for data read or written once per pass (non-temporal: the line
// is not kept in the caches for reuse); NT = false gives the plain access.
template<bool NT> __device__ __forceinline__ T2 ldx(const T2* p) {
  if constexpr (NT) {
    v2d v = __builtin_nontemporal_load((const GLOBAL v2d*)p);
    return make_double2(v.x, v.y);
  } else {
    return gld(p);
  }
}
template<bool NT> __device__ __forceinline__ void stx(T2* p, T2 x) {
  if constexpr (NT) {
    v2d v = {x.x, x.y};
    __builtin_nontemporal_store(v, (GLOBAL v2d*)p);
  } else {
    gst(p, x);
  }
}
__device__ __forceinline__ int2 gld(const int2* p) {
  v2i v = *(const GLOBAL v2i*)p;
  return make_int2(v.x, v.y);
}
__device__ __forceinline__ void gst(int2* p, int2 x) {
  v2i v = {x.x, x.y};
  *(GLOBAL v2i*)p = v;
}
__device__ __forceinline__ int64_t gld(const int64_t* p) { return *(const GLOBAL int64_t*)p; }
__device__ __forceinline__ uint64_t gld(const uint64_t* p) { return *(const GLOBAL uint64_t*)p; }
__device__ __forceinline__ void gst(int64_t* p, int64_t x) { *(GLOBAL int64_t*)p = x; }
__device__ __forceinline__ LineW gld(const LineW* p) {
  const GLOBAL LineW* g = (const GLOBAL LineW*)p;
  LineW r;
  r.w[0] = g->w[0];
  r.w[1] = g->w[1];
  r.iw[0] = g->iw[0];
  r.iw[1] = g->iw[1];
  r.e[0] = g->e[0];
  r.e[1] = g->e[1];
  return r;
}
__device__ __forceinline__ ColW gld(const ColW* p) {
  const GLOBAL ColW* g = (const GLOBAL ColW*)p;
  ColW r;
  r.w = g->w;
  r.iw = g->iw;
  r.e = g->e;
  r.pad = 0;
  return r;
}

// ---------------------------------------------------------------- complex helpers
__device__ __forceinline__ T2 cadd(T2 a, T2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ T2 csub(T2 a, T2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ T2 conj2(T2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
// a * conj(b)
__device__ __forceinline__ T2 cmulc(T2 a, T2 b) {
  return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y));
}
__device__ __forceinline__ T2 csqr(T2 a) { return make_double2(fma(a.x, a.x, -a.y * a.y), 2 * a.x * a.y); }
__device__ __forceinline__ T2 cscale(T2 a, double s) { return make_double2(a.x * s, a.y * s); }
// Multiply by exp(-i pi/2) = -i (forward) or +i (inverse).
template<bool INV> __device__ __forceinline__ T2 mulMinusI(T2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}
// Multiply by exp(-+2 pi i j / R) with compile-time constants (after unrolling).
template<int R, bool INV> __device__ __forceinline__ T2 rootMul(T2 a, int j) {
  j %= R;
  if (j == 0) { return a; }
  if (4 * j == R) { return mulMinusI<INV>(a); }
  if (2 * j == R) { return make_double2(-a.x, -a.y); }
  if (4 * j == 3 * R) { return mulMinusI<!INV>(a); }
  double c = Roots<R>::c[j];
  double s = INV ? Roots<R>::s[j] : -Roots<R>::s[j];
  return make_double2(fma(a.x, c, -a.y * s), fma(a.x, s, a.y * c));
}

// ---------------------------------------------------------------- small DFTs
// Dft<R, INV>::run(a): in-place R-point DFT, natural order in and out,
// kernel exp(-2 pi i/R) forward and exp(+2 pi i/R) inverse (no scaling).
template<int R, bool INV> struct Dft;

template<bool INV> struct Dft<2, INV> {
  __device__ __forceinline__ static void run(T2* a) {
    T2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template<bool INV> struct Dft<4, INV> {
  __device__ __forceinline__ static void run(T2* a) {
    T2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    T2 t2 = cadd(a[1], a[3]), t3 = mulMinusI<INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

// Odd prime P: symmetric sums S_k = a_k + a_{P-k}, differences D_k = a_k - a_{P-k}.
template<int P, bool INV> struct DftOdd {
  __device__ __forceinline__ static void run(T2* a) {
    constexpr int h = (P - 1) / 2;
    T2 S[h], D[h];
    T2 a0 = a[0];
    T2 sum = a0;
#pragma unroll
    for (int k = 1; k <= h; ++k) {
      S[k - 1] = cadd(a[k], a[P - k]);
      D[k - 1] = csub(a[k], a[P - k]);
      sum = cadd(sum, S[k - 1]);
    }
#pragma unroll
    for (int m = 1; m <= h; ++m) {
      T2 A = a0;
      T2 B = make_double2(0, 0);
#pragma unroll
      for (int k = 1; k <= h; ++k) {
        int idx = (k * m) % P;
        double c = Roots<P>::c[idx], s = Roots<P>::s[idx];
        A.x = fma(c, S[k - 1].x, A.x);
        A.y = fma(c, S[k - 1].y, A.y);
        B.x = fma(s, D[k - 1].x, B.x);
        B.y = fma(s, D[k - 1].y, B.y);
      }
      // forward: X_m = A - iB, X_{P-m} = A + iB; inverse swaps the signs.
      T2 mIB = make_double2(B.y, -B.x);  // -i*B
      T2 pIB = make_double2(-B.y, B.x);  // +i*B
      a[m] = INV ? cadd(A, pIB) : cadd(A, mIB);
      a[P - m] = INV ? cadd(A, mIB) : cadd(A, pIB);
    }
    a[0] = sum;
  }
};
template<bool INV> struct Dft<3, INV> : DftOdd<3, INV> {};
template<bool INV> struct Dft<5, INV> : DftOdd<5, INV> {};
template<bool INV> struct Dft<7, INV> : DftOdd<7, INV> {};
template<bool INV> struct Dft<11, INV> : DftOdd<11, INV> {};
template<bool INV> struct Dft<13, INV> : DftOdd<13, INV> {};

// Composite R = P*Q by Cooley-Tukey in registers: input index Q*k1 + k2,
// output index m1 + P*m2, inner twiddle exp(-+2 pi i k2*m1 / R).
template<int P, int Q, bool INV> struct DftCT {
  __device__ __forceinline__ static void run(T2* a) {
    constexpr int R = P * Q;
    T2 b[Q][P];
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) {
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) { b[k2][k1] = a[Q * k1 + k2]; }
      Dft<P, INV>::run(b[k2]);
#pragma unroll
      for (int m1 = 1; m1 < P; ++m1) { b[k2][m1] = rootMul<R, INV>(b[k2][m1], k2 * m1); }
    }
#pragma unroll
    for (int m1 = 0; m1 < P; ++m1) {
      T2 c[Q];
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) { c[k2] = b[k2][m1]; }
      Dft<Q, INV>::run(c);
#pragma unroll
      for (int m2 = 0; m2 < Q; ++m2) { a[m1 + P * m2] = c[m2]; }
    }
  }
};
template<bool INV> struct Dft<6, INV> : DftCT<2, 3, INV> {};
template<bool INV> struct Dft<8, INV> : DftCT<2, 4, INV> {};
template<bool INV> struct Dft<9, INV> : DftCT<3, 3, INV> {};
template<bool INV> struct Dft<10, INV> : DftCT<2, 5, INV> {};
template<bool INV> struct Dft<12, INV> : DftCT<4, 3, INV> {};
template<bool INV> struct Dft<14, INV> : DftCT<2, 7, INV> {};
template<bool INV> struct Dft<15, INV> : DftCT<3, 5, INV> {};
template<bool INV> struct Dft<16, INV> : DftCT<4, 4, INV> {};

// ---------------------------------------------------------------- word helpers
// Rounding constant 3*2^51: for |v| < 2^51, v + RNDVAL holds rint(v) in its low mantissa bits.
#define OWL_RNDVAL 6755399441055744.0
#define OWL_RNDVAL_BITS 0x4338000000000000LL

__device__ __forceinline__ int64_t lowBits64(int64_t u, int b) { return (int64_t)((uint64_t)u << (64 - b)) >> (64 - b); }

// Slot s of a frequency row -> frequency index a (see job.h).
__device__ __forceinline__ int slotFreq(int s, int W) {
  int q = s >> 1;
  return (s & 1) ? (q ? W - q : W / 2) : q;
}

// Weight data for word (line, col, rep).
struct WordW {
  double w, iw;
  int bits;
};
__device__ __forceinline__ WordW wordWeight(const JobDesc& J, const LineW& lw, const ColW& cw, int rep) {
  uint32_t e = lw.e[rep] + cw.e;
  bool wrap = e >= J.N;
  uint32_t ex = wrap ? e - J.N : e;
  WordW r;
  r.bits = J.bitsLo + (ex < J.rE);
  r.w = lw.w[rep] * cw.w;
  r.iw = lw.iw[rep] * cw.iw;
  if (wrap) {
    r.w *= 0.5;
    r.iw *= 2.0;
  }
  return r;
}

// Round v to the nearest integer (|v| < 2^51); track the round-off.
__device__ __forceinline__ int64_t roundWord(double v, float& roe) {
  double t = v + OWL_RNDVAL;
  double r = t - OWL_RNDVAL;
  roe = fmaxf(roe, (float)fabs(v - r));
  return __double_as_longlong(t) - OWL_RNDVAL_BITS;
}

// A launch measures round-off only when its slot is < OWL_ROE_SIZE (the host measures a
// window of iterations, as the reference does, src/Gpu.cpp:1459-1464); the slot is
// uniform, so unmeasured launches skip the reduction and the atomic.
__device__ __forceinline__ void roeReport(const JobDesc& J, uint32_t slot, float roe) {
  if (slot >= OWL_ROE_SIZE) { return; }
  // wave-level max, then one atomic per wave (float bits of non-negative values order as uints)
  for (int off = 32; off > 0; off >>= 1) { roe = fmaxf(roe, __shfl_xor(roe, off, 64)); }
  if ((threadIdx.x & 63) == 0) { atomicMax(J.roe + slot, __float_as_uint(roe)); }
}

#define OWL_SPIN_LIMIT (1u << 24)

// Stairway carries are tagged: bits 48..63 hold the launch tag, bits 0..47 the
// signed carry. A consumer polls its column's carry until the tag is current, so
// no flags, fences or barriers are needed between the groups.
__device__ __forceinline__ uint64_t packCarry(int64_t c, uint32_t tag) {
  return ((uint64_t)c & 0x0000FFFFFFFFFFFFull) | ((uint64_t)tag << 48);
}
__device__ __forceinline__ int64_t unpackCarry(uint64_t v) { return (int64_t)(v << 16) >> 16; }

__device__ __forceinline__ uint64_t pollCarry(const JobDesc& J, uint64_t* p, uint32_t tag) {
  uint64_t v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t n = 0;
  while ((uint32_t)(v >> 48) != tag) {
    __builtin_amdgcn_s_sleep(1);
    if (++n >= OWL_SPIN_LIMIT) {
      atomicOr(J.flags + 2, 1u);  // hand-off timed out: flagged to the host, kernel still exits
      break;
    }
    v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return v;
}

// Ticket order of the fused row pass (the runtime-plan k_row, and k_rowx below 8 blocks; k_rowx
// otherwise uses the chunked per-XCD order of fusedItem): groups are ordered by start, not by
// blockIdx: each group takes the next ticket of the job's counter, and the ticket is its block. A group then only waits
// for groups that have already started, so the carry hand-off progresses even when other
// kernels (other queues) hold the rest of the device. Every fused launch takes exactly
// nb + 1 tickets; the group that takes the last one (no other group of the launch touches
// the counter after it) sets the counter back to 0 for the next launch.
__device__ __forceinline__ int fusedBlock(const JobDesc& J, int* slot) {
  if (threadIdx.x == 0) {
    const int t = int(atomicAdd((unsigned long long*)J.ticket, 1ull) % (unsigned long long)(J.nb + 1));
    if (t == (int)J.nb) { __hip_atomic_store(J.ticket, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    *slot = t;
  }
  __syncthreads();
  return *slot;
}

// N polls issued together: every carry's first load is in flight at once, then only the
// ones whose tag is not current yet are loaded again.
template<int N>
__device__ __forceinline__ void pollStart(uint64_t* const* p, uint64_t* v) {
#pragma unroll
  for (int k = 0; k < N; ++k) { v[k] = __hip_atomic_load(p[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
}
// Continue polling from values v already loaded by pollStart (independent work can run between).
template<int N>
__device__ __forceinline__ void pollFinish(const JobDesc& J, uint64_t* const* p, uint32_t tag, uint64_t* v) {
  uint32_t n = 0;
  while (true) {
    bool ready = true;
#pragma unroll
    for (int k = 0; k < N; ++k) { ready = ready && (uint32_t)(v[k] >> 48) == tag; }
    if (ready) { break; }
    __builtin_amdgcn_s_sleep(1);
    if (++n >= OWL_SPIN_LIMIT) {
      atomicOr(J.flags + 2, 1u);  // hand-off timed out: flagged to the host, kernel still exits
      break;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if ((uint32_t)(v[k] >> 48) != tag) { v[k] = __hip_atomic_load(p[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    }
  }
}
template<int N>
__device__ __forceinline__ void pollCarries(const JobDesc& J, uint64_t* const* p, uint32_t tag, uint64_t* v) {
  pollStart<N>(p, v);
  pollFinish<N>(J, p, tag, v);
}

// Round-off of a workgroup: each wave's max goes to part[wave]; after a barrier, thread 0
// folds them and makes the group's single atomic (one per group instead of one per wave).
__device__ __forceinline__ void roeWave(uint32_t slot, float roe, float* part) {
  if (slot >= OWL_ROE_SIZE) { return; }
  for (int off = 32; off > 0; off >>= 1) { roe = fmaxf(roe, __shfl_xor(roe, off, 64)); }
  if ((threadIdx.x & 63) == 0) { part[threadIdx.x >> 6] = roe; }
}
__device__ __forceinline__ void roeGroup(const JobDesc& J, uint32_t slot, const float* part, int nWaves) {
  if (threadIdx.x == 0 && slot < OWL_ROE_SIZE) {
    float m = part[0];
    for (int w = 1; w < nWaves; ++w) { m = fmaxf(m, part[w]); }
    atomicMax(J.roe + slot, __float_as_uint(m));
  }
}

// ---------------------------------------------------------------- column pass
__device__ __forceinline__ T2 midTw(const JobDesc& J, int a, int l) {
  return cmul(gld(J.twMidA + a * 64 + (l & 63)), gld(J.twMidB + a * J.nHi + (l >> 6)));
}

// Real-signal pairing for the squaring (see the derivation in the host docs):
// with D = X - conj(Xp), t = (1 + w) D^2 / 4:  X <- X^2 - t,  Xp <- Xp^2 - conj(t),
// where w = exp(-2 pi i K / ND) for X = Z_K and Xp = Z_{ND-K}.
__device__ __forceinline__ void pairSq(T2& X, T2& Xp, T2 w) {
  T2 D = csub(X, conj2(Xp));
  T2 t = cscale(cmul(make_double2(1.0 + w.x, w.y), csqr(D)), 0.25);
  T2 nX = csub(csqr(X), t);
  Xp = csub(csqr(Xp), conj2(t));
  X = nX;
}
// Same for a product with a second transformed operand (Y, Yp).
__device__ __forceinline__ void pairMul(T2& X, T2& Xp, T2 Y, T2 Yp, T2 w) {
  T2 D = csub(X, conj2(Xp));
  T2 E = csub(Y, conj2(Yp));
  T2 t = cscale(cmul(make_double2(1.0 + w.x, w.y), cmul(D, E)), 0.25);
  T2 nX = csub(cmul(X, Y), t);
  Xp = csub(cmul(Xp, Yp), conj2(t));
  X = nX;
}

