// Shape-specialized row and column passes for gfx950.
//
// Same algorithm, layout and modes as k_row / k_col in kernels.hip, with the
// FFT plan fixed at compile time (no integer division, no radix dispatch, every
// loop unrolled) and fewer LDS round trips:
//   k_colx: the first forward stage is fused with the global load and the mid
//           twiddle, the last inverse stage with the conjugate twiddle and the
//           store (forward plan [C0,C1,C2], inverse plan [C2,C1,C0]).
//   k_rowx: the last inverse stage leaves each thread holding whole columns
//           (T = W/F0 threads, columns tid + k*T for all R lines), so rounding,
//           the carry chain, the stairway hand-off and the weighting all run in
//           registers and feed the first forward stage directly
//           (forward plan [F0,F1,F2], inverse plan [F2,F1,F0]).
// LDS is padded by one element per 16 (index i -> i + i/16) so that the
// stride-R writes of the first stages do not hit one bank, and consecutive
// line (row pass) or column (column pass) regions start 4 or 8 elements apart
// in the 64-bank cycle, so that the slot-order load and store phases, whose 16
// lanes of a b128 access span several regions, spread over the banks.
#include "kcommon.h"

__device__ __forceinline__ int padi(int i) { return i + (i >> 4); }
// region strides in 16-byte elements (the host sizes LDS with the same formulas)
template<int W> constexpr int rowRegion() { return W + W / 16 + 4; }
template<int H> constexpr int colRegion() { return H + H / 16 + 8; }

// One LDS -> LDS radix-R autosort stage over NFFT FFTs of length L stored in
// padded regions; Ns is the product of the earlier radices. tw[m] = exp(-2 pi i m/L).
// PRIV: the caller allows wave-private stages. A stage with 64 butterflies per FFT
// (or a one-wave group) reads and writes only its own wave's regions; LDS operations of
// one wave complete in order, so such a stage needs no workgroup barrier, only a compiler
// fence. The caller puts a __syncthreads() after the last stage when the next phase
// reads across waves.
template<int L, int REG, int NFFT, int T, bool INV, int R, int Ns, bool PRIV = false>
__device__ __forceinline__ void xstage(T2* lds, const T2* __restrict__ tw, int tid) {
  constexpr bool priv = PRIV && (T == 64 || L / R == 64);
  constexpr int NB = L / R;
  constexpr int TOT = NB * NFFT;
  constexpr int MB = (TOT + T - 1) / T;
  constexpr int STRIDE = L / (Ns * R);
  T2 v[MB][R];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int g = tid + m * T;
    if (TOT % T == 0 || g < TOT) {
      const int f = g / NB, j = g % NB;
      const T2* x = lds + f * REG;
#pragma unroll
      for (int k = 0; k < R; ++k) { v[m][k] = x[padi(j + k * NB)]; }
      if (Ns > 1) {
        const int p = j % Ns;
#pragma unroll
        for (int k = 1; k < R; ++k) {
          T2 w = gld(tw + p * STRIDE * k);
          v[m][k] = INV ? cmulc(v[m][k], w) : cmul(v[m][k], w);
        }
      }
      Dft<R, INV>::run(v[m]);
    }
  }
  if (priv) {
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int g = tid + m * T;
    if (TOT % T == 0 || g < TOT) {
      const int f = g / NB, j = g % NB;
      T2* x = lds + f * REG;
      const int p = j % Ns;
      const int base = (j - p) * R + p;
#pragma unroll
      for (int k = 0; k < R; ++k) { x[padi(base + k * Ns)] = v[m][k]; }
    }
  }
  if (priv) {
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

template<int L, int REG, int NFFT, int T, bool INV, bool PRIV, int Ns, int... Rs> struct XStages;
template<int L, int REG, int NFFT, int T, bool INV, bool PRIV, int Ns> struct XStages<L, REG, NFFT, T, INV, PRIV, Ns> {
  __device__ __forceinline__ static void run(T2*, const T2*, int) {}
};
template<int L, int REG, int NFFT, int T, bool INV, bool PRIV, int Ns, int R, int... Rs>
struct XStages<L, REG, NFFT, T, INV, PRIV, Ns, R, Rs...> {
  __device__ __forceinline__ static void run(T2* lds, const T2* tw, int tid) {
    xstage<L, REG, NFFT, T, INV, R, Ns, PRIV>(lds, tw, tid);
    XStages<L, REG, NFFT, T, INV, PRIV, Ns * R, Rs...>::run(lds, tw, tid);
  }
};

// ---------------------------------------------------------------- column pass
// Grid (W/2, batch), T threads, LDS 2*colRegion<H>()*16 bytes. Modes as k_col.
// J.twMidS[a*C0 + k] = exp(-2 pi i a*k*(H/C0) / ND).
template<int H, int T, int C0, int C1, int C2>
__global__ __launch_bounds__(T) void k_colx(const JobDesc* __restrict__ jobs, int mode, int fi, int oi) {
  static_assert(C0 * C1 * C2 == H, "plan");
  extern __shared__ T2 lds[];
  constexpr int REG = colRegion<H>();
  constexpr int NB0 = H / C0;
  constexpr int NBT = 2 * NB0;
  constexpr int MB0 = (NBT + T - 1) / T;
  const JobDesc& J = jobs[blockIdx.y];
  const int W = J.W, R = J.R, lgR = J.lgR;
  const int q = blockIdx.x, tid = threadIdx.x;
  const int a0 = q, a1 = q ? W - q : W / 2;
  T2* io = J.freq[fi];
  const size_t colBase = (size_t)2 * q * R;
  const size_t blockStride = (size_t)W * R;

  // forward stage 0 (radix C0, Ns = 1), fused with the load and the mid twiddle.
  // Lane order within 2R lanes: (column h, line-in-block r), so 2R lanes read 2R*16
  // contiguous bytes of one line block.
#pragma unroll
  for (int m = 0; m < MB0; ++m) {
    const int g = tid + m * T;
    if (NBT % T == 0 || g < NBT) {
      const int h = (g >> lgR) & 1;
      const int j = ((g >> (lgR + 1)) << lgR) | (g & (R - 1));
      const int a = h ? a1 : a0;
      const T2 base = midTw(J, a, j);
      const T2* S = J.twMidS + a * C0;
      T2 v[C0];
#pragma unroll
      for (int k = 0; k < C0; ++k) {
        const int l = j + k * NB0;
        T2 x = gld(io + (size_t)(l >> lgR) * blockStride + colBase + (h << lgR) + (l & (R - 1)));
        v[k] = cmul(x, k ? cmul(base, gld(S + k)) : base);
      }
      Dft<C0, false>::run(v);
#pragma unroll
      for (int k = 0; k < C0; ++k) { lds[h * REG + padi(j * C0 + k)] = v[k]; }
    }
  }
  __syncthreads();
  XStages<H, REG, 2, T, false, false, C0, C1, C2>::run(lds, J.twCol, tid);

  const int per = 2 * R;
  if (mode == COL_FWD) {
    for (int e = tid; e < 2 * H; e += T) {
      int lb = e / per;
      int wi = e - lb * per;
      int h = wi >> lgR;
      int b = lb * R + (wi & (R - 1));
      gst(io + (size_t)lb * blockStride + colBase + wi, lds[h * REG + padi(b)]);
    }
    return;
  }

  const T2* other = (mode == COL_MUL) ? J.freq[oi] : nullptr;
  auto spos = [&](int h, int b) -> size_t { return (size_t)(b >> lgR) * blockStride + colBase + (h << lgR) + (b & (R - 1)); };
  const T2* twCol = J.twCol;
  if (q) {
    const T2 wa = gld(J.twPair + a0);
    for (int b = tid; b < H; b += T) {
      int bp = H - 1 - b;
      T2 X = lds[padi(b)], Xp = lds[REG + padi(bp)];
      T2 w = cmul(wa, gld(twCol + b));
      if (other) {
        pairMul(X, Xp, gld(other + spos(0, b)), gld(other + spos(1, bp)), w);
      } else {
        pairSq(X, Xp, w);
      }
      lds[padi(b)] = X;
      lds[REG + padi(bp)] = Xp;
    }
  } else {
    const T2 wa = gld(J.twPair + W / 2);
    for (int i = tid; i <= H; i += T) {
      if (i <= H / 2) {
        int b = i, bp = (H - i) % H;
        T2 X = lds[padi(b)], Xp = lds[padi(bp)];
        T2 w = gld(twCol + b);
        if (other) {
          pairMul(X, Xp, gld(other + spos(0, b)), gld(other + spos(0, bp)), w);
        } else {
          pairSq(X, Xp, w);
        }
        lds[padi(b)] = X;
        if (bp != b) { lds[padi(bp)] = Xp; }
      } else {
        int b = i - H / 2 - 1, bp = H - 1 - b;
        T2 X = lds[REG + padi(b)], Xp = lds[REG + padi(bp)];
        T2 w = cmul(wa, gld(twCol + b));
        if (other) {
          pairMul(X, Xp, gld(other + spos(1, b)), gld(other + spos(1, bp)), w);
        } else {
          pairSq(X, Xp, w);
        }
        lds[REG + padi(b)] = X;
        lds[REG + padi(bp)] = Xp;
      }
    }
  }
  __syncthreads();
  XStages<H, REG, 2, T, true, false, 1, C2, C1>::run(lds, twCol, tid);

  // last inverse stage (radix C0, Ns = NB0), fused with the conjugate mid twiddle and the store
#pragma unroll
  for (int m = 0; m < MB0; ++m) {
    const int g = tid + m * T;
    if (NBT % T == 0 || g < NBT) {
      const int h = (g >> lgR) & 1;
      const int j = ((g >> (lgR + 1)) << lgR) | (g & (R - 1));
      const int a = h ? a1 : a0;
      T2 v[C0];
#pragma unroll
      for (int k = 0; k < C0; ++k) { v[k] = lds[h * REG + padi(j + k * NB0)]; }
#pragma unroll
      for (int k = 1; k < C0; ++k) { v[k] = cmulc(v[k], gld(twCol + j * k)); }
      Dft<C0, true>::run(v);
      const T2 base = midTw(J, a, j);
      const T2* S = J.twMidS + a * C0;
#pragma unroll
      for (int k = 0; k < C0; ++k) {
        const int l = j + k * NB0;
        gst(io + (size_t)(l >> lgR) * blockStride + colBase + (h << lgR) + (l & (R - 1)),
            cmulc(v[k], k ? cmul(base, gld(S + k)) : base));
      }
    }
  }
}

// ---------------------------------------------------------------- column pass, one region
// The two columns of the pair go through one LDS region in turn (half the LDS of k_colx,
// so two groups share a CU), and the real-signal pairing runs in registers: the last
// forward stage (radix C2, T = H/C2 butterflies) leaves thread t holding X0[t + T*k];
// column 1's last stage is computed for butterfly T-1-t, so the partner of X0[b],
// X1[H-1-b], is X1[(T-1-t) + T*(C2-1-k)] in the same thread. The inverse's first stage
// (radix C2, Ns = 1) reads exactly those registers. The pair q = 0 (columns 0 and W/2,
// each paired with itself) goes through LDS. Plan: forward [16, C1, C2], inverse
// [C2, C1, 16]; J.twMidS[a*16 + k] = exp(-2 pi i a*k*(H/16) / ND).
template<int H, int C1, int C2>
__global__ __launch_bounds__(H / C2, 2) void k_coly(const JobDesc* __restrict__ jobs, int mode, int fi, int oi) {
  constexpr int C0 = 16;
  constexpr int T = H / C2;
  constexpr int NB0 = H / C0;
  static_assert(C0 * C1 * C2 == H && T % 64 == 0 && NB0 <= T, "plan");
  constexpr int REG = colRegion<H>();
  extern __shared__ T2 lds[];
  const JobDesc& J = jobs[blockIdx.y];
  const int W = J.W, R = J.R, lgR = J.lgR;
  const int q = blockIdx.x, tid = threadIdx.x;
  const int a0 = q, a1 = q ? W - q : W / 2;
  T2* io = J.freq[fi];
  const size_t colBase = (size_t)2 * q * R;
  const size_t blockStride = (size_t)W * R;
  const T2* twCol = J.twCol;
  auto gpos = [&](int h, int l) -> size_t {
    return (size_t)(l >> lgR) * blockStride + colBase + (h << lgR) + (l & (R - 1));
  };
  const int jl1 = q ? T - 1 - tid : tid;  // column 1's last-stage butterfly

  T2 x0[C2], x1[C2];

  // forward FFT of column h (mid twiddle for column index a), last stage for butterfly jl
  auto forward = [&](int h, int a, int jl, T2* out) {
    if (tid < NB0) {
      const int j = tid;
      const T2 base = midTw(J, a, j);
      const T2* S = J.twMidS + a * C0;
      T2 v[C0];
#pragma unroll
      for (int k = 0; k < C0; ++k) {
        T2 x = gld(io + gpos(h, j + k * NB0));
        v[k] = cmul(x, k ? cmul(base, gld(S + k)) : base);
      }
      Dft<C0, false>::run(v);
#pragma unroll
      for (int k = 0; k < C0; ++k) { lds[padi(j * C0 + k)] = v[k]; }
    }
    __syncthreads();
    XStages<H, REG, 1, T, false, false, C0, C1>::run(lds, twCol, tid);
#pragma unroll
    for (int k = 0; k < C2; ++k) { out[k] = lds[padi(jl + k * T)]; }
#pragma unroll
    for (int k = 1; k < C2; ++k) { out[k] = cmul(out[k], gld(twCol + jl * k)); }
    Dft<C2, false>::run(out);
    __syncthreads();
  };

  // inverse FFT of column h from registers (butterfly jl of the first stage), then store
  auto inverse = [&](int h, int a, int jl, T2* in) {
    Dft<C2, true>::run(in);
#pragma unroll
    for (int k = 0; k < C2; ++k) { lds[padi(jl * C2 + k)] = in[k]; }
    __syncthreads();
    XStages<H, REG, 1, T, true, false, C2, C1>::run(lds, twCol, tid);
    if (tid < NB0) {
      const int j = tid;
      T2 v[C0];
#pragma unroll
      for (int k = 0; k < C0; ++k) { v[k] = lds[padi(j + k * NB0)]; }
#pragma unroll
      for (int k = 1; k < C0; ++k) { v[k] = cmulc(v[k], gld(twCol + j * k)); }
      Dft<C0, true>::run(v);
      const T2 base = midTw(J, a, j);
      const T2* S = J.twMidS + a * C0;
#pragma unroll
      for (int k = 0; k < C0; ++k) {
        gst(io + gpos(h, j + k * NB0), cmulc(v[k], k ? cmul(base, gld(S + k)) : base));
      }
    }
    __syncthreads();
  };

  forward(0, a0, tid, x0);
  forward(1, a1, jl1, x1);

  if (mode == COL_FWD) {
#pragma unroll
    for (int k = 0; k < C2; ++k) {
      gst(io + gpos(0, tid + k * T), x0[k]);
      gst(io + gpos(1, jl1 + k * T), x1[k]);
    }
    return;
  }

  const T2* other = (mode == COL_MUL) ? J.freq[oi] : nullptr;
  if (q) {
    const T2 wa = gld(J.twPair + a0);
#pragma unroll
    for (int k = 0; k < C2; ++k) {
      const int b = tid + k * T, bp = H - 1 - b;
      const T2 w = cmul(wa, gld(twCol + b));
      if (other) {
        pairMul(x0[k], x1[C2 - 1 - k], gld(other + gpos(0, b)), gld(other + gpos(1, bp)), w);
      } else {
        pairSq(x0[k], x1[C2 - 1 - k], w);
      }
    }
  } else {
    // column 0 pairs b with (H - b) % H, column W/2 pairs b with H - 1 - b
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      T2* x = h ? x1 : x0;
#pragma unroll
      for (int k = 0; k < C2; ++k) { lds[padi(tid + k * T)] = x[k]; }
      __syncthreads();
      const T2 wa = gld(J.twPair + (h ? W / 2 : 0));
      const int last = h ? H / 2 - 1 : H / 2;
      for (int b = tid; b <= last; b += T) {
        const int bp = h ? H - 1 - b : (H - b) % H;
        T2 X = lds[padi(b)], Xp = lds[padi(bp)];
        const T2 w = h ? cmul(wa, gld(twCol + b)) : gld(twCol + b);
        if (other) {
          pairMul(X, Xp, gld(other + gpos(h, b)), gld(other + gpos(h, bp)), w);
        } else {
          pairSq(X, Xp, w);
        }
        lds[padi(b)] = X;
        if (bp != b) { lds[padi(bp)] = Xp; }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < C2; ++k) { x[k] = lds[padi(tid + k * T)]; }
      __syncthreads();
    }
  }

  inverse(0, a0, tid, x0);
  inverse(1, a1, jl1, x1);
}

// ---------------------------------------------------------------- row pass
// Grid (nb + 1 for ROW_FUSED else nb, batch), T = W/F0 threads, LDS R*rowRegion<W>()*16 bytes.
// Modes, options and the tagged stairway as k_row.
template<int W, int R, int F0, int F1, int F2>
__global__ __launch_bounds__(W / F0) void k_rowx(const JobDesc* __restrict__ jobs, int mode, int fi, int wi, int opts,
                                                 uint32_t roeSlot, uint32_t tag) {
  static_assert(F0 * F1 * F2 == W, "plan");
  constexpr int T = W / F0;
  constexpr int LW = R * W;
  constexpr int REG = rowRegion<W>();
  extern __shared__ T2 lds[];
  const JobDesc& J = jobs[blockIdx.y];
  const int nb = J.nb;
  const int g = blockIdx.x;
  const int lb = (g == nb) ? 0 : g;
  const int tid = threadIdx.x;
  const bool ll = opts & 1, mul3 = opts & 2;
  T2* blk = J.freq[fi] + (size_t)lb * LW;

  // Thread tid owns columns c = tid + k*T (k < F0) of the R lines of the block.
  T2 v[R][F0];
  int2 d[R][F0];
  ColW cw[F0];
#pragma unroll
  for (int k = 0; k < F0; ++k) { cw[k] = gld(J.colW + tid + k * T); }

  if (mode == ROW_LEADIN) {
    const int2* src = J.words[wi] + (size_t)lb * LW;
#pragma unroll
    for (int k = 0; k < F0; ++k) {
#pragma unroll
      for (int r = 0; r < R; ++r) { d[r][k] = gld(src + (tid + k * T) * R + r); }
    }
  } else {
#pragma unroll
    for (int i = 0; i < LW / T; ++i) {
      const int e = tid + i * T;
      const int s = e / R, r = e % R;
      lds[r * REG + padi(slotFreq(s, W))] = gld(blk + e);
    }
    __syncthreads();
    XStages<W, REG, R, T, true, true, 1, F2, F1>::run(lds, J.twRow, tid);
    __syncthreads();
    // last inverse stage (radix F0, Ns = T): butterfly tid of each line -> columns tid + k*T
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int k = 0; k < F0; ++k) { v[r][k] = lds[r * REG + padi(tid + k * T)]; }
#pragma unroll
      for (int k = 1; k < F0; ++k) { v[r][k] = cmulc(v[r][k], gld(J.twRow + tid * k)); }
      Dft<F0, true>::run(v[r]);
    }

    // unweight, round, carry along the block's 2R words per column
    float roe = 0;
    int64_t s[F0];
#pragma unroll
    for (int k = 0; k < F0; ++k) { s[k] = 0; }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int l = lb * R + r;
      const LineW lw = gld(J.lineW + l);
#pragma unroll
      for (int k = 0; k < F0; ++k) {
        const int c = tid + k * T;
        int32_t dd[2];
#pragma unroll
        for (int rep = 0; rep < 2; ++rep) {
          WordW ww = wordWeight(J, lw, cw[k], rep);
          int64_t u = roundWord((rep ? v[r][k].y : v[r][k].x) * ww.iw, roe);
          if (mul3) { u *= 3; }
          if (ll && l == 0 && c == 0 && rep == 0) { u -= 2; }
          u += s[k];
          int64_t t = lowBits64(u, ww.bits);
          s[k] = (u - t) >> ww.bits;
          dd[rep] = (int32_t)t;
        }
        d[r][k] = make_int2(dd[0], dd[1]);
      }
    }
    roeReport(J, roeSlot, roe);

    if (mode == ROW_CARRYOUT) {
      int2* dst = J.words[wi] + (size_t)lb * LW;
#pragma unroll
      for (int k = 0; k < F0; ++k) {
        const int c = tid + k * T;
        gst(J.carry[0] + (size_t)lb * W + c, s[k]);
#pragma unroll
        for (int r = 0; r < R; ++r) { gst(dst + c * R + r, d[r][k]); }
      }
      return;
    }

    if (g < nb) {
#pragma unroll
      for (int k = 0; k < F0; ++k) {
        __hip_atomic_store(J.stair + (size_t)g * W + tid + k * T, packCarry(s[k], tag), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (g == 0) { return; }
    const int pred = g - 1;
#pragma unroll
    for (int k = 0; k < F0; ++k) {
      const int c = tid + k * T;
      const int srcCol = (g == nb) ? (c == 0 ? W - 1 : c - 1) : c;
      int64_t carry = unpackCarry(pollCarry(J, J.stair + (size_t)pred * W + srcCol, tag));
      if (g == nb) { pollCarry(J, J.stair + c, tag); }  // group 0 has read block 0
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (carry) {
          const LineW lw = gld(J.lineW + lb * R + r);
          WordW w0 = wordWeight(J, lw, cw[k], 0);
          int64_t t = (int64_t)d[r][k].x + carry;
          int64_t nd = lowBits64(t, w0.bits);
          carry = (t - nd) >> w0.bits;
          d[r][k].x = (int32_t)nd;
          if (r == R - 1) {
            d[r][k].y += (int32_t)carry;
            carry = 0;
          } else if (carry) {
            WordW w1 = wordWeight(J, lw, cw[k], 1);
            t = (int64_t)d[r][k].y + carry;
            nd = lowBits64(t, w1.bits);
            carry = (t - nd) >> w1.bits;
            d[r][k].y = (int32_t)nd;
          }
        }
      }
    }
  }

  // weight, then forward stage 0 (radix F0, Ns = 1) straight from registers
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const LineW lw = gld(J.lineW + lb * R + r);
#pragma unroll
    for (int k = 0; k < F0; ++k) {
      WordW w0 = wordWeight(J, lw, cw[k], 0);
      WordW w1 = wordWeight(J, lw, cw[k], 1);
      v[r][k] = make_double2(d[r][k].x * w0.w, d[r][k].y * w1.w);
    }
    Dft<F0, false>::run(v[r]);
  }
  __syncthreads();  // every thread has finished reading the inverse's LDS
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int k = 0; k < F0; ++k) { lds[r * REG + padi(tid * F0 + k)] = v[r][k]; }
  }
  __syncthreads();
  XStages<W, REG, R, T, false, true, F0, F1, F2>::run(lds, J.twRow, tid);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < LW / T; ++i) {
    const int e = tid + i * T;
    const int s = e / R, r = e % R;
    gst(blk + e, lds[r * REG + padi(slotFreq(s, W))]);
  }
}

// ---------------------------------------------------------------- table
struct OwlKernelEntry {
  const char* name;
  const void* fn;
  unsigned maxThreads;
};

// Names: colx<H> and rowx<W>r<R>; the host uses them when the shape matches
// (column plan [C0,C1,C2] with C0 = 16 for twMidS, row plan [F0,F1,F2]).
extern "C" const OwlKernelEntry owl_kernel_table_fixed[] = {
    {"colx3840", (const void*)&k_colx<3840, 512, 16, 16, 15>, 512},
    {"colx768", (const void*)&k_colx<768, 128, 16, 16, 3>, 128},
    {"colx512", (const void*)&k_colx<512, 64, 16, 16, 2>, 64},
    {"coly3840", (const void*)&k_coly<3840, 16, 15>, 256},
    {"coly768", (const void*)&k_coly<768, 4, 12>, 64},
    {"coly512", (const void*)&k_coly<512, 4, 8>, 64},
    {"rowx1024r4", (const void*)&k_rowx<1024, 4, 4, 16, 16>, 256},
    {"rowx1024r2", (const void*)&k_rowx<1024, 2, 4, 16, 16>, 256},
    {"rowx256r2", (const void*)&k_rowx<256, 2, 4, 4, 16>, 64},
    {nullptr, nullptr, 0},
};
