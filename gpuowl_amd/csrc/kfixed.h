// Shape-specialized row and column passes for gfx950 (templates; the kfix_*.hip files
// instantiate them for the compiled shapes and export their name tables).
//
// Same algorithm, layout and modes as k_row / k_col in kernels.hip, with the
// FFT plan fixed at compile time (no integer division, no radix dispatch, every
// loop unrolled) and fewer LDS round trips:
//   k_colx: the first forward stage is fused with the global load and the mid
//           twiddle, the last inverse stage with the conjugate twiddle and the
//           store (forward plan [C0,C1,C2], inverse plan [C1,C2,C0]).
//   k_rowx: the last inverse stage leaves each thread holding whole columns
//           (T = W/F0 threads, columns tid + k*T for all R lines), so rounding,
//           the carry chain, the stairway hand-off and the weighting all run in
//           registers and feed the first forward stage directly
//           (forward plan [F0,F1,F2], inverse plan [F2,F1,F0]).
// LDS bank model (gfx950, 64 banks of 4 bytes, 16-byte accesses, fitted to rocprofv3's
// SQ_LDS_IDX_ACTIVE over probe patterns): a read serves 32 lanes per two cycles, taking as many
// cycles as the most distinct 16-byte slots among them that share a slot index mod 16; a write
// serves 8 lanes per cycle, as many cycles as the most distinct slots sharing an index mod 8.
// LDS is padded by one element per 16 (index i -> i + i/16) so that the strided accesses of the
// FFT stages spread over the banks, and consecutive line (row pass) or column (column pass)
// regions start 6 (row pass, 4 or 8 lines) or 4 elements apart mod 16, which keeps the writes
// that span several regions (the row pass's slot-order scatter, the column pass's first stage)
// on distinct slots mod 8.
#pragma once

#include <type_traits>

#include "kcommon.h"

__device__ __forceinline__ int padi(int i) { return i + (i >> 4); }
// Column-pass region layout: element i at lx(i), the low three bits of i XORed with bits 4..6
// (each 16-element run keeps its place, permuted), no padding. Consecutive elements read by a
// half-wave stay in distinct bank quads within each read sub-group, and the R consecutive outputs
// of a radix-R butterfly for 8 consecutive lanes (stride R across the lanes) land on 8 distinct
// slots mod 8. (In the row pass, whose LDS phases overlap the other group's work, the padded layout
// measured as fast: there the fewer address operations weigh as much as the fewer bank conflicts.)
__device__ __forceinline__ int lx(int i) { return i ^ ((i >> 4) & 7); }
template<bool XL> __device__ __forceinline__ int lpos(int i) {
  if constexpr (XL) {
    return lx(i);
  } else {
    return padi(i);
  }
}
// region strides in 16-byte elements (the host sizes LDS with the same formulas)
template<int W, int R> constexpr int rowRegion() { return W + W / 16 + ((R == 4 || R == 8) ? 6 : 4); }
template<int H> constexpr int colRegion() { return H + H / 16 + 4; }  // k_coly (padded layout)
template<int H> constexpr int colRegionX() { return H + 4; }          // k_colx (lx layout)

// One LDS -> LDS radix-R autosort stage over NFFT FFTs of length L stored in
// padded regions; Ns is the product of the earlier radices. The twiddle of butterfly input k
// at p = j mod Ns is exp(-2 pi i p*k/(Ns*R)) = tw(p*STRIDE*k) with tw(m) = exp(-2 pi i m/L); the
// provider's st<Ns*R>(p, k, STRIDE) returns it (from a table of the (Ns*R)-th roots when it has one).
// PRIV: the caller allows wave-private stages. A stage with 64 butterflies per FFT
// (or a one-wave group) reads and writes only its own wave's regions; LDS operations of
// one wave complete in order, so such a stage needs no workgroup barrier, only a compiler
// fence. The caller puts a __syncthreads() after the last stage when the next phase
// reads across waves.
template<int L, int T, int R, bool PRIV> constexpr bool stagePrivate() { return PRIV && (T == 64 || L / R == 64); }

template<int L, int REG, int NFFT, int T, bool INV, int R, int Ns, bool PRIV = false, bool XL = false, class TW>
__device__ __forceinline__ void xstage(T2* lds, const TW& tw, int tid) {
  constexpr bool priv = stagePrivate<L, T, R, PRIV>();
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
      if constexpr (XL && NB % 128 == 0) {  // lx(j + k*NB) = lx(j) + k*NB
        const T2* xj = x + lx(j);
#pragma unroll
        for (int k = 0; k < R; ++k) { v[m][k] = xj[k * NB]; }
      } else if constexpr (XL && NB % 64 == 0) {  // lx(j + k*NB) = (lx(j) ^ 4*(k & 1)) + k*NB
        const T2* x0 = x + lx(j);
        const T2* x1 = x + (lx(j) ^ 4);
#pragma unroll
        for (int k = 0; k < R; ++k) { v[m][k] = (k & 1 ? x1 : x0)[k * NB]; }
      } else if constexpr (XL) {
#pragma unroll
        for (int k = 0; k < R; ++k) { v[m][k] = x[lx(j + k * NB)]; }
      } else if constexpr (NB % 16 == 0) {  // padi(j + k*NB) = padi(j) + k*(NB + NB/16)
        const T2* xj = x + padi(j);
#pragma unroll
        for (int k = 0; k < R; ++k) { v[m][k] = xj[k * (NB + NB / 16)]; }
      } else {
#pragma unroll
        for (int k = 0; k < R; ++k) { v[m][k] = x[padi(j + k * NB)]; }
      }
      if constexpr (Ns > 1) {
        const int p = j % Ns;
#pragma unroll
        for (int k = 1; k < R; ++k) {
          T2 w = tw.template st<Ns * R, R>(p, k, STRIDE);
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
      if constexpr (XL && Ns % 128 == 0) {  // lx(base + k*Ns) = lx(base) + k*Ns
        T2* xb = x + lx(base);
#pragma unroll
        for (int k = 0; k < R; ++k) { xb[k * Ns] = v[m][k]; }
      } else if constexpr (XL && Ns == 1 && 16 % R == 0) {  // base = j*R: lx(base + k) = lx(base) ^ k
        const int lb = lx(base);
#pragma unroll
        for (int k = 0; k < R; ++k) { x[lb ^ k] = v[m][k]; }
      } else if constexpr (XL) {
#pragma unroll
        for (int k = 0; k < R; ++k) { x[lx(base + k * Ns)] = v[m][k]; }
      } else if constexpr (Ns % 16 == 0) {  // padi(base + k*Ns) = padi(base) + k*(Ns + Ns/16)
        T2* xb = x + padi(base);
#pragma unroll
        for (int k = 0; k < R; ++k) { xb[k * (Ns + Ns / 16)] = v[m][k]; }
      } else if constexpr (Ns == 1 && 16 % R == 0) {  // base = j*R: the R outputs stay in one 16-run
        T2* xb = x + padi(base);
#pragma unroll
        for (int k = 0; k < R; ++k) { xb[k] = v[m][k]; }
      } else {
#pragma unroll
        for (int k = 0; k < R; ++k) { x[padi(base + k * Ns)] = v[m][k]; }
      }
    }
  }
  if (priv) {
    __builtin_amdgcn_wave_barrier();
  } else {
    __syncthreads();
  }
}

// XL: regions in the lx layout (row pass) instead of padi (column pass)
template<int L, int REG, int NFFT, int T, bool INV, bool PRIV, bool XL, int Ns, int... Rs> struct XStages;
template<int L, int REG, int NFFT, int T, bool INV, bool PRIV, bool XL, int Ns>
struct XStages<L, REG, NFFT, T, INV, PRIV, XL, Ns> {
  template<class TW> __device__ __forceinline__ static void run(T2*, const TW&, int) {}
};
template<int L, int REG, int NFFT, int T, bool INV, bool PRIV, bool XL, int Ns, int R, int... Rs>
struct XStages<L, REG, NFFT, T, INV, PRIV, XL, Ns, R, Rs...> {
  template<class TW> __device__ __forceinline__ static void run(T2* lds, const TW& tw, int tid) {
    xstage<L, REG, NFFT, T, INV, R, Ns, PRIV, XL>(lds, tw, tid);
    if constexpr (sizeof...(Rs) > 0) {
      // a workgroup-wide stage after a wave-private one reads what other waves wrote
      constexpr int next[] = {Rs...};
      if constexpr (stagePrivate<L, T, R, PRIV>() && !stagePrivate<L, T, next[0], PRIV>()) { __syncthreads(); }
    }
    XStages<L, REG, NFFT, T, INV, PRIV, XL, Ns * R, Rs...>::run(lds, tw, tid);
  }
};

// exp(-2 pi i m / L) from a table in global memory
struct GlobalTw {
  const T2* t;
  __device__ __forceinline__ T2 operator()(int m) const { return gld(t + m); }
  template<int NR, int RR> __device__ __forceinline__ T2 st(int p, int k, int stride) const { return (*this)(p * stride * k); }
  __device__ __forceinline__ T2 pk(int p, int k) const { return (*this)(p * k); }
};

// Exact product tables for the stages whose twiddles are n-th roots with n = Ns*R small (Ns <= 16
// butterfly phases p, R <= 16): entry (p, k) = w_n^(p*k) for p < P and 1 <= k <= K, one correctly
// rounded value per twiddle instead of the product of two, stored at p*KS + k - 1 with an odd row
// stride KS >= K. The 16 lanes of one LDS access phase read (p, k) for 16 consecutive p and one k:
// 16 entries KS apart, which an odd KS puts in 16 distinct 16-byte bank slots (64 banks). (A table
// of the n-th roots indexed by p*k puts them 2 to 16 in one slot for most k, padded or not.)
// `full` holds w_L^m at m, and w_n = w_L^stride.
template<int P, int K, int KS>
__device__ __forceinline__ void fillProd(const T2* full, int stride, T2* tab, int tid, int T) {
  static_assert(KS % 2 == 1 && KS >= K, "product table stride");
  for (int i = tid; i < P * KS; i += T) {
    const int p = i / KS, k = i % KS + 1;
    tab[i] = k <= K ? gld(full + p * k * stride) : make_double2(0, 0);
  }
}

// The low-part tables TA[i] = w^i (i < 64) below are read at i = (p*k*stride) & 63 for the lanes
// p of a wave and a fixed k; for even k many lanes of one LDS access phase would then hit the same
// banks (k = 8: eight distinct entries i = 0, 8, 16, ... in one bank row of 64 4-byte banks). TA
// is therefore stored with one pad slot after every 16 entries (entry i at i + i/16, 68 slots),
// which spreads those entries over the banks (about half the bank cycles of the unpadded table
// for the row pass's stages, a fifth less for the column pass's).
__device__ __forceinline__ constexpr int taPos(int i) { return i + (i >> 4); }
constexpr int TA_SLOTS = 68;
// table position p of TA (p < TA_SLOTS) -> entry index, -1 for a pad slot
__device__ __forceinline__ constexpr int taEntry(int p) { return (p % 17 == 16) ? -1 : 16 * (p / 17) + p % 17; }

// Row-pass twiddles staged in LDS (ROWTW_SIZE entries after the line regions and the
// round-off slots): w_W^m = TA[m & 63] * TB[m >> 6], TA[i] = w_W^i, TB[j] = w_W^(64 j),
// each twiddle a product of two correctly rounded table entries. Used for W >= 512; at
// W = 256 the fill (68 entries for a 16 KB block) costs more than the global reads save.
// The radix-F1 stages whose twiddles are (F1*F2)-th roots or coarser (the inverse stage before the
// last, Ns = F2, and the second forward stage, Ns = F0 dividing F2) read them from the product
// table of the (F1*F2)-th roots after it (fillProd: F2 rows, stride F1 - 1; at most T = W/F0 =
// F1*F2 slots are reserved): phase p of such a stage reads row p * (F1*F2) / (Ns*F1).
//
// The other twiddles are w_W^(p*k) for a phase p and k < 16 (the last forward stage, Ns*R = W,
// and the last inverse stage, p = tid, k < F0). TA[p*k & 63] for 16 consecutive p hits as few
// as a quarter of the bank slots for odd k; from W = 1024 up they are instead
// P2[(p >> 4) * k] * P1[p & 15][k] with P1[t][k] = w_W^(t*k) (t < 16, k < 16: a product table
// as above, stride 15) and P2[j] = w_W^(16 j) (j < W/16; one entry for the 16 lanes of an access
// phase): again the product of two correctly rounded entries, now without bank conflicts. Layout
// from W = 1024: the (F1*F2)-th-root product table (T slots), P1 (240), P2 (W/16), no TA/TB.
constexpr int ROWTW_SIZE = TA_SLOTS + 64;
template<int W, int T> constexpr int rowTwSlots() { return W >= 1024 ? T + 240 + W / 16 : ROWTW_SIZE + T; }
template<int W, int F1, int F2>
struct RowTw {
  static_assert(F1 % 2 == 0 && F2 * (F1 - 1) <= F1 * F2, "row product table");
  static constexpr bool PK = W >= 1024;
  static constexpr int PT = PK ? 0 : ROWTW_SIZE, P1 = F1 * F2, P2 = F1 * F2 + 240;
  const T2* tab;
  __device__ __forceinline__ T2 operator()(int m) const {
    static_assert(!PK, "no TA/TB table from W = 1024 up");
    return cmul(tab[taPos(m & 63)], tab[TA_SLOTS + (m >> 6)]);
  }
  // w_W^(p*k) for p*k < W and 1 <= k < 16
  __device__ __forceinline__ T2 pk(int p, int k) const {
    if constexpr (PK) {
      return cmul(tab[P2 + (p >> 4) * k], tab[P1 + (p & 15) * 15 + k - 1]);
    } else {
      return (*this)(p * k);
    }
  }
  template<int NR, int RR> __device__ __forceinline__ T2 st(int p, int k, int stride) const {
    if constexpr (RR == F1 && (F1 * F2) % NR == 0) {
      return tab[PT + p * ((F1 * F2) / NR) * (F1 - 1) + k - 1];
    } else if constexpr (NR == W && RR <= 16) {
      return pk(p, k);
    } else {
      return (*this)(p * stride * k);
    }
  }
};
template<int W, int T, int F1, int F2>
__device__ __forceinline__ void fillRowTw(const T2* twRow, T2* tab, int tid) {
  static_assert(W % 64 == 0 && W / 64 <= 64 && W % T == 0, "twiddle table");
  using TW = RowTw<W, F1, F2>;
  if constexpr (TW::PK) {
    fillProd<16, 15, 15>(twRow, 1, tab + TW::P1, tid, T);
    for (int j = tid; j < W / 16; j += T) { tab[TW::P2 + j] = gld(twRow + 16 * j); }
  } else {
    constexpr int N = TA_SLOTS + W / 64;
#pragma unroll
    for (int p = tid; p < N; p += T) {
      const int e = p < TA_SLOTS ? taEntry(p) : 64 * (p - TA_SLOTS);
      tab[p] = e < 0 ? make_double2(0, 0) : gld(twRow + e);
    }
  }
  fillProd<F2, F1 - 1, F1 - 1>(twRow, W / (F1 * F2), tab + TW::PT, tid, T);
}

// Column-pass twiddles staged in LDS after the data regions (colTwSize<H>() entries):
// w_H^m = TA[m & 63] * TB[m >> 6] for m < H, and for column h of the pair the mid
// twiddles w_ND^(a*l) = MA[h][l & 63] * MB[h][l >> 6] and S[h][k] = w_ND^(a*k*H/16).
// Each twiddle is a product of two correctly rounded table entries. Layout: TA (72 slots,
// padded as above), TB[H/64], then per column MA[64], MB[H/64], S[16] (MA is read at
// consecutive l, which needs no padding); sized per H so that small-H groups keep their LDS
// footprint (and so the number of groups per CU) as low as the data allows.
template<int H> constexpr int colTwPer() { return 64 + H / 64 + 16; }
template<int H> constexpr int colTwSize() { return TA_SLOTS + H / 64 + 2 * colTwPer<H>(); }
// With ROOTS (H >= 2304, plan [16, 16, H/256]), the second forward stage (Ns = 16, R = 16: 256th
// roots) and the second inverse stage (Ns = 16, R = H/256: (H/16)-th roots) read exact product
// tables (fillProd; roots, after the pair tables): A, 16 rows of stride 15, then B, 16 rows of stride
// (H/256 - 1) | 1, unless H/16 = 256 and A serves both.
// Below H = 4096 the stages whose twiddles are w_H^(p*k) (the last forward stage, Ns = 256, and the
// last inverse stage, Ns = H/16) read them as P2[(p >> 4) * k] * P1[p & 15][k], as the row pass
// does (P1: w_H^(t*k), 16 rows of stride 15; P2: w_H^(16 j), j < H/16), after B. (At H = 4096
// they stay with TA and TB: the two-pair pass would not fit in LDS.)
template<int H> constexpr int colProdStrideB() { return (H / 256 - 1) | 1; }
template<int H> constexpr int colRootSlots() {
  return 16 * 15 + (H == 4096 ? 0 : 16 * colProdStrideB<H>() + 240 + H / 16);
}
template<int H, bool ROOTS = false>
struct ColTw {
  static_assert(H % 64 == 0 && H / 64 <= 64, "twiddle table");
  static constexpr int MID = TA_SLOTS + H / 64;  // start of column 0's part
  const T2* tab;
  const T2* roots = nullptr;
  __device__ __forceinline__ T2 operator()(int m) const { return cmul(tab[taPos(m & 63)], tab[TA_SLOTS + (m >> 6)]); }
  static constexpr bool PK = ROOTS && H != 4096;
  static constexpr int P1 = PK ? 16 * 15 + 16 * colProdStrideB<H>() : 0, P2 = P1 + 240;
  // w_H^(p*k) for p*k < H and 1 <= k < 16
  __device__ __forceinline__ T2 pk(int p, int k) const {
    if constexpr (PK) {
      return cmul(roots[P2 + (p >> 4) * k], roots[P1 + (p & 15) * 15 + k - 1]);
    } else {
      return (*this)(p * k);
    }
  }
  template<int NR, int RR> __device__ __forceinline__ T2 st(int p, int k, int stride) const {
    if constexpr (ROOTS && NR == 256 && RR == 16) {
      return roots[p * 15 + k - 1];
    } else if constexpr (ROOTS && NR == H / 16 && RR == H / 256) {
      return roots[16 * 15 + p * colProdStrideB<H>() + k - 1];
    } else if constexpr (NR == H && RR <= 16) {
      return pk(p, k);
    } else {
      return (*this)(p * stride * k);
    }
  }
  __device__ __forceinline__ T2 mid(int h, int l) const {
    const T2* p = tab + MID + h * colTwPer<H>();
    return cmul(p[l & 63], p[64 + (l >> 6)]);
  }
  __device__ __forceinline__ T2 s(int h, int k) const { return tab[MID + h * colTwPer<H>() + 64 + H / 64 + k]; }
};
// Entry i of the twiddle table of the column pair (a0, a1).
template<int H>
__device__ __forceinline__ T2 colTwEntry(const JobDesc& J, int i, int a0, int a1) {
  constexpr int NHI = H / 64, MID = ColTw<H>::MID, PER = colTwPer<H>();
  T2 v;
  if (i < TA_SLOTS) {
    const int e = taEntry(i);
    v = e < 0 ? make_double2(0, 0) : gld(J.twCol + e);
  } else if (i < MID) {
    v = gld(J.twCol + 64 * (i - TA_SLOTS));
  } else {
    const int h = (i - MID) / PER, r = (i - MID) % PER;
    const int a = h ? a1 : a0;
    if (r < 64) {
      v = gld(J.twMidA + a * 64 + r);
    } else if (r < 64 + NHI) {
      v = gld(J.twMidB + a * J.nHi + (r - 64));
    } else {
      v = gld(J.twMidS + a * 16 + (r - 64 - NHI));
    }
  }
  return v;
}
template<int H>
__device__ __forceinline__ void fillColTw(const JobDesc& J, T2* tab, int a0, int a1, int tid, int T) {
  for (int i = tid; i < colTwSize<H>(); i += T) { tab[i] = colTwEntry<H>(J, i, a0, a1); }
}
template<int H>
__device__ __forceinline__ void fillColRoots(const JobDesc& J, T2* roots, int tid, int T) {
  static_assert(H % 256 == 0 && H / 256 <= 16, "column product tables");
  fillProd<16, 15, 15>(J.twCol, H / 256, roots, tid, T);
  if constexpr (H != 4096) {
    fillProd<16, H / 256 - 1, colProdStrideB<H>()>(J.twCol, 16, roots + 16 * 15, tid, T);
    using TW = ColTw<H, true>;
    fillProd<16, 15, 15>(J.twCol, 1, roots + TW::P1, tid, T);
    for (int j = tid; j < H / 16; j += T) { roots[TW::P2 + j] = gld(J.twCol + 16 * j); }
  }
}

// ---------------------------------------------------------------- column pass
// MULTI = false: grid (W/2, batch), group q handles the column pair q; LDS: two
// colRegionX<H>() regions (lx layout) and one twiddle table (colTwSize<H>() entries).
// MULTI = true: grid (G, batch) with G <= W/2, group x handles the pairs x, x + G, x + 2G,
// ... below W/2; LDS: the two regions and two tables. T threads; modes as k_col.
// For H >= 2304 the root tables (colRootSlots<H>() entries) follow the twiddle table(s).
// With several pairs per group, the next pair's column data and twiddle-table entries
// are requested right after the current pair's first stage and stay in registers while
// the current pair is transformed, squared and stored, so that HBM reads overlap the
// LDS work instead of each group waiting for its loads with nothing else to do. The next
// table goes to the other table buffer; a barrier at the end of each pair keeps the
// regions and that buffer free until every wave is done with them.
template<int H, int T, int C0, int C1, int C2, bool MULTI = false>
__global__ __launch_bounds__(T) void k_colx(const JobDesc* __restrict__ jobs, int mode, int fi, int oi) {
  static_assert(C0 * C1 * C2 == H && C0 == 16, "plan");
  extern __shared__ T2 lds[];
  constexpr int REG = colRegionX<H>();
  constexpr int NB0 = H / C0;
  constexpr int NBT = 2 * NB0;
  constexpr int MB0 = (NBT + T - 1) / T;
  constexpr int TWS = colTwSize<H>();
  constexpr int MT = (TWS + T - 1) / T;
  // exact root tables for two of the stages where LDS allows (one group per CU anyway)
  constexpr bool ROOTS = H >= 2304;
  T2* const roots = lds + 2 * REG + (MULTI ? 2 : 1) * TWS;
  const JobDesc& J = jobs[blockIdx.y];
  const int W = J.W, R = J.R, lgR = J.lgR;
  const int tid0 = threadIdx.x;
  // the column data streams through with non-temporal accesses from H = 768 up (at 512 x
  // 256 the 2 MB buffers stay cached between passes and plain accesses measured faster)
  constexpr bool NT = H >= 768;
  T2* io = J.freq[fi];
  const T2* other = (mode == COL_MUL) ? J.freq[oi] : nullptr;
  const size_t blockStride = (size_t)W * R;
  const int per = 2 * R;
  // Element (line l, column h of pair q) lies at offset (l/R)*W*R + 2qR + hR + l%R, below ND < 2^32.
  // The host runs this pass only when R divides H/16 = NB0, so for the lines l = j + k*NB0 of a
  // stage-0 butterfly the offset is off(j) + k*kStride: one add per element.
  const uint32_t kStride = (uint32_t)(NB0 >> lgR) * (uint32_t)(W * R);
  auto offset0 = [&](int q, int h, int j) -> uint32_t {
    return (uint32_t)(j >> lgR) * (uint32_t)(W * R) + (uint32_t)(2 * q * R + (h << lgR) + (j & (R - 1)));
  };

  // Lane order within 2R lanes: (column h, line-in-block r), so 2R lanes read 2R*16
  // contiguous bytes of one line block.
  T2 xin[MB0][C0];
  auto loadPair = [&](int q, int tid) {
#pragma unroll
    for (int m = 0; m < MB0; ++m) {
      const int g = tid + m * T;
      if (NBT % T == 0 || g < NBT) {
        const int h = (g >> lgR) & 1;
        const int j = ((g >> (lgR + 1)) << lgR) | (g & (R - 1));
        const uint32_t o = offset0(q, h, j);
#pragma unroll
        for (int k = 0; k < C0; ++k) { xin[m][k] = ldx<NT>(io + (o + k * kStride)); }
      }
    }
  };

  int q = blockIdx.x;
  // the column data is requested before the twiddle table is filled, so the two latencies overlap
  loadPair(q, tid0);
  fillColTw<H>(J, lds + 2 * REG, q, q ? W - q : W / 2, tid0, T);
  if constexpr (ROOTS) { fillColRoots<H>(J, roots, tid0, T); }
  __syncthreads();
  for (int it = 0;; ++it) {
    // a fresh copy of the thread index per pair: offsets derived from it are recomputed
    // in each pair instead of being hoisted out of the loop and held in registers
    int tid = tid0;
    if constexpr (MULTI) { asm volatile("" : "+v"(tid)); }
    const ColTw<H, ROOTS> tw{lds + 2 * REG + (it & 1) * TWS, roots};
    const int a0 = q, a1 = q ? W - q : W / 2;
    const size_t colBase = (size_t)2 * q * R;
    const T2 wa = gld(J.twPair + (q ? a0 : a1));  // column a0's pairing twiddle, or column W/2's for q = 0
    const int qn = MULTI ? q + (int)gridDim.x : W / 2;

    // forward stage 0 (radix C0, Ns = 1), fused with the load and the mid twiddle
#pragma unroll
    for (int m = 0; m < MB0; ++m) {
      const int g = tid + m * T;
      if (NBT % T == 0 || g < NBT) {
        const int h = (g >> lgR) & 1;
        const int j = ((g >> (lgR + 1)) << lgR) | (g & (R - 1));
        const T2 base = tw.mid(h, j);
        T2 v[C0];
#pragma unroll
        for (int k = 0; k < C0; ++k) { v[k] = cmul(xin[m][k], k ? cmul(base, tw.s(h, k)) : base); }
        Dft<C0, false>::run(v);
#pragma unroll
        for (int k = 0; k < C0; ++k) { lds[h * REG + (lx(j * C0) ^ k)] = v[k]; }
      }
    }
    // prefetch of the next pair: table entries first (written to LDS below, so their wait
    // does not cover the column data), then the column data
    T2 twn[MT];
    if (MULTI && qn < W / 2) {
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int i = tid + m * T;
        if (TWS % T == 0 || i < TWS) { twn[m] = colTwEntry<H>(J, i, qn, W - qn); }
      }
      loadPair(qn, tid);
    }
    __syncthreads();
    XStages<H, REG, 2, T, false, false, true, C0, C1, C2>::run(lds, tw, tid);
    if (MULTI && qn < W / 2) {
      T2* tabn = lds + 2 * REG + ((it + 1) & 1) * TWS;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int i = tid + m * T;
        if (TWS % T == 0 || i < TWS) { tabn[i] = twn[m]; }
      }
    }

    if (mode == COL_FWD) {
      for (int e = tid; e < 2 * H; e += T) {
        int lb = e / per;
        int wi = e - lb * per;
        int h = wi >> lgR;
        int b = lb * R + (wi & (R - 1));
        stx<NT>(io + (size_t)lb * blockStride + colBase + wi, lds[h * REG + lx(b)]);
      }
    } else {
      auto spos = [&](int h, int b) -> size_t { return (size_t)(b >> lgR) * blockStride + colBase + (h << lgR) + (b & (R - 1)); };
      if (q) {
        for (int b = tid; b < H; b += T) {
          int bp = H - 1 - b;
          T2 X = lds[lx(b)], Xp = lds[REG + lx(bp)];
          T2 w = cmul(wa, tw(b));
          if (other) {
            pairMul(X, Xp, gld(other + spos(0, b)), gld(other + spos(1, bp)), w);
          } else {
            pairSq(X, Xp, w);
          }
          lds[lx(b)] = X;
          lds[REG + lx(bp)] = Xp;
        }
      } else {
        for (int i = tid; i <= H; i += T) {
          if (i <= H / 2) {
            int b = i, bp = (H - i) % H;
            T2 X = lds[lx(b)], Xp = lds[lx(bp)];
            T2 w = tw(b);
            if (other) {
              pairMul(X, Xp, gld(other + spos(0, b)), gld(other + spos(0, bp)), w);
            } else {
              pairSq(X, Xp, w);
            }
            lds[lx(b)] = X;
            if (bp != b) { lds[lx(bp)] = Xp; }
          } else {
            int b = i - H / 2 - 1, bp = H - 1 - b;
            T2 X = lds[REG + lx(b)], Xp = lds[REG + lx(bp)];
            T2 w = cmul(wa, tw(b));
            if (other) {
              pairMul(X, Xp, gld(other + spos(1, b)), gld(other + spos(1, bp)), w);
            } else {
              pairSq(X, Xp, w);
            }
            lds[REG + lx(b)] = X;
            lds[REG + lx(bp)] = Xp;
          }
        }
      }
      __syncthreads();
      // inverse stages radix C1 then C2 (the reverse of the forward order would put the radix-C2
      // stage first, whose stride-C2 writes hit few LDS banks for C2 = 15)
      XStages<H, REG, 2, T, true, false, true, 1, C1, C2>::run(lds, tw, tid);

      // last inverse stage (radix C0, Ns = NB0), fused with the conjugate mid twiddle and the store
#pragma unroll
      for (int m = 0; m < MB0; ++m) {
        const int g = tid + m * T;
        if (NBT % T == 0 || g < NBT) {
          const int h = (g >> lgR) & 1;
          const int j = ((g >> (lgR + 1)) << lgR) | (g & (R - 1));
          T2 v[C0];
#pragma unroll
          for (int k = 0; k < C0; ++k) { v[k] = lds[h * REG + lx(j + k * NB0)]; }
#pragma unroll
          for (int k = 1; k < C0; ++k) { v[k] = cmulc(v[k], tw.pk(j, k)); }
          Dft<C0, true>::run(v);
          const T2 base = tw.mid(h, j);
          const uint32_t o = offset0(q, h, j);
#pragma unroll
          for (int k = 0; k < C0; ++k) { stx<NT>(io + (o + k * kStride), cmulc(v[k], k ? cmul(base, tw.s(h, k)) : base)); }
        }
      }
    }
    if (!MULTI || qn >= W / 2) { break; }
    __syncthreads();
    q = qn;
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
// [C2, C1, 16]; LDS (colRegion<H>() + colTwSize<H>())*16 bytes, twiddles from the LDS table.
template<int H, int C1, int C2>
__global__ __launch_bounds__(H / C2, 2) void k_coly(const JobDesc* __restrict__ jobs, int mode, int fi, int oi) {
  constexpr int C0 = 16;
  constexpr int T = H / C2;
  constexpr int NB0 = H / C0;
  static_assert(C0 * C1 * C2 == H && T % 64 == 0 && NB0 <= T, "plan");
  constexpr int REG = colRegion<H>();
  extern __shared__ T2 lds[];
  const ColTw<H> tw{lds + REG};
  const JobDesc& J = jobs[blockIdx.y];
  const int W = J.W, R = J.R, lgR = J.lgR;
  const int q = blockIdx.x, tid = threadIdx.x;
  const int a0 = q, a1 = q ? W - q : W / 2;
  T2* io = J.freq[fi];
  const size_t colBase = (size_t)2 * q * R;
  const size_t blockStride = (size_t)W * R;
  auto gpos = [&](int h, int l) -> size_t {
    return (size_t)(l >> lgR) * blockStride + colBase + (h << lgR) + (l & (R - 1));
  };
  const int jl1 = q ? T - 1 - tid : tid;  // column 1's last-stage butterfly
  fillColTw<H>(J, lds + REG, a0, a1, tid, T);
  __syncthreads();

  T2 x0[C2], x1[C2];

  // forward FFT of column h (mid twiddle for column index a), last stage for butterfly jl
  auto forward = [&](int h, int jl, T2* out) {
    if (tid < NB0) {
      const int j = tid;
      const T2 base = tw.mid(h, j);
      T2 v[C0];
#pragma unroll
      for (int k = 0; k < C0; ++k) {
        T2 x = gld(io + gpos(h, j + k * NB0));
        v[k] = cmul(x, k ? cmul(base, tw.s(h, k)) : base);
      }
      Dft<C0, false>::run(v);
#pragma unroll
      for (int k = 0; k < C0; ++k) { lds[padi(j * C0 + k)] = v[k]; }
    }
    __syncthreads();
    XStages<H, REG, 1, T, false, false, false, C0, C1>::run(lds, tw, tid);
#pragma unroll
    for (int k = 0; k < C2; ++k) { out[k] = lds[padi(jl + k * T)]; }
#pragma unroll
    for (int k = 1; k < C2; ++k) { out[k] = cmul(out[k], tw(jl * k)); }
    Dft<C2, false>::run(out);
    __syncthreads();
  };

  // inverse FFT of column h from registers (butterfly jl of the first stage), then store
  auto inverse = [&](int h, int jl, T2* in) {
    Dft<C2, true>::run(in);
#pragma unroll
    for (int k = 0; k < C2; ++k) { lds[padi(jl * C2 + k)] = in[k]; }
    __syncthreads();
    XStages<H, REG, 1, T, true, false, false, C2, C1>::run(lds, tw, tid);
    if (tid < NB0) {
      const int j = tid;
      T2 v[C0];
#pragma unroll
      for (int k = 0; k < C0; ++k) { v[k] = lds[padi(j + k * NB0)]; }
#pragma unroll
      for (int k = 1; k < C0; ++k) { v[k] = cmulc(v[k], tw(j * k)); }
      Dft<C0, true>::run(v);
      const T2 base = tw.mid(h, j);
#pragma unroll
      for (int k = 0; k < C0; ++k) {
        gst(io + gpos(h, j + k * NB0), cmulc(v[k], k ? cmul(base, tw.s(h, k)) : base));
      }
    }
    __syncthreads();
  };

  forward(0, tid, x0);
  forward(1, jl1, x1);

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
      const T2 w = cmul(wa, tw(b));
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
        const T2 w = h ? cmul(wa, tw(b)) : tw(b);
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

  inverse(0, tid, x0);
  inverse(1, jl1, x1);
}

// ---------------------------------------------------------------- row pass
// Weight data of one line read through the scalar cache: the line index is uniform across the
// group (readfirstlane), so the loads are scalar and the values live in SGPRs. Vector loads of
// them would also sit in the in-order vector memory counter behind a prefetch still in flight.
#define CONSTANT __attribute__((address_space(4)))
struct LineWs {
  double w0, w1, iw0, iw1;
  uint32_t e0, e1;
};
__device__ __forceinline__ LineWs lineWs(const LineW* p) {
  const CONSTANT LineW* q = (const CONSTANT LineW*)p;
  return {q->w[0], q->w[1], q->iw[0], q->iw[1], q->e[0], q->e[1]};
}
// Tagged stairway access through the global address space.
__device__ __forceinline__ uint64_t stairLoad(const uint64_t* p) {
  return __hip_atomic_load((const GLOBAL uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stairStore(uint64_t* p, uint64_t v) {
  __hip_atomic_store((GLOBAL uint64_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Polls N tagged carries until every tag is current. All N loads are in flight together and a
// round that finds any of them stale reloads all N (a current tag stays current for the rest of
// the launch), so the loop body has no per-carry branches. Bounded: a time-out is flagged to the
// host and the kernel still exits.
template<int N>
__device__ __forceinline__ void pollAll(const JobDesc& J, uint64_t* const* p, uint32_t tag, uint64_t* v) {
#pragma unroll
  for (int k = 0; k < N; ++k) { v[k] = stairLoad(p[k]); }
  for (uint32_t n = 0;; ++n) {
    bool ready = true;
#pragma unroll
    for (int k = 0; k < N; ++k) { ready &= (uint32_t)(v[k] >> 48) == tag; }
    if (ready) { break; }
    if (n >= OWL_SPIN_LIMIT) {
      atomicOr(J.flags + 2, 1u);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
#pragma unroll
    for (int k = 0; k < N; ++k) { v[k] = stairLoad(p[k]); }
  }
}

// What one group of a fused row launch does: block lb; for a first visit (first) only the
// carry-out is published, for a redo (redo) the block is transformed again and nothing is
// published; otherwise it waits for block pred's carries (wrap: the carries of the last block,
// which enter block 0 one column to the right) and publishes its own.
struct FusedItem {
  int lb, pred;
  bool first, redo, wrap, none;
};
// Chunked order (opts bit 4, nb >= 8, grid 8 * (ceil(nb/8) + 1)): group i runs on XCD i mod 8
// (workgroups go to the eight XCDs in turn) and XCD x works through one contiguous chunk of
// blocks B_x .. B_{x+1}-1: its first group visits B_x first (publish only), the next groups take
// the following blocks, and its last group redoes B_{x+1}, the head of the next chunk (block 0
// after the last chunk) once the chunk's last carries are in. Every wait is then on a group of
// the same XCD with a lower index, which that XCD has already started, or on the first group of
// a chunk, which never waits: no start counter is needed. Groups beyond a chunk's items exit.
// Results do not depend on the XCD assignment (any group order gives the same residues); only
// forward progress while kernels of other queues hold parts of the device relies on it.
// Ticket order (otherwise, grid nb + 1): the group's start ticket is its block; ticket nb redoes
// block 0.
// Item of group i of a chunked launch over nb >= 8 blocks (also compiled for the host, where a
// CPU test checks the mapping).
__host__ __device__ inline FusedItem chunkedItem(int nb, int i) {
  FusedItem it{0, 0, false, false, false, false};
  const int x = i & 7, j = i >> 3, q = nb >> 3, rem = nb & 7;
  const int b0 = x * q + (x < rem ? x : rem), c = q + (x < rem ? 1 : 0);
  if (j > c) {
    it.none = true;
  } else if (j == 0) {
    it.lb = b0;
    it.first = true;
  } else if (j < c) {
    it.lb = b0 + j;
    it.pred = it.lb - 1;
  } else {
    const int h = b0 + c;  // the next chunk's head; nb after the last chunk
    it.lb = (h == nb) ? 0 : h;
    it.pred = h - 1;
    it.redo = true;
    it.wrap = h == nb;
  }
  return it;
}
__device__ __forceinline__ FusedItem fusedItem(const JobDesc& J, int opts, int* slot) {
  const int nb = J.nb;
  if (opts & 4) { return chunkedItem(nb, blockIdx.x); }
  FusedItem it{0, 0, false, false, false, false};
  const int g = __builtin_amdgcn_readfirstlane(fusedBlock(J, slot));
  it.lb = (g == nb) ? 0 : g;
  it.pred = g - 1;
  it.first = g == 0;
  it.redo = g == nb;
  it.wrap = g == nb;
  return it;
}

// Pieces of the row pass over one block of R lines, T = W/F0 threads. Thread tid owns columns
// c = tid + k*T (k < F0) of the R lines; v[r][k] holds the inverse transform of (line r,
// column c), then its balanced digits (as doubles), then the weighted digits.
template<int W, int R, int F0, int F1, int F2> struct RowPass {
  static constexpr int T = W / F0;
  static constexpr int LW = R * W;
  static constexpr int NL = LW / T;  // state elements per thread
  static constexpr int REG = rowRegion<W, R>();
  static constexpr bool LDS_TW = W >= 512;
  // the block streams through with non-temporal accesses from W = 512 up
  static constexpr bool NT = W >= 512;
  using Tw = std::conditional_t<LDS_TW, RowTw<W, F1, F2>, GlobalTw>;

  // Thread tid owns the columns colBase(tid) + k*T (k < F0): colBase permutes the low four bits of
  // tid so that the first forward stage's LDS writes (positions colBase*F0 + k, one per lane) of
  // 8 consecutive lanes fall in 8 distinct bank slots (LDS writes serve 8 lanes per cycle, 16-byte
  // slots mod 8); with colBase = tid they pair up for F0 = 2, 4 and 8. The host lays out the word
  // metadata with the same permutation.
  __device__ __forceinline__ static int colBase(int tid) {
    const int b = tid & 15, hi = tid & ~15;
    if constexpr (F0 == 2) {
      return hi | (b & 3) | ((b >> 3) << 2) | (((b >> 2) & 1) << 3);
    } else if constexpr (F0 == 4) {
      return hi | (b & 1) | (((b >> 3) & 1) << 1) | (((b >> 1) & 3) << 2);
    } else if constexpr (F0 == 8) {
      return hi | ((b >> 3) & 1) | ((b & 7) << 1);
    } else {
      return tid;
    }
  }

  // element e = s*R + r of the block (slot s, line r) for e = tid + i*T
  __device__ __forceinline__ static void load(const T2* blk, T2* P, int tid) {
#pragma unroll
    for (int i = 0; i < NL; ++i) { P[i] = ldx<NT>(blk + tid + i * T); }
  }
  // ... into the line regions in frequency order
  // LDS position of the block's element e = tid + i*T (slot s = e/R, line r = e%R) in the line
  // regions, frequency order. When T/(2R) is a multiple of 16 the slot of element i is
  // s0 + i*T/R with the same parity as s0, so its frequency moves by +-i*T/(2R) from s0's and its
  // padded position by +-i*(T/(2R))*17/16: one multiply-add per element instead of the division
  // and the slot-to-frequency mapping (the one exception, frequency W/2 in slot 1, is i = 0).
  static constexpr bool LINEAR_SLOTS = T % (2 * R) == 0 && (T / (2 * R)) % 16 == 0;
  struct SlotPos {
    int base, delta, first;
    __device__ __forceinline__ SlotPos(int tid) {
      if constexpr (LINEAR_SLOTS) {
        const int s0 = tid / R, r = tid % R, q0 = s0 >> 1;
        const bool odd = s0 & 1;
        base = r * REG + (odd ? padi(W - q0) : padi(q0));
        delta = (odd ? -1 : 1) * (T / (2 * R)) * 17 / 16;
        first = r * REG + padi(slotFreq(s0, W));
      } else {
        base = delta = first = 0;
      }
    }
    __device__ __forceinline__ int operator()(int tid, int i) const {
      if constexpr (LINEAR_SLOTS) {
        return i == 0 ? first : base + i * delta;
      } else {
        const int e = tid + i * T;
        return (e % R) * REG + padi(slotFreq(e / R, W));
      }
    }
  };
  __device__ __forceinline__ static void scatter(T2* lds, const T2* P, int tid) {
    const SlotPos pos(tid);
#pragma unroll
    for (int i = 0; i < NL; ++i) { lds[pos(tid, i)] = P[i]; }
  }
  // inverse W-point FFT of the R lines (LDS stages, then the radix-F0 stage into registers)
  __device__ __forceinline__ static void inverse(T2* lds, const Tw& tw, int tid, int cb, T2 (&v)[R][F0]) {
    XStages<W, REG, R, T, true, true, false, 1, F2, F1>::run(lds, tw, tid);
    __syncthreads();
    // last inverse stage (radix F0, Ns = T): butterfly cb of each line -> columns cb + k*T
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int k = 0; k < F0; ++k) { v[r][k] = lds[r * REG + padi(cb + k * T)]; }
#pragma unroll
      for (int k = 1; k < F0; ++k) { v[r][k] = cmulc(v[r][k], tw.pk(cb, k)); }
      Dft<F0, true>::run(v[r]);
    }
  }
  // Unweight, round, carry along the block's 2R words per column. The carry chain runs in
  // FP64: every value is an integer of magnitude below 2^51, so sums, scaling by powers of two
  // and floor are exact. Word u of b bits splits into the carry floor(u/2^b + 1/2) and the
  // balanced digit u - carry*2^b in [-2^(b-1), 2^(b-1)), as the integer chain would.
  // cs[k] = the block's carry-out of column k; roe = max round-off.
  // Word metadata (J.rowMeta, one 64-bit word per thread and block): for the thread's word
  // w = (k*R + r)*2 + rep, bit w says whether it is big (bitsLo + 1 bits) and bit 32 + w whether
  // its weight exponent wrapped (weight halved, inverse weight doubled).
  static_assert(F0 * R * 2 <= 32, "row metadata");
  __device__ __forceinline__ static int bigBit(uint64_t meta, int k, int r, int rep) {
    return (int)((uint32_t)meta >> ((k * R + r) * 2 + rep)) & 1;
  }
  __device__ __forceinline__ static int wrapBit(uint64_t meta, int k, int r, int rep) {
    return (int)((uint32_t)(meta >> 32) >> ((k * R + r) * 2 + rep)) & 1;
  }
  __device__ __forceinline__ static void carry(const JobDesc& J, const LineW* lwBase, const ColW* cw, uint64_t meta,
                                               T2 (&v)[R][F0], double* cs, double& roe, bool llHere, bool mul3) {
#pragma unroll
    for (int k = 0; k < F0; ++k) { cs[k] = 0; }
    const double m3 = mul3 ? 3.0 : 1.0;
    const int bitsLo = J.bitsLo;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const LineWs lw = lineWs(lwBase + r);
#pragma unroll
      for (int k = 0; k < F0; ++k) {
#pragma unroll
        for (int rep = 0; rep < 2; ++rep) {
          const int b = bitsLo + bigBit(meta, k, r, rep);
          const double iw = __builtin_amdgcn_ldexp((rep ? lw.iw1 : lw.iw0) * cw[k].iw, wrapBit(meta, k, r, rep));
          const double x = rep ? v[r][k].y : v[r][k].x;
          const double t = fma(x, iw, OWL_RNDVAL) - OWL_RNDVAL;
          roe = fmax(roe, fabs(fma(x, iw, -t)));
          double u = fma(t, m3, cs[k]);
          if (llHere && r == 0 && k == 0 && rep == 0) { u -= 2; }
          const double c = floor(__builtin_amdgcn_ldexp(u, -b) + 0.5);
          const double dgt = u - __builtin_amdgcn_ldexp(c, b);
          cs[k] = c;
          if (rep) {
            v[r][k].y = dgt;
          } else {
            v[r][k].x = dgt;
          }
        }
      }
    }
  }
  // Ripples the predecessor's carries through each column's 2R words until they die.
  __device__ __forceinline__ static void ripple(const JobDesc& J, uint64_t meta, T2 (&v)[R][F0], const uint64_t* got) {
    const int bitsLo = J.bitsLo;
#pragma unroll
    for (int k = 0; k < F0; ++k) {
      double carry = (double)unpackCarry(got[k]);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (carry != 0) {
          const int b0 = bitsLo + bigBit(meta, k, r, 0);
          double t = v[r][k].x + carry;
          carry = floor(__builtin_amdgcn_ldexp(t, -b0) + 0.5);
          v[r][k].x = t - __builtin_amdgcn_ldexp(carry, b0);
          if (r == R - 1) {
            v[r][k].y += carry;
            carry = 0;
          } else if (carry != 0) {
            const int b1 = bitsLo + bigBit(meta, k, r, 1);
            t = v[r][k].y + carry;
            carry = floor(__builtin_amdgcn_ldexp(t, -b1) + 0.5);
            v[r][k].y = t - __builtin_amdgcn_ldexp(carry, b1);
          }
        }
      }
    }
  }
  // Weight, then the forward W-point FFT's stage 0 (radix F0, Ns = 1), in registers.
  __device__ __forceinline__ static void weightStage0(const LineW* lwBase, const ColW* cw, uint64_t meta, T2 (&v)[R][F0]) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const LineWs lw = lineWs(lwBase + r);
#pragma unroll
      for (int k = 0; k < F0; ++k) {
        const double w0 = __builtin_amdgcn_ldexp(lw.w0 * cw[k].w, -wrapBit(meta, k, r, 0));
        const double w1 = __builtin_amdgcn_ldexp(lw.w1 * cw[k].w, -wrapBit(meta, k, r, 1));
        v[r][k] = make_double2(v[r][k].x * w0, v[r][k].y * w1);
      }
      Dft<F0, false>::run(v[r]);
    }
  }
  // The other forward stages in LDS. The caller has made sure that no thread still reads the
  // line regions.
  __device__ __forceinline__ static void forwardLds(const T2 (&v)[R][F0], T2* lds, const Tw& tw, int tid, int cb) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
      for (int k = 0; k < F0; ++k) { lds[r * REG + padi(cb * F0 + k)] = v[r][k]; }
    }
    __syncthreads();
    XStages<W, REG, R, T, false, true, false, F0, F1, F2>::run(lds, tw, tid);
    __syncthreads();
  }
  // the block from the line regions, slot order
  __device__ __forceinline__ static void store(T2* blk, const T2* lds, int tid) {
    const SlotPos pos(tid);
#pragma unroll
    for (int i = 0; i < NL; ++i) { stx<NT>(blk + tid + i * T, lds[pos(tid, i)]); }
  }
  // the stairway carries this block waits for: block it.pred's, shifted by one column when they
  // wrap around into block 0
  __device__ __forceinline__ static void pollPredecessor(const JobDesc& J, const FusedItem& it, int cb, uint32_t tag,
                                                         uint64_t* got) {
    uint64_t* src[F0];
#pragma unroll
    for (int k = 0; k < F0; ++k) {
      const int c = cb + k * T;
      src[k] = J.stair + (size_t)it.pred * W + (it.wrap ? (c == 0 ? W - 1 : c - 1) : c);
    }
    pollAll<F0>(J, src, tag, got);
    if (it.redo) {  // the block's first visit has read the block before the redo rewrites it
#pragma unroll
      for (int k = 0; k < F0; ++k) { src[k] = J.stair + (size_t)it.lb * W + cb + k * T; }
      uint64_t seen[F0];
      pollAll<F0>(J, src, tag, seen);
    }
  }
  __device__ __forceinline__ static void publish(const JobDesc& J, int g, int cb, uint32_t tag, const double* cs) {
#pragma unroll
    for (int k = 0; k < F0; ++k) { stairStore(J.stair + (size_t)g * W + cb + k * T, packCarry((int64_t)cs[k], tag)); }
  }
};

// Grid (ROW_FUSED: nb + 1, or 8 * (ceil(nb/8) + 1) in chunked order; else nb; batch), T = W/F0
// threads, LDS R*rowRegion<W, R>()*16 + 64
// (+ rowTwSlots<W, T>()*16 for W >= 512) bytes: line regions, per-wave round-off, twiddle tables.
// Modes, options and the tagged stairway as k_row.
template<int W, int R, int F0, int F1, int F2>
__global__ __launch_bounds__(W / F0) void k_rowx(const JobDesc* __restrict__ jobs, int mode, int fi, int wi, int opts,
                                                 uint32_t roeSlot, uint32_t tag) {
  static_assert(F0 * F1 * F2 == W, "plan");
  using P = RowPass<W, R, F0, F1, F2>;
  constexpr int T = P::T, LW = P::LW, REG = P::REG;
  extern __shared__ T2 lds[];
  float* roePart = reinterpret_cast<float*>(lds + R * REG);  // one float per wave, after the line regions
  T2* twTab = lds + R * REG + 4;                               // after the 64 bytes of round-off slots
  const JobDesc& J = jobs[blockIdx.y];
  const typename P::Tw tw{P::LDS_TW ? twTab : J.twRow};
  const int nb = J.nb;
  __shared__ int ticketSlot;
  // the block index is uniform: line weights then come through the scalar cache
  FusedItem it{(int)blockIdx.x, 0, false, false, false, false};
  if (mode == ROW_FUSED) { it = fusedItem(J, opts, &ticketSlot); }
  if (it.none) { return; }
  const int lb = it.lb;
  const int tid = threadIdx.x, cb = P::colBase(tid);
  const bool ll = opts & 1, mul3 = opts & 2;
  T2* blk = J.freq[fi] + (size_t)lb * LW;
  if constexpr (P::LDS_TW) { fillRowTw<W, T, F1, F2>(J.twRow, twTab, tid); }  // read after the first barrier below

  T2 v[R][F0];
  ColW cw[F0];
#pragma unroll
  for (int k = 0; k < F0; ++k) { cw[k] = gld(J.colW + cb + k * T); }
  const uint64_t meta = gld(J.rowMeta + (size_t)lb * T + tid);
  const LineW* lwBase = J.lineW + lb * R;

  if (mode == ROW_LEADIN) {
    const int2* src = J.words[wi] + (size_t)lb * LW;
#pragma unroll
    for (int k = 0; k < F0; ++k) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int2 d = gld(src + (cb + k * T) * R + r);
        v[r][k] = make_double2(d.x, d.y);
      }
    }
  } else {
    {
      T2 x[P::NL];
      P::load(blk, x, tid);
      P::scatter(lds, x, tid);
    }
    __syncthreads();
    P::inverse(lds, tw, tid, cb, v);
    double roe = 0;
    double cs[F0];
    P::carry(J, lwBase, cw, meta, v, cs, roe, ll && lb == 0 && cb == 0, mul3);
    if (mode == ROW_CARRYOUT) {
      roeReport(J, roeSlot, (float)roe);
      int2* dst = J.words[wi] + (size_t)lb * LW;
#pragma unroll
      for (int k = 0; k < F0; ++k) {
        const int c = cb + k * T;
        gst(J.carry[0] + (size_t)lb * W + c, (int64_t)cs[k]);
#pragma unroll
        for (int r = 0; r < R; ++r) { gst(dst + c * R + r, make_int2((int)v[r][k].x, (int)v[r][k].y)); }
      }
      return;
    }
    if (!it.redo) { P::publish(J, lb, cb, tag, cs); }
    if (it.first) {
      roeReport(J, roeSlot, (float)roe);
      return;
    }
    roeWave(roeSlot, (float)roe, roePart);
    uint64_t got[F0];
    P::pollPredecessor(J, it, cb, tag, got);
    P::ripple(J, meta, v, got);
  }
  P::weightStage0(lwBase, cw, meta, v);
  __syncthreads();  // every thread has finished reading the inverse's LDS
  if (mode == ROW_FUSED) { roeGroup(J, roeSlot, roePart, T / 64); }
  P::forwardLds(v, lds, tw, tid, cb);
  P::store(blk, lds, tid);
}

// ---------------------------------------------------------------- name tables
struct OwlKernelEntry {
  const char* name;
  const void* fn;
  unsigned maxThreads;
};
