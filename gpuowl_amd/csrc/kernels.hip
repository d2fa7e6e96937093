// HIP kernels for one PRP/LL modular squaring mod 2^E - 1 as an FP64 IBDWT, for gfx950.
//
// One steady-state iteration is two in-place passes over the FP64 state:
//   k_row (ROW_FUSED)  per block of R lines: inverse W-point FFT, unweight, round,
//                      carry, stairway carry hand-off, weight, forward W-point FFT.
//                      Replaces the reference carryFused (src/cl/carryfused.cl:21-231).
//   k_col (COL_SQUARE) per pair of columns: mid twiddle, forward H-point FFT,
//                      real-signal pairing + square, inverse FFT, mid twiddle.
//                      Replaces fftMiddleIn + tailSquare + fftMiddleOut
//                      (src/cl/fftmiddlein.cl:8-48, src/cl/tailsquare.cl:225-313,
//                      src/cl/fftmiddleout.cl:8-58) in one pass.
// Boundary kernels: k_row ROW_LEADIN (fftP, src/cl/fftp.cl:10-33), k_row ROW_CARRYOUT
// (fftW + carryA, src/cl/fftw.cl:9-19, src/cl/carry.cl:9-41), k_carryb (carryB,
// src/cl/carryb.cl:5-32), k_col COL_MUL / COL_FWD (tailMul, src/cl/tailmul.cl:19-127),
// k_words_in/out (transposeIn/Out, src/cl/transpose.cl:8-40), k_residue
// (readResidue, src/cl/etc.cl:9-15), k_is_equal (isEqual, src/cl/etc.cl:34-41),
// k_sum64 (sum64, src/cl/etc.cl:19-29).
#include "kcommon.h"

// ---------------------------------------------------------------- LDS Stockham FFT
// Elements per thread staged in registers during one stage. The host only
// chooses plans where ceil(L*nfft/R / T) * R <= OWL_EPT for every stage.
#define OWL_EPT 16

// One radix-R autosort stage over nfft independent L-point FFTs stored
// back to back in LDS. Ns = product of the radices of the earlier stages.
// tw[m] = exp(-2 pi i m / L). Not inlined: with every radix inlined into one
// kernel the register allocator does not share registers between the stages.
template<int R, bool INV, int T>
__device__ __noinline__ void fftStage(T2* lds, int L, int nfft, int Ns, const T2* __restrict__ tw, int tid) {
  constexpr int MB = OWL_EPT / R;
  const int nbf = L / R;
  const int total = nbf * nfft;
  const int stride = L / (Ns * R);
  T2 v[MB][R];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    int g = tid + m * T;
    if (g < total) {
      int f = g / nbf;
      int j = g - f * nbf;
      const T2* x = lds + f * L;
      int p = j % Ns;
#pragma unroll
      for (int k = 0; k < R; ++k) { v[m][k] = x[j + k * nbf]; }
      if (p) {
        int st = p * stride;
#pragma unroll
        for (int k = 1; k < R; ++k) {
          T2 w = gld(tw + st * k);
          v[m][k] = INV ? cmulc(v[m][k], w) : cmul(v[m][k], w);
        }
      }
      Dft<R, INV>::run(v[m]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    int g = tid + m * T;
    if (g < total) {
      int f = g / nbf;
      int j = g - f * nbf;
      T2* x = lds + f * L;
      int p = j % Ns;
      int base = (j - p) * R + p;
#pragma unroll
      for (int k = 0; k < R; ++k) { x[base + k * Ns] = v[m][k]; }
    }
  }
  __syncthreads();
}

// Full mixed-radix FFT; the caller must have synchronized after filling LDS.
template<bool INV, int T>
__device__ void fftLds(T2* lds, int L, int nfft, const uint32_t* plan, int nStages, const T2* __restrict__ tw, int tid) {
  int Ns = 1;
  for (int s = 0; s < nStages; ++s) {
    int R = plan[s];
    switch (R) {
      case 2: fftStage<2, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 3: fftStage<3, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 4: fftStage<4, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 5: fftStage<5, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 6: fftStage<6, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 7: fftStage<7, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 8: fftStage<8, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 9: fftStage<9, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 10: fftStage<10, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 11: fftStage<11, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 12: fftStage<12, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 13: fftStage<13, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 14: fftStage<14, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 15: fftStage<15, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      case 16: fftStage<16, INV, T>(lds, L, nfft, Ns, tw, tid); break;
      default: break;  // the host never builds such a plan
    }
    Ns *= R;
  }
}

// ---------------------------------------------------------------- row pass
// Grid (nb + 1, batch) for ROW_FUSED, (nb, batch) otherwise; LDS R*W*16 bytes.
// opts: bit0 = LL (-2 as carry-in of word 0), bit1 = MUL3 (x3 before the carry).
// ROW_FUSED: freq[fi] -> freq[fi] in place (group nb redoes block 0 with the
//   wrap-around carry, as the reference carryFused does with group H). tag must
//   differ from the previous fused launch's (1..0xFFFE).
// ROW_LEADIN: words[wi] -> freq[fi].
// ROW_CARRYOUT: freq[fi] -> words[wi] + carry[0] (block-normalized words).
template<int T>
__global__ __launch_bounds__(T) void k_row(const JobDesc* __restrict__ jobs, int mode, int fi, int wi, int opts,
                                           uint32_t roeSlot, uint32_t tag) {
  extern __shared__ T2 lds[];
  const JobDesc& J = jobs[blockIdx.y];
  const int W = J.W, R = J.R, nb = J.nb;
  __shared__ int ticketSlot;
  const int g = (mode == ROW_FUSED) ? fusedBlock(J, &ticketSlot) : (int)blockIdx.x;
  const int lb = (g == nb) ? 0 : g;
  const int tid = threadIdx.x;
  const int LW = R * W;
  T2* blk = J.freq[fi] + (size_t)lb * LW;
  int2* ws = (int2*)lds;  // digit view of an LDS slot (thread-private during the carry phase)
  const bool ll = opts & 1, mul3 = opts & 2;

  if (mode == ROW_LEADIN) {
    const int2* src = J.words[wi] + (size_t)lb * LW;
    for (int c = tid; c < W; c += T) {
      for (int r = 0; r < R; ++r) { ws[2 * (r * W + c)] = gld(src + c * R + r); }
    }
  } else {
    for (int e = tid; e < LW; e += T) {
      int s = e / R;
      int r = e - s * R;
      lds[r * W + slotFreq(s, W)] = gld(blk + e);
    }
    __syncthreads();
    fftLds<true, T>(lds, W, R, J.rowPlan, J.nRow, J.twRow, tid);

    float roe = 0;
    for (int c = tid; c < W; c += T) {
      const ColW cw = gld(J.colW + c);
      int64_t s = 0;
      for (int r = 0; r < R; ++r) {
        const int l = lb * R + r;
        const LineW lw = gld(J.lineW + l);
        T2 z = lds[r * W + c];
        int32_t d[2];
#pragma unroll
        for (int rep = 0; rep < 2; ++rep) {
          WordW ww = wordWeight(J, lw, cw, rep);
          int64_t u = roundWord((rep ? z.y : z.x) * ww.iw, roe);
          if (mul3) { u *= 3; }
          if (ll && l == 0 && c == 0 && rep == 0) { u -= 2; }
          u += s;
          int64_t dd = lowBits64(u, ww.bits);
          s = (u - dd) >> ww.bits;
          d[rep] = (int32_t)dd;
        }
        ws[2 * (r * W + c)] = make_int2(d[0], d[1]);
      }
      if (mode == ROW_CARRYOUT) {
        gst(J.carry[0] + (size_t)lb * W + c, s);
      } else if (g < nb) {
        __hip_atomic_store(J.stair + (size_t)g * W + c, packCarry(s, tag), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    roeReport(J, roeSlot, roe);

    if (mode == ROW_CARRYOUT) {
      int2* dst = J.words[wi] + (size_t)lb * LW;
      for (int c = tid; c < W; c += T) {
        for (int r = 0; r < R; ++r) { gst(dst + c * R + r, ws[2 * (r * W + c)]); }
      }
      return;
    }

    // Stairway: block g's carries were published above; take the previous block's.
    // Group 0 does not write: group nb redoes block 0 with the wrap-around carry.
    if (g == 0) { return; }
    const int pred = g - 1;
    for (int c = tid; c < W; c += T) {
      // block 0 takes the carry of the last block of the previous column (word order
      // wraps from column W-1 to column 0 because 2^E == 1 mod 2^E - 1)
      int srcCol = (g == nb) ? (c == 0 ? W - 1 : c - 1) : c;
      int64_t carry = unpackCarry(pollCarry(J, J.stair + (size_t)pred * W + srcCol, tag));
      if (g == nb) {
        // block 0 is rewritten in place: wait until group 0 has published a carry,
        // which it computes only after every one of its loads of block 0 has returned
        pollCarry(J, J.stair + c, tag);
      }
      // Ripple the incoming carry through this block's 2R words; what survives the
      // block (rare: the host picks R so the carry dies well inside it) stays in the
      // last word. Leaving it all in the second word, as a one-step hand-off does,
      // makes the iteration diverge below about 12 bits per word.
      const ColW cw = gld(J.colW + c);
      for (int r = 0; r < R && carry; ++r) {
        const LineW lw = gld(J.lineW + lb * R + r);
        int2 d = ws[2 * (r * W + c)];
        WordW w0 = wordWeight(J, lw, cw, 0);
        int64_t t = (int64_t)d.x + carry;
        int64_t nd = lowBits64(t, w0.bits);
        carry = (t - nd) >> w0.bits;
        d.x = (int32_t)nd;
        if (r == R - 1) {
          d.y += (int32_t)carry;
          carry = 0;
        } else if (carry) {
          WordW w1 = wordWeight(J, lw, cw, 1);
          t = (int64_t)d.y + carry;
          nd = lowBits64(t, w1.bits);
          carry = (t - nd) >> w1.bits;
          d.y = (int32_t)nd;
        }
        ws[2 * (r * W + c)] = d;
      }
    }
  }

  // Weight the digits and run the forward row FFT.
  for (int c = tid; c < W; c += T) {
    const ColW cw = gld(J.colW + c);
    for (int r = 0; r < R; ++r) {
      const LineW lw = gld(J.lineW + lb * R + r);
      int2 d = ws[2 * (r * W + c)];
      WordW w0 = wordWeight(J, lw, cw, 0);
      WordW w1 = wordWeight(J, lw, cw, 1);
      lds[r * W + c] = make_double2(d.x * w0.w, d.y * w1.w);
    }
  }
  __syncthreads();
  fftLds<false, T>(lds, W, R, J.rowPlan, J.nRow, J.twRow, tid);
  for (int e = tid; e < LW; e += T) {
    int s = e / R;
    int r = e - s * R;
    gst(blk + e, lds[r * W + slotFreq(s, W)]);
  }
}

// ---------------------------------------------------------------- column pass
// Grid (W/2, batch); LDS 2*H*16 bytes. Group q owns slots 2q and 2q+1.
// COL_SQUARE: freq[fi] <- square, in place.   COL_MUL: freq[fi] <- freq[fi] * spectrum freq[oi].
// COL_FWD: freq[fi] <- forward column transform only (the spectrum of a second operand).
template<int T>
__global__ __launch_bounds__(T) void k_col(const JobDesc* __restrict__ jobs, int mode, int fi, int oi) {
  extern __shared__ T2 lds[];
  const JobDesc& J = jobs[blockIdx.y];
  const int W = J.W, H = J.H, R = J.R;
  const int q = blockIdx.x;
  const int tid = threadIdx.x;
  const int a0 = q, a1 = q ? W - q : W / 2;
  T2* io = J.freq[fi];
  const int per = 2 * R;
  const size_t colBase = (size_t)2 * q * R;

  for (int e = tid; e < 2 * H; e += T) {
    int lb = e / per;
    int wi = e - lb * per;
    int h = wi / R;
    int l = lb * R + wi - h * R;
    T2 v = gld(io + (size_t)lb * W * R + colBase + wi);
    lds[h * H + l] = cmul(v, midTw(J, h ? a1 : a0, l));
  }
  __syncthreads();
  fftLds<false, T>(lds, H, 2, J.colPlan, J.nCol, J.twCol, tid);

  if (mode == COL_FWD) {
    for (int e = tid; e < 2 * H; e += T) {
      int lb = e / per;
      int wi = e - lb * per;
      int h = wi / R;
      int b = lb * R + wi - h * R;
      gst(io + (size_t)lb * W * R + colBase + wi, lds[h * H + b]);
    }
    return;
  }

  const T2* other = (mode == COL_MUL) ? J.freq[oi] : nullptr;
  // position of spectrum bin b of half h in a COL_FWD output buffer
  auto spos = [&](int h, int b) -> size_t { return (size_t)(b / R) * W * R + colBase + h * R + (b % R); };

  const T2* twCol = J.twCol;
  if (q) {
    const T2 wa = gld(J.twPair + a0);
    for (int b = tid; b < H; b += T) {
      int bp = H - 1 - b;
      T2 X = lds[b], Xp = lds[H + bp];
      T2 w = cmul(wa, gld(twCol + b));
      if (other) {
        pairMul(X, Xp, gld(other + spos(0, b)), gld(other + spos(1, bp)), w);
      } else {
        pairSq(X, Xp, w);
      }
      lds[b] = X;
      lds[H + bp] = Xp;
    }
  } else {
    // column 0 pairs b with H-b (b = 0 and b = H/2 pair with themselves);
    // column W/2 pairs b with H-1-b.
    const T2 wa = gld(J.twPair + W / 2);
    for (int i = tid; i <= H; i += T) {
      if (i <= H / 2) {
        int b = i, bp = (H - i) % H;
        T2 X = lds[b], Xp = lds[bp];
        T2 w = gld(twCol + b);
        if (other) {
          pairMul(X, Xp, gld(other + spos(0, b)), gld(other + spos(0, bp)), w);
        } else {
          pairSq(X, Xp, w);
        }
        lds[b] = X;
        if (bp != b) { lds[bp] = Xp; }
      } else {
        int b = i - H / 2 - 1, bp = H - 1 - b;
        T2 X = lds[H + b], Xp = lds[H + bp];
        T2 w = cmul(wa, gld(twCol + b));
        if (other) {
          pairMul(X, Xp, gld(other + spos(1, b)), gld(other + spos(1, bp)), w);
        } else {
          pairSq(X, Xp, w);
        }
        lds[H + b] = X;
        lds[H + bp] = Xp;
      }
    }
  }
  __syncthreads();
  fftLds<true, T>(lds, H, 2, J.colPlan, J.nCol, twCol, tid);

  for (int e = tid; e < 2 * H; e += T) {
    int lb = e / per;
    int wi = e - lb * per;
    int h = wi / R;
    int l = lb * R + wi - h * R;
    gst(io + (size_t)lb * W * R + colBase + wi, cmulc(lds[h * H + l], midTw(J, h ? a1 : a0, l)));
  }
}

// ---------------------------------------------------------------- carry propagation
// Word order of the (line block, column) carries: rank = c * nb + lb, the block of lines
// [lb*R, lb*R + R) in column c holds words 2*(l + H*c) + rep, so consecutive ranks are
// consecutive runs of 2R words and rank nb*W - 1 wraps around to rank 0.

// Adds one carry into the 2R words of the block at `rank` and onwards in word order until it
// dies (or has gone once around); the digits stay balanced.
__device__ void rippleCarry(const JobDesc& J, int2* words, uint32_t rank, int64_t carry) {
  const uint32_t W = J.W, R = J.R, nb = J.nb, total = nb * W;
  for (uint32_t n = 0; carry && n < total; ++n, rank = (rank + 1 == total) ? 0 : rank + 1) {
    const uint32_t lb = rank % nb, c = rank / nb;
    int2* blk = words + (size_t)lb * W * R + (size_t)c * R;
    const ColW cw = J.colW[c];
    for (uint32_t r = 0; r < R && carry; ++r) {
      const LineW lw = J.lineW[lb * R + r];
      int2 d = blk[r];
      const WordW w0 = wordWeight(J, lw, cw, 0);
      int64_t x = (int64_t)d.x + carry;
      int64_t nd = lowBits64(x, w0.bits);
      carry = (x - nd) >> w0.bits;
      d.x = (int32_t)nd;
      if (carry) {
        const WordW w1 = wordWeight(J, lw, cw, 1);
        x = (int64_t)d.y + carry;
        nd = lowBits64(x, w1.bits);
        carry = (x - nd) >> w1.bits;
        d.y = (int32_t)nd;
      }
      blk[r] = d;
    }
  }
}

// One thread per (line block, column): add the carry-out of the preceding block
// (in word order) and propagate through this block's 2R words. A carry still
// left after the block goes to carry[co], raises flags[0] and marks its rank in the
// job's fix map (bit rank of map[], bit rank/64 of the summary after it) for k_carry_fix.
__global__ __launch_bounds__(256) void k_carryb(const JobDesc* __restrict__ jobs, int wi, int ci, int co) {
  const JobDesc& J = jobs[blockIdx.y];
  const int W = J.W, R = J.R, nb = J.nb;
  int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= nb * W) { return; }
  int lb = idx / W, c = idx - lb * W;
  int plb = lb ? lb - 1 : nb - 1;
  int pc = lb ? c : (c ? c - 1 : W - 1);
  int64_t carry = J.carry[ci][(size_t)plb * W + pc];
  if (carry) {
    int2* blk = J.words[wi] + (size_t)lb * W * R + (size_t)c * R;
    const ColW cw = J.colW[c];
    for (int r = 0; r < R && carry; ++r) {
      const LineW lw = J.lineW[lb * R + r];
      int2 d = blk[r];
      WordW w0 = wordWeight(J, lw, cw, 0);
      int64_t t = (int64_t)d.x + carry;
      int64_t nd = lowBits64(t, w0.bits);
      carry = (t - nd) >> w0.bits;
      d.x = (int32_t)nd;
      if (carry) {
        WordW w1 = wordWeight(J, lw, cw, 1);
        t = (int64_t)d.y + carry;
        nd = lowBits64(t, w1.bits);
        carry = (t - nd) >> w1.bits;
        d.y = (int32_t)nd;
      }
      blk[r] = d;
    }
  }
  J.carry[co][(size_t)lb * W + c] = carry;
  if (carry) {
    const uint32_t rank = (uint32_t)c * nb + lb;
    const uint32_t nMap = (nb * W + 63) / 64;
    atomicOr((unsigned long long*)J.fixMap + rank / 64, 1ull << (rank % 64));
    atomicOr((unsigned long long*)J.fixMap + nMap + rank / 4096, 1ull << ((rank / 64) % 64));
    atomicOr(J.flags, 1u);
  }
}

// Completes, in word order, the rare carries that k_carryb could not absorb inside their
// block (flags[0] set); exits at once when there are none. The group reads the fix map's
// summary 256 words at a time (clearing it), and thread 0 visits only the marked ranks,
// in ascending order, rippling each carry onwards; map and flag are left cleared.
__global__ __launch_bounds__(256) void k_carry_fix(const JobDesc* __restrict__ jobs, int wi, int ci) {
  const JobDesc& J = jobs[blockIdx.y];
  __shared__ unsigned long long sum[256];
  if (__hip_atomic_load(J.flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) { return; }
  const uint32_t nb = J.nb, W = J.W, total = nb * W;
  const uint32_t nMap = (total + 63) / 64, nSum = (nMap + 63) / 64;
  unsigned long long* map = (unsigned long long*)J.fixMap;
  for (uint32_t s0 = 0; s0 < nSum; s0 += 256) {
    const uint32_t i = s0 + threadIdx.x;
    sum[threadIdx.x] = 0;
    if (i < nSum) {
      sum[threadIdx.x] = map[nMap + i];
      if (sum[threadIdx.x]) { map[nMap + i] = 0; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t k = 0; k < 256 && s0 + k < nSum; ++k) {
        for (unsigned long long bits = sum[k]; bits; bits &= bits - 1) {
          const uint32_t mw = (s0 + k) * 64 + __builtin_ctzll(bits);
          unsigned long long m = map[mw];
          map[mw] = 0;
          for (; m; m &= m - 1) {
            const uint32_t rank = mw * 64 + __builtin_ctzll(m);
            const uint32_t lb = rank % nb, c = rank / nb;
            int64_t* slot = J.carry[ci] + (size_t)lb * W + c;
            const int64_t carry = *slot;
            *slot = 0;
            rippleCarry(J, J.words[wi], rank + 1 == total ? 0 : rank + 1, carry);
          }
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { J.flags[0] = 0; }
}

// ---------------------------------------------------------------- state I/O
// Sequential words (word k at seq[k]) <-> line-block layout of int2 word pairs.
__global__ __launch_bounds__(256) void k_words_in(const JobDesc* __restrict__ jobs, int wi, const int32_t* __restrict__ seq) {
  const JobDesc& J = jobs[blockIdx.y];
  uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= J.ND) { return; }
  uint32_t l = j % J.H, c = j / J.H;
  J.words[wi][(size_t)(l / J.R) * J.W * J.R + (size_t)c * J.R + l % J.R] = make_int2(seq[2 * j], seq[2 * j + 1]);
}

__global__ __launch_bounds__(256) void k_words_out(const JobDesc* __restrict__ jobs, int wi, int32_t* __restrict__ seq) {
  const JobDesc& J = jobs[blockIdx.y];
  uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= J.ND) { return; }
  uint32_t l = j % J.H, c = j / J.H;
  int2 v = J.words[wi][(size_t)(l / J.R) * J.W * J.R + (size_t)c * J.R + l % J.R];
  seq[2 * j] = v.x;
  seq[2 * j + 1] = v.y;
}

// small[0..31] = the top 32 words (k = N-32 .. N-1), small[32..63] = the bottom 32 words.
__global__ __launch_bounds__(64) void k_residue(const JobDesc* __restrict__ jobs, int wi) {
  const JobDesc& J = jobs[blockIdx.y];
  int t = threadIdx.x;
  if (t >= 32) { return; }
  uint32_t j = (t < 16) ? J.ND - 16 + t : t - 16;
  uint32_t l = j % J.H, c = j / J.H;
  int2 v = J.words[wi][(size_t)(l / J.R) * J.W * J.R + (size_t)c * J.R + l % J.R];
  J.small[2 * t] = v.x;
  J.small[2 * t + 1] = v.y;
}

__global__ __launch_bounds__(256) void k_is_equal(const JobDesc* __restrict__ jobs, int wa, int wb) {
  const JobDesc& J = jobs[blockIdx.y];
  const int64_t* a = (const int64_t*)J.words[wa];
  const int64_t* b = (const int64_t*)J.words[wb];
  bool diff = false;
  for (uint32_t p = blockIdx.x * 256 + threadIdx.x; p < J.ND; p += gridDim.x * 256) { diff |= a[p] != b[p]; }
  if (diff) { atomicOr(J.flags + 1, 1u); }
}

// 64-bit sum of a words buffer, for verified reads (src/Gpu.cpp:677-705).
__global__ __launch_bounds__(256) void k_sum64(const JobDesc* __restrict__ jobs, int wi) {
  const JobDesc& J = jobs[blockIdx.y];
  const uint64_t* a = (const uint64_t*)J.words[wi];
  uint64_t s = 0;
  for (uint32_t p = blockIdx.x * 256 + threadIdx.x; p < J.ND; p += gridDim.x * 256) { s += a[p]; }
  for (int off = 32; off > 0; off >>= 1) { s += __shfl_xor(s, off, 64); }
  if ((threadIdx.x & 63) == 0) { atomicAdd((unsigned long long*)J.sum, (unsigned long long)s); }
}

// Fill n u32 with v (used instead of memset so that it can sit inside a captured graph).
__global__ __launch_bounds__(256) void k_fill32(uint32_t* p, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) { p[i] = v; }
}

__global__ __launch_bounds__(256) void k_copy_words(const JobDesc* __restrict__ jobs, int dst, int src) {
  const JobDesc& J = jobs[blockIdx.y];
  for (uint32_t p = blockIdx.x * 256 + threadIdx.x; p < J.ND; p += gridDim.x * 256) { J.words[dst][p] = J.words[src][p]; }
}

// ---------------------------------------------------------------- kernel table
// Looked up by name through owl_hip_kernel_get (the replacement of the reference's
// named kernel loading, src/clwrap.cpp:278-284).
struct OwlKernelEntry {
  const char* name;
  const void* fn;
  unsigned maxThreads;
};

extern "C" const OwlKernelEntry owl_kernel_table[] = {
    {"row64", (const void*)&k_row<64>, 64},
    {"row128", (const void*)&k_row<128>, 128},
    {"row256", (const void*)&k_row<256>, 256},
    {"row512", (const void*)&k_row<512>, 512},
    {"col64", (const void*)&k_col<64>, 64},
    {"col128", (const void*)&k_col<128>, 128},
    {"col256", (const void*)&k_col<256>, 256},
    {"col512", (const void*)&k_col<512>, 512},
    {"carryB", (const void*)&k_carryb, 256},
    {"carryFix", (const void*)&k_carry_fix, 256},
    {"wordsIn", (const void*)&k_words_in, 256},
    {"wordsOut", (const void*)&k_words_out, 256},
    {"readResidue", (const void*)&k_residue, 64},
    {"isEqual", (const void*)&k_is_equal, 256},
    {"sum64", (const void*)&k_sum64, 256},
    {"fill32", (const void*)&k_fill32, 256},
    {"copyWords", (const void*)&k_copy_words, 256},
    {nullptr, nullptr, 0},
};

// ---------------------------------------------------------------- staging without allocation
// The digit readout (k_words_out) stages N int32 in device memory. owl_gpu_read_digits
// allocates that staging buffer per call; owl_gpu_read_digits_noalloc stages through this
// file-scope array instead, for callers that must not allocate. Each device holds its own
// copy of the array until the process ends. In host code its name is not a device address:
// owlNoAllocScratch() resolves the address once per device (with that device current) and
// kernels only ever receive that pointer. Calls on one device share the array, so
// overlapping calls on one device race; every call overwrites the N words it reads back.
static __device__ __align__(256) int32_t scratch[OWL_NOALLOC_MAX_WORDS];

#include <mutex>

extern "C" int owlNoAllocScratch(int device, int32_t** ptr) {
  static std::mutex mu;
  static int32_t* addr[OWL_MAX_DEVICES] = {};
  if (device < 0 || device >= OWL_MAX_DEVICES) { return 1; }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) { return 2; }
    if (cur != device && hipSetDevice(device) != hipSuccess) { return 2; }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) { (void)hipSetDevice(cur); }
    if (e != hipSuccess || !p) { return 3; }
    addr[device] = static_cast<int32_t*>(p);
  }
  *ptr = addr[device];
  return 0;
}
