// Per-exponent job descriptor shared by the host engine and the HIP kernels.
//
// The reference bakes every exponent constant into the OpenCL source as -D
// defines and JIT-compiles per exponent (src/Gpu.cpp:187-291). Here every such
// constant is a runtime field of JobDesc, so one offline-compiled gfx950 code
// object serves any exponent, and a batch of exponents is an array of JobDesc.
//
// Geometry. N words = 2*ND, ND = W*H complex FP64 values. Complex index
// j = l + H*c (line l in [0,H), column c in [0,W)); word k = 2*j + rep.
// This is the reference's word order (kAt, src/Gpu.cpp:40). W is the
// reference WIDTH, H is the reference BIG_HEIGHT = MIDDLE*SMALL_HEIGHT.
//
// HBM layout ("line blocks"). Lines are grouped into blocks of R lines; element
// (l, s) lives at  (l/R)*W*R + s*R + l%R. In the frequency domain s is a
// column "slot": slot 2q holds frequency q and slot 2q+1 frequency W-q
// (q in [1, W/2)); slot 0 holds frequency 0 and slot 1 frequency W/2. Thus the
// two columns that the real-signal pairing couples sit side by side, and a
// column-pass group reads 2*R*16 contiguous bytes per line block. In the
// integer (time) domain s is the column c itself.
#pragma once
#include <stdint.h>

#define OWL_MAX_STAGES 8
#define OWL_ROE_SIZE 100000
// Largest N (words) whose digits the no-allocation readout stages in its static array
// (2^23 int32 = 32 MiB per device; covers FFTs up to 8M words).
#define OWL_NOALLOC_MAX_WORDS (1u << 23)
#define OWL_MAX_DEVICES 64

// Per-line IBDWT weight data: words 2*(l + H*c) + rep for rep in {0,1}.
// e[rep] = step*(2l+rep) mod N; w = 2^(e/N); iw = 2^(-e/N).
struct LineW {
  double w[2];
  double iw[2];
  uint32_t e[2];
};

// Per-column weight data: e = step*2H*c mod N; w = 2^(e/N);
// iw = 2^(-e/N) / ND (the inverse-FFT scale is folded in here).
struct ColW {
  double w;
  double iw;
  uint32_t e;
  uint32_t pad;
};

struct JobDesc {
  uint32_t E, N, ND, W, H, R, nb;   // nb = H / R line blocks
  uint32_t bitsLo;                  // E / N
  uint32_t rE;                      // E % N: word k is "big" iff extra(k) < rE
  uint32_t nHi;                     // ceil(H / 64): rows of the split mid twiddle table
  uint32_t lgR;                     // log2(R)
  uint32_t nRow, nCol;              // number of FFT stages
  uint32_t rowPlan[OWL_MAX_STAGES]; // radices of the W-point row FFT
  uint32_t colPlan[OWL_MAX_STAGES]; // radices of the H-point column FFT

  const double2* twRow;   // W entries: exp(-2 pi i m / W)
  const double2* twCol;   // H entries: exp(-2 pi i m / H)
  const double2* twMidA;  // W*64: exp(-2 pi i a*lo / ND), lo < 64
  const double2* twMidB;  // W*nHi: exp(-2 pi i a*64*hi / ND)
  const double2* twPair;  // W entries: exp(-2 pi i a / ND)
  const double2* twMidS;  // W*16: exp(-2 pi i a*k*(H/16) / ND), k < 16 (shape-specialized column pass)
  const LineW* lineW;     // H entries
  const ColW* colW;       // W entries

  double2* freq[3];       // [0] state, [1] scratch, [2] second-operand spectrum
  int2* words[4];         // [0] data, [1] aux, [2] check, [3] base: integer words (line-block layout)
  int64_t* carry[2];      // carry shuttle (nb*W) and carry-propagation ping-pong
  uint64_t* stair;        // nb*W tagged carries of the fused row pass (tag << 48 | 48-bit carry)
  uint64_t* ticket;       // fused row-pass start counter: every launch takes nb+1 tickets
  uint32_t* roe;          // OWL_ROE_SIZE max-round-off slots (float bits)
  uint32_t* flags;        // [0] carry leftover, [1] compare mismatch
  uint64_t* rowMeta;      // shape-specialized row pass: nb*T words (T = W/F0 threads), word (block, tid):
                          // bit w = big word, bit 32+w = weight exponent wrapped, for the thread's
                          // words w = (k*R + r)*2 + rep (column tid + k*T, line block*R + r)
  uint64_t* fixMap;       // ceil(nb*W/64) words: bit rank set where a carry left its block, then
                          // ceil(nb*W/4096) summary words: bit w set where map word w is nonzero
  int32_t* small;         // 64 ints for the residue read
  uint64_t* sum;          // checksum output
};

// Row-pass modes.
enum { ROW_FUSED = 0, ROW_LEADIN = 1, ROW_CARRYOUT = 2 };
// Column-pass modes.
enum { COL_SQUARE = 0, COL_MUL = 1, COL_FWD = 2 };
