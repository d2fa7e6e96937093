#include "state.h"

#include <stdexcept>

Words compactBits(const std::vector<int>& digits, u32 E) {
  if (digits.empty()) { return {}; }
  const u32 N = u32(digits.size());
  Words out;
  out.reserve(nWords(E));
  u64 acc = 0;   // pending bits
  u32 have = 0;  // number of pending bits
  int borrow = 0;
  for (u32 k = 0; k < N; ++k) {
    u32 b = bitlen(N, E, k);
    i64 v = i64(digits[k]) + borrow;
    borrow = 0;
    if (v < 0) {
      v += i64(1) << b;
      borrow = -1;
    }
    if (v < 0 || v >= (i64(1) << b)) { throw std::runtime_error("compactBits: digit out of range"); }
    acc |= u64(v) << have;
    have += b;
    while (have >= 32) {
      out.push_back(u32(acc));
      acc >>= 32;
      have -= 32;
    }
  }
  if (have) { out.push_back(u32(acc)); }
  // A negative value V is stored as V + 2^E; subtracting 1 gives V + (2^E - 1).
  for (size_t p = 0; borrow && p < out.size(); ++p) {
    i64 v = i64(out[p]) + borrow;
    out[p] = u32(v);
    borrow = v < 0 ? -1 : 0;
  }
  if (out.size() != nWords(E)) { throw std::runtime_error("compactBits: size"); }
  return out;
}

std::vector<int> expandBits(const Words& words, u32 N, u32 E) {
  if (E % 32 == 0) { throw std::invalid_argument("expandBits: E multiple of 32"); }
  std::vector<int> out(N);
  u64 acc = 0;
  u32 have = 0;
  size_t next = 0;
  for (u32 k = 0; k < N; ++k) {
    u32 b = bitlen(N, E, k);
    if (have < b) {
      if (next >= words.size()) { throw std::invalid_argument("expandBits: too few words"); }
      acc += u64(words[next++]) << have;
      have += 32;
    }
    // take the low b bits as a signed digit; a negative digit borrows 1 from what remains
    u64 low = acc & ((u64(1) << b) - 1);
    int d = int(i64(low << (64 - b)) >> (64 - b));
    acc >>= b;
    have -= b;
    if (d < 0) { acc += 1; }
    out[k] = d;
  }
  if (next != words.size()) { throw std::invalid_argument("expandBits: too many words"); }
  // What is left above bit E is a carry of 0 or 1 that wraps to word 0 (2^E == 1).
  out[0] += int(acc);
  return out;
}

u64 residueFromEnds(const int* w, u32 N, u32 E) {
  int carry = 0;
  for (int i = 0; i < 32; ++i) { carry = (w[i] + carry < 0) ? -1 : 0; }
  u64 res = 0;
  u32 hasBits = 0;
  for (u32 k = 0; k < 32 && hasBits < 64; ++k) {
    u32 len = bitlen(N, E, k);
    i64 v = i64(w[32 + k]) + carry;
    carry = v < 0 ? -1 : 0;
    if (v < 0) { v += i64(1) << len; }
    res |= u64(v) << hasBits;
    hasBits += len;
  }
  return res;
}
