// Word layout of a residue mod 2^E - 1 in N balanced digits.
// Same definitions as the reference src/state.h:14-17 and src/state.cpp:26-115:
// word k has bitlen(k) = E/N + isBigWord(k) bits, isBigWord(k) <=> extra(k) + step < N,
// extra(k) = step*k mod N, step = N - E mod N; digits are balanced
// (-2^(b-1) <= d < 2^(b-1)); Words are the canonical residue as (E-1)/32+1 u32, low first.
#pragma once

#include <cstdint>
#include <vector>

using u32 = uint32_t;
using u64 = uint64_t;
using i64 = int64_t;
using Words = std::vector<u32>;

constexpr u32 stepOf(u32 N, u32 E) { return N - (E % N); }
constexpr u32 extraOf(u32 N, u32 E, u32 k) { return u32(u64(stepOf(N, E)) * k % N); }
constexpr bool isBigWord(u32 N, u32 E, u32 k) { return extraOf(N, E, k) + stepOf(N, E) < N; }
constexpr u32 bitlen(u32 N, u32 E, u32 k) { return E / N + isBigWord(N, E, k); }
inline u32 nWords(u32 E) { return (E - 1) / 32 + 1; }
inline u64 res64(const Words& w) { return w.empty() ? 0 : (u64(w.size() > 1 ? w[1] : 0) << 32) | w[0]; }

// Balanced digits (sequential word order) -> canonical packed words. Empty input = zero.
Words compactBits(const std::vector<int>& digits, u32 E);
// Packed words -> balanced digits in sequential word order.
std::vector<int> expandBits(const Words& words, u32 N, u32 E);
// res64 from the 32 top words (k = N-32..N-1) and 32 bottom words of a normalized
// residue, borrowing from the top when it is negative (src/Gpu.cpp:955-975).
u64 residueFromEnds(const int* top32bottom32, u32 N, u32 E);
