// Twiddle and weight generation in extended precision (x87 long double).
// Plays the role of the reference src/Trig.cpp and src/TrigBufCache.cpp:22-120
// (root1 / roundTrig): every table entry is computed from an exactly reduced
// angle and rounded once to double.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

using u32 = uint32_t;
using u64 = uint64_t;

// exp(-2 pi i k / n) = (cos(2 pi k/n), -sin(2 pi k/n)), k reduced exactly to one octant.
std::pair<double, double> root1(u32 n, u64 k);

// Mixed-radix plan for an L-point FFT: radices in {2..16}, largest first.
// Empty if L has a prime factor above 13.
std::vector<u32> fftPlan(u32 L);

// True if every stage of plan keeps at most ept elements per thread when
// nfft L-point FFTs are done by T threads.
bool planFits(const std::vector<u32>& plan, u32 L, u32 nfft, u32 T, u32 ept);
