#include "Trig.h"

#include <cmath>

std::pair<double, double> root1(u32 n, u64 k) {
  k %= n;
  // 8k = o*n + rem: angle = (pi/4)(o + rem/n), x = (pi/4) rem/n, y = (pi/4) - x.
  u64 K = 8 * k;
  u32 o = u32(K / n);
  u64 rem = K % n;
  const long double q = 0.785398163397448309615660845819875721L;  // pi/4
  long double x = q * (long double)rem / (long double)n;
  long double y = q * (long double)(n - rem) / (long double)n;
  long double cx = cosl(x), sx = sinl(x), cy = cosl(y), sy = sinl(y);
  long double c = 0, s = 0;
  switch (o) {
    case 0: c = cx; s = sx; break;
    case 1: c = sy; s = cy; break;
    case 2: c = -sx; s = cx; break;
    case 3: c = -cy; s = sy; break;
    case 4: c = -cx; s = -sx; break;
    case 5: c = -sy; s = -cy; break;
    case 6: c = sx; s = -cx; break;
    default: c = cy; s = -sy; break;
  }
  // exact values at multiples of pi/2
  if (rem == 0 && o % 2 == 0) {
    const int cc[4] = {1, 0, -1, 0}, ss[4] = {0, 1, 0, -1};
    c = cc[o / 2];
    s = ss[o / 2];
  }
  return {double(c), double(-s)};
}

std::vector<u32> fftPlan(u32 L) {
  std::vector<u32> plan;
  u32 odd = L;
  while (odd % 2 == 0) { odd /= 2; }
  u32 pow2 = L / odd;
  while (pow2 >= 16) {
    plan.push_back(16);
    pow2 /= 16;
  }
  for (u32 f : {15u, 9u, 3u, 5u, 7u, 11u, 13u}) {
    while (odd % f == 0) {
      plan.push_back(f);
      odd /= f;
    }
  }
  if (odd != 1) { return {}; }
  if (pow2 > 1) {
    // fold a small power of two into an odd radix when it stays <= 16
    bool merged = false;
    for (u32& r : plan) {
      if (r != 16 && r * pow2 <= 16 && r % 2 == 1) {
        r *= pow2;
        merged = true;
        break;
      }
    }
    if (!merged) { plan.push_back(pow2); }
  }
  return plan;
}

bool planFits(const std::vector<u32>& plan, u32 L, u32 nfft, u32 T, u32 ept) {
  if (plan.empty()) { return false; }
  u64 prod = 1;
  for (u32 r : plan) {
    if (r < 2 || r > 16 || r > ept) { return false; }
    u64 butterflies = u64(L / r) * nfft;
    u64 perThread = (butterflies + T - 1) / T;
    if (perThread * r > ept || perThread > ept / r) { return false; }
    prod *= r;
  }
  return prod == L;
}
