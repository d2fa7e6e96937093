#include "TrigBufCache.h"

#include "Trig.h"

#include <map>
#include <mutex>
#include <vector>

namespace {
double2 d2(std::pair<double, double> p) { return double2{p.first, p.second}; }
}  // namespace

std::shared_ptr<const TrigTables> TrigBufCache::get(u32 W, u32 H, owl_queue q) {
  static std::mutex mtx;
  static std::map<std::pair<u32, u32>, std::weak_ptr<const TrigTables>> cache;
  std::lock_guard<std::mutex> lock(mtx);
  auto key = std::make_pair(W, H);
  if (auto it = cache.find(key); it != cache.end()) {
    if (auto sp = it->second.lock()) { return sp; }
  }
  const u32 ND = W * H;
  const u32 nHi = (H + 63) / 64;
  std::vector<double2> row(W), col(H), midA(size_t(W) * 64), midB(size_t(W) * nHi), pair(W), midS(size_t(W) * 16);
  for (u32 m = 0; m < W; ++m) { row[m] = d2(root1(W, m)); }
  for (u32 m = 0; m < H; ++m) { col[m] = d2(root1(H, m)); }
  for (u32 a = 0; a < W; ++a) {
    pair[a] = d2(root1(ND, a));
    for (u32 lo = 0; lo < 64; ++lo) { midA[size_t(a) * 64 + lo] = d2(root1(ND, u64(a) * lo)); }
    for (u32 hi = 0; hi < nHi; ++hi) { midB[size_t(a) * nHi + hi] = d2(root1(ND, u64(a) * 64 * hi % ND)); }
    for (u32 k = 0; k < 16; ++k) { midS[size_t(a) * 16 + k] = d2(root1(ND, u64(a) * k * (H / 16) % ND)); }
  }
  auto t = std::make_shared<TrigTables>();
  t->row = Buffer<double2>(q, row);
  t->col = Buffer<double2>(q, col);
  t->midA = Buffer<double2>(q, midA);
  t->midB = Buffer<double2>(q, midB);
  t->pair = Buffer<double2>(q, pair);
  t->midS = Buffer<double2>(q, midS);
  cache[key] = t;
  return t;
}
