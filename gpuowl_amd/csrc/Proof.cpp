#include "Proof.h"

#include "Saver.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <memory>
#include <stdexcept>

// ---------------------------------------------------------------- SHA3-256
namespace {

// Keccak-f[1600] constants derived from their definitions in FIPS 202 (3.2.2, 3.2.5)
struct KeccakConsts {
  u64 rc[24];
  int rot[25];  // by lane index x + 5y
  KeccakConsts() {
    auto rcBit = [](int t) {
      if (t % 255 == 0) { return 1u; }
      unsigned r = 1;
      for (int i = 1; i <= t % 255; ++i) {
        r <<= 1;
        const unsigned b8 = (r >> 8) & 1;
        r ^= b8 | (b8 << 4) | (b8 << 5) | (b8 << 6);
        r &= 0xFF;
      }
      return r & 1;
    };
    for (int i = 0; i < 24; ++i) {
      rc[i] = 0;
      for (int j = 0; j <= 6; ++j) {
        if (rcBit(j + 7 * i)) { rc[i] |= u64(1) << ((1 << j) - 1); }
      }
    }
    rot[0] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; ++t) {
      rot[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
      const int nx = y, ny = (2 * x + 3 * y) % 5;
      x = nx, y = ny;
    }
  }
};
const KeccakConsts& kc() {
  static const KeccakConsts c;
  return c;
}

inline u64 rotl(u64 v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }

void keccakF(u64* a) {
  const KeccakConsts& k = kc();
  for (int round = 0; round < 24; ++round) {
    u64 c[5], b[25];
    for (int x = 0; x < 5; ++x) { c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20]; }
    for (int x = 0; x < 5; ++x) {
      const u64 d = c[(x + 4) % 5] ^ rotl(c[(x + 1) % 5], 1);
      for (int y = 0; y < 5; ++y) { a[x + 5 * y] ^= d; }
    }
    for (int x = 0; x < 5; ++x) {  // rho and pi: B[y][2x + 3y] = rot(A[x][y])
      for (int y = 0; y < 5; ++y) { b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(a[x + 5 * y], k.rot[x + 5 * y]); }
    }
    for (int y = 0; y < 5; ++y) {  // chi
      for (int x = 0; x < 5; ++x) { a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]); }
    }
    a[0] ^= k.rc[round];  // iota
  }
}

u64 loadLE(const unsigned char* p) {
  u64 v = 0;
  for (int i = 7; i >= 0; --i) { v = (v << 8) | p[i]; }
  return v;
}

}  // namespace

Sha3_256::Sha3_256() { std::memset(st_, 0, sizeof st_); }

void Sha3_256::absorbBlock(const unsigned char* p) {
  for (int i = 0; i < 17; ++i) { st_[i] ^= loadLE(p + 8 * i); }
  keccakF(st_);
}

Sha3_256& Sha3_256::update(const void* data, size_t n) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  while (n) {
    const size_t take = std::min(n, sizeof buf_ - fill_);
    std::memcpy(buf_ + fill_, p, take);
    fill_ += take, p += take, n -= take;
    if (fill_ == sizeof buf_) {
      absorbBlock(buf_);
      fill_ = 0;
    }
  }
  return *this;
}

std::array<u64, 4> Sha3_256::finish() {
  std::memset(buf_ + fill_, 0, sizeof buf_ - fill_);
  buf_[fill_] ^= 0x06;  // SHA-3 domain bits and the first pad bit
  buf_[sizeof buf_ - 1] ^= 0x80;
  absorbBlock(buf_);
  fill_ = 0;
  return {st_[0], st_[1], st_[2], st_[3]};
}

std::array<u64, 4> proofHash(u32 E, const Words& w) {
  if (w.size() * 4 < (E - 1) / 8 + 1) { throw std::invalid_argument("proof hash: words"); }
  return Sha3_256{}.update(w.data(), (E - 1) / 8 + 1).finish();
}

std::array<u64, 4> proofHash(u32 E, const std::array<u64, 4>& prefix, const Words& w) {
  if (w.size() * 4 < (E - 1) / 8 + 1) { throw std::invalid_argument("proof hash: words"); }
  return Sha3_256{}.update(prefix.data(), 32).update(w.data(), (E - 1) / 8 + 1).finish();
}

// ---------------------------------------------------------------- points
std::vector<u32> proofPoints(u32 E, u32 power) {
  if (!(E & 1) || power < 1 || power > 12) { throw std::invalid_argument("proof power / exponent"); }
  std::vector<u32> pts{0};
  for (u32 p = 0, span = (E + 1) / 2; p < power; ++p, span = (span + 1) / 2) {
    const size_t n = pts.size();
    for (size_t i = 0; i < n; ++i) { pts.push_back(pts[i] + span); }
  }
  pts.front() = E;
  std::sort(pts.begin(), pts.end());
  return pts;
}

// ---------------------------------------------------------------- file
namespace {
struct FileCloser {
  void operator()(FILE* f) const { fclose(f); }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;
constexpr const char* kHeader = "PRP PROOF\nVERSION=2\nHASHSIZE=64\nPOWER=%u\nNUMBER=M%u\n";
}  // namespace

void PrpProof::save(const std::string& path) const {
  const size_t nBytes = (E - 1) / 8 + 1;
  FilePtr f{fopen(path.c_str(), "wb")};
  if (!f) { throw std::invalid_argument("cannot write proof " + path); }
  bool ok = fprintf(f.get(), kHeader, unsigned(middles.size()), E) > 0;
  auto put = [&](const Words& w) {
    if (w.size() * 4 < nBytes) { throw std::invalid_argument("proof residue too short"); }
    ok = ok && fwrite(w.data(), 1, nBytes, f.get()) == nBytes;
  };
  put(B);
  for (const Words& m : middles) { put(m); }
  if (!ok || fflush(f.get())) { throw std::invalid_argument("proof write failed: " + path); }
}

PrpProof PrpProof::load(const std::string& path) {
  FilePtr f{fopen(path.c_str(), "rb")};
  if (!f) { throw std::invalid_argument("cannot read proof " + path); }
  unsigned power = 0, E = 0;
  char nl = 0;
  if (fscanf(f.get(), "PRP PROOF\nVERSION=2\nHASHSIZE=64\nPOWER=%u\nNUMBER=M%u%c", &power, &E, &nl) != 3 || nl != '\n' ||
      power < 1 || power > 12 || E < 2) {
    throw std::invalid_argument("invalid proof header in " + path);
  }
  PrpProof p;
  p.E = E;
  const size_t nBytes = (E - 1) / 8 + 1;
  auto get = [&]() {
    Words w(nWords(E), 0);
    if (fread(w.data(), 1, nBytes, f.get()) != nBytes) { throw std::invalid_argument("short proof file " + path); }
    return w;
  };
  p.B = get();
  for (unsigned i = 0; i < power; ++i) { p.middles.push_back(get()); }
  return p;
}

// ---------------------------------------------------------------- MD5
// RFC 1321. The additive constants are floor(|sin(i + 1)| * 2^32) (section 3.4), computed once
// in long double; the per-round shift amounts are the RFC's four rows of four.
namespace {
struct Md5Consts {
  u32 T[64];
  Md5Consts() {
    for (int i = 0; i < 64; ++i) { T[i] = u32(std::floor(std::fabs(std::sin((long double)(i + 1))) * 4294967296.0L)); }
  }
};
inline u32 rotl32(u32 v, int n) { return (v << n) | (v >> (32 - n)); }

void md5Block(u32* h, const unsigned char* p) {
  static const Md5Consts c;
  static const int S[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
  u32 X[16];
  for (int i = 0; i < 16; ++i) { X[i] = u32(p[4 * i]) | u32(p[4 * i + 1]) << 8 | u32(p[4 * i + 2]) << 16 | u32(p[4 * i + 3]) << 24; }
  u32 a = h[0], b = h[1], cc = h[2], d = h[3];
  for (int i = 0; i < 64; ++i) {
    const int round = i / 16;
    u32 f;
    int k;
    switch (round) {
      case 0: f = (b & cc) | (~b & d), k = i; break;
      case 1: f = (b & d) | (cc & ~d), k = (5 * i + 1) % 16; break;
      case 2: f = b ^ cc ^ d, k = (3 * i + 5) % 16; break;
      default: f = cc ^ (b | ~d), k = (7 * i) % 16; break;
    }
    const u32 t = d;
    d = cc;
    cc = b;
    b = b + rotl32(a + f + c.T[i] + X[k], S[round][i % 4]);
    a = t;
  }
  h[0] += a, h[1] += b, h[2] += cc, h[3] += d;
}

struct Md5 {
  u32 h[4] = {0x67452301, 0xefcdab89, 0x98badcfe, 0x10325476};
  unsigned char buf[64];
  size_t fill = 0;
  u64 total = 0;
  void update(const unsigned char* p, size_t n) {
    total += n;
    while (n) {
      const size_t take = std::min(n, 64 - fill);
      std::memcpy(buf + fill, p, take);
      fill += take, p += take, n -= take;
      if (fill == 64) {
        md5Block(h, buf);
        fill = 0;
      }
    }
  }
  std::string hex() {
    const u64 bits = total * 8;
    const unsigned char one = 0x80, zero = 0;
    update(&one, 1);
    while (fill != 56) { update(&zero, 1); }
    unsigned char len[8];
    for (int i = 0; i < 8; ++i) { len[i] = (unsigned char)(bits >> (8 * i)); }
    update(len, 8);
    static const char* digits = "0123456789abcdef";
    std::string s;
    for (u32 w : h) {
      for (int i = 0; i < 4; ++i) {
        const unsigned byte = (w >> (8 * i)) & 0xFF;
        s += digits[byte >> 4];
        s += digits[byte & 15];
      }
    }
    return s;
  }
};
}  // namespace

std::string md5Hex(const void* data, size_t n) {
  Md5 m;
  m.update(static_cast<const unsigned char*>(data), n);
  return m.hex();
}

std::string md5File(const std::string& path) {
  FilePtr f{fopen(path.c_str(), "rb")};
  if (!f) { throw std::invalid_argument("cannot read " + path); }
  Md5 m;
  unsigned char buf[65536];
  size_t got = 0;
  while ((got = fread(buf, 1, sizeof buf, f.get())) > 0) { m.update(buf, got); }
  return m.hex();
}

// ---------------------------------------------------------------- residue store
ProofStore::ProofStore(const std::string& dir, u32 E, u32 power)
    : dir_{dir}, E_{E}, power_{power}, points_{proofPoints(E, power)} {
  std::filesystem::create_directories(dir_);
}

u32 ProofStore::next(u32 k) const {
  auto it = std::upper_bound(points_.begin(), points_.end(), k);
  return it == points_.end() ? ~0u : *it;
}

void ProofStore::save(u32 k, const Words& w) const {
  if (!std::binary_search(points_.begin(), points_.end(), k)) { throw std::invalid_argument("not a proof point"); }
  Words full = w;
  full.resize(nWords(E_), 0);
  const std::string path = dir_ + "/" + std::to_string(k);
  FilePtr f{fopen(path.c_str(), "wb")};
  const u32 crc = crc32(full.data(), full.size() * 4);
  if (!f || fwrite(&crc, 4, 1, f.get()) != 1 || fwrite(full.data(), 4, full.size(), f.get()) != full.size() ||
      fflush(f.get())) {
    throw std::invalid_argument("cannot write proof residue " + path);
  }
}

Words ProofStore::load(u32 k) const {
  const std::string path = dir_ + "/" + std::to_string(k);
  FilePtr f{fopen(path.c_str(), "rb")};
  u32 crc = 0;
  Words w(nWords(E_));
  if (!f || fread(&crc, 4, 1, f.get()) != 1 || fread(w.data(), 4, w.size(), f.get()) != w.size() ||
      crc32(w.data(), w.size() * 4) != crc) {
    throw std::invalid_argument("missing or damaged proof residue " + path);
  }
  return w;
}
