#include "Worktodo.h"

#include <cctype>
#include <cstdio>
#include <stdexcept>
#include <ctime>
#include <sys/utsname.h>

namespace {

std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  size_t start = 0;
  while (true) {
    size_t p = s.find(sep, start);
    out.push_back(s.substr(start, p == std::string::npos ? std::string::npos : p - start));
    if (p == std::string::npos) { return out; }
    start = p + 1;
  }
}

std::string trimEnd(std::string s) {
  while (!s.empty() && (s.back() == '\n' || s.back() == '\r' || s.back() == ' ' || s.back() == '\t')) { s.pop_back(); }
  return s;
}

bool isAid(const std::string& s) {
  if (s.size() != 32) { return false; }
  for (char c : s) {
    if (!std::isxdigit(static_cast<unsigned char>(c))) { return false; }
  }
  return true;
}

// a decimal number made of digits only (0 when not)
unsigned long long number(const std::string& s) {
  if (s.empty() || s.size() > 19) { return 0; }
  unsigned long long v = 0;
  for (char c : s) {
    if (c < '0' || c > '9') { return 0; }
    v = v * 10 + unsigned(c - '0');
  }
  return v;
}

}  // namespace

WorkTask parseWorktodoLine(const std::string& raw) {
  const std::string line = trimEnd(raw);
  WorkTask t;
  if (line.empty() || line[0] == '#') { return t; }
  const auto eq = line.find('=');
  if (eq == std::string::npos || line.find('=', eq + 1) != std::string::npos) { return t; }
  const std::string kind = line.substr(0, eq);
  std::vector<std::string> f = split(line.substr(eq + 1), ',');
  if (kind == "PRP" || kind == "PRPDC" || kind == "Test" || kind == "DoubleCheck") {
    if (!f.empty() && (f[0].empty() || f[0] == "N/A")) { f.erase(f.begin()); }
    std::string aid;
    if (!f.empty() && isAid(f[0])) {
      aid = f[0];
      f.erase(f.begin());
    }
    // "k,b,n,c" with k = 1, b = 2, c = -1 names 2^n - 1; otherwise the exponent comes first
    const std::string e = (f.size() >= 4 && f[0] == "1" && f[1] == "2" && f[3] == "-1") ? f[2] : (f.empty() ? "" : f[0]);
    const unsigned long long E = number(e);
    if (E > 1000 && E < (1ull << 32)) {
      t.kind = (kind == "PRP" || kind == "PRPDC") ? WorkTask::PRP : WorkTask::LL;
      t.exponent = u32(E);
      t.aid = aid;
    }
  } else if (kind == "Cert") {
    if (f.size() == 6 && isAid(f[0]) && f[1] == "1" && f[2] == "2" && f[4] == "-1") {
      const unsigned long long E = number(f[3]), S = number(f[5]);
      if (E > 1000 && E < (1ull << 32) && S > 100 && S < (1ull << 32)) {
        t.kind = WorkTask::CERT;
        t.exponent = u32(E);
        t.aid = f[0];
        t.squarings = u32(S);
      }
    }
  }
  return t;
}

int bestWorktodoLine(const std::vector<std::string>& lines) {
  int best = -1;
  WorkTask b;
  for (size_t i = 0; i < lines.size(); ++i) {
    WorkTask t = parseWorktodoLine(lines[i]);
    if (t.kind == WorkTask::NONE) { continue; }
    const bool tCert = t.kind == WorkTask::CERT, bCert = b.kind == WorkTask::CERT;
    if (best < 0 || (tCert && !bCert) || (tCert == bCert && t.exponent < b.exponent)) {
      best = int(i);
      b = t;
    }
  }
  return best;
}

Words div9(const Words& x, u32 E) {
  const u32 n = nWords(E);
  if (E < 33 || x.size() > n) { throw std::invalid_argument("div9: operand"); }
  // M = 2^E - 1. Find t < 9 with x + t*M = 0 (mod 9): 2^32 = 4 (mod 9).
  u32 xm = 0;
  for (size_t i = x.size(); i-- > 0;) { xm = u32((u64(xm) * 4 + x[i] % 9) % 9); }
  u32 p2 = 1;
  for (u32 i = 0; i < E; ++i) { p2 = p2 * 2 % 9; }
  const u32 mm = (p2 + 8) % 9;
  u32 t = 0;
  while (t < 9 && (xm + t * mm) % 9) { ++t; }
  if (t == 9) { throw std::invalid_argument("div9: 9 has no inverse modulo 2^E - 1"); }
  // y = x + t*2^E - t as n + 1 words, then y / 9 from the top word down
  std::vector<u64> y(n + 1, 0);
  for (size_t i = 0; i < x.size(); ++i) { y[i] = x[i]; }
  y[E / 32] += u64(t) << (E % 32);
  for (u32 i = E / 32; i < n && (y[i] >> 32); ++i) {
    y[i + 1] += y[i] >> 32;
    y[i] &= 0xFFFFFFFFu;
  }
  u64 borrow = t;
  for (u32 i = 0; i <= n && borrow; ++i) {
    if (y[i] >= borrow) {
      y[i] -= borrow;
      borrow = 0;
    } else {
      y[i] = y[i] + (u64(1) << 32) - borrow;
      borrow = 1;
    }
  }
  Words q(n, 0);
  u64 r = 0;
  for (u32 i = n + 1; i-- > 0;) {
    const u64 cur = (r << 32) | y[i];
    if (i < n) {
      q[i] = u32(cur / 9);
    } else if (cur / 9) {
      throw std::logic_error("div9: quotient overflow");
    }
    r = cur % 9;
  }
  return q;
}

std::string res2048Hex(const Words& w) {
  std::string s;
  char buf[16];
  for (int i = 63; i >= 0; --i) {
    snprintf(buf, sizeof buf, "%08x", size_t(i) < w.size() ? w[i] : 0u);
    s += buf;
  }
  return s;
}

namespace {
// a JSON string (quotes and backslashes escaped, which the reference leaves as they are)
std::string str(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') { o += '\\'; }
    o += c;
  }
  return o + "\"";
}
std::string field(const std::string& key, const std::string& jsonValue) { return str(key) + ":" + jsonValue; }
std::string object(const std::vector<std::string>& fields) {
  std::string s;
  for (const std::string& f : fields) {
    if (f.empty()) { continue; }
    s += (s.empty() ? "{" : ", ") + f;
  }
  return s.empty() ? "" : s + "}";
}
std::string maybe(const std::string& key, const std::string& value) { return value.empty() ? "" : field(key, str(value)); }
std::string hex16(u64 x) {
  char buf[24];
  snprintf(buf, sizeof buf, "%016llx", (unsigned long long)x);
  return buf;
}
}  // namespace

std::string resultJson(const ResultInfo& r) {
  if (r.kind != WorkTask::PRP && r.kind != WorkTask::LL && r.kind != WorkTask::CERT) {
    throw std::invalid_argument("result kind");
  }
  std::vector<std::string> f;
  if (r.kind == WorkTask::CERT) {
    if (r.sha3Hash.size() != 64) { throw std::invalid_argument("CERT result: sha3-hash must be 64 hex digits"); }
    f = {field("worktype", str("Cert")), field("exponent", std::to_string(r.exponent)),
         field("sha3-hash", str(r.sha3Hash)), field("squarings", std::to_string(r.squarings)),
         field("fft-length", std::to_string(r.fftLength)), field("shift-count", "0"),
         field("error-code", str("00000000"))};
  } else if (r.kind == WorkTask::PRP) {
    f = {field("status", str(r.isPrime ? "P" : "C")), field("exponent", std::to_string(r.exponent)),
         field("worktype", str("PRP-3")), field("res64", str(hex16(r.res64))), field("res2048", str(r.res2048)),
         field("residue-type", "1"), field("errors", object({field("gerbicz", std::to_string(r.nErrors))})),
         field("fft-length", std::to_string(r.fftLength))};
    if (r.proofPower > 0) {
      f.push_back(field("proof", object({field("version", "1"), field("power", std::to_string(r.proofPower)),
                                         field("hashsize", "64"), field("md5", str(r.proofMd5))})));
    }
  } else {
    f = {field("status", str(r.isPrime ? "P" : "C")), field("exponent", std::to_string(r.exponent)),
         field("worktype", str("LL")), field("res64", str(hex16(r.res64))),
         field("fft-length", std::to_string(r.fftLength)), field("shift-count", "0"),
         field("error-code", str("00000000"))};
  }
  utsname u{};
  uname(&u);
  constexpr int kLinux64 = 8;  // the PrimeNet platform code the reference reports on 64-bit Linux
  f.push_back(field("program", object({field("name", str(r.programName)), field("version", str(r.programVersion)),
                                       field("port", std::to_string(kLinux64)),
                                       field("os", object({field("os", str(u.sysname)), maybe("version", u.release),
                                                           maybe("architecture", u.machine)}))})));
  f.push_back(maybe("user", r.user));
  f.push_back(maybe("aid", r.aid));
  f.push_back(maybe("uid", r.uid));
  char ts[32];
  const time_t now = time(nullptr);
  tm g{};
  gmtime_r(&now, &g);
  strftime(ts, sizeof ts, "%Y-%m-%d %H:%M:%S", &g);
  f.push_back(field("timestamp", str(ts)));
  return object(f);
}
