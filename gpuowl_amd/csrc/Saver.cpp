#include "Saver.h"

#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <memory>
#include <type_traits>

namespace fs = std::filesystem;

namespace {

// byte-at-a-time table of the reflected CRC-32 polynomial
struct CrcTable {
  u32 t[256];
  CrcTable() {
    for (u32 i = 0; i < 256; ++i) {
      u32 c = i;
      for (int b = 0; b < 8; ++b) { c = (c & 1) ? (c >> 1) ^ 0xEDB88320u : c >> 1; }
      t[i] = c;
    }
  }
};

// Header lines of the reference's formats (src/Saver.cpp:17-30). The printf forms give
// the same bytes the reference writes with its scanf strings.
constexpr const char* kPRP13Scan = "OWL PRP 13 N=1*2^%u-1 k=%u block=%u res64=%16" SCNx64 " err=%u time=%lf";
constexpr const char* kPRP13Print = "OWL PRP 13 N=1*2^%u-1 k=%u block=%u res64=%016" PRIx64 " err=%u time=%f\n";
constexpr const char* kPRP12Scan = "OWL PRP 12 %u %u %u %16" SCNx64 " %u %u";
constexpr const char* kLL13Scan = "OWL LL 13 N=1*2^%u-1 k=%u time=%lf";
constexpr const char* kLL13Print = "OWL LL 13 N=1*2^%u-1 k=%u time=%f\n";
constexpr const char* kLL1Scan = "OWL LL 1 E=%u k=%u CRC=%u";

struct FileCloser {
  void operator()(FILE* f) const { fclose(f); }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

FilePtr openRead(const std::string& path) {
  FilePtr f{fopen(path.c_str(), "rb")};
  if (!f) { throw SavefileError("cannot open " + path); }
  return f;
}

std::string readLine(FILE* f, const std::string& path) {
  std::string s;
  for (int c; (c = fgetc(f)) != EOF;) {
    if (c == '\n') { return s; }
    s.push_back(char(c));
    if (s.size() > 512) { break; }
  }
  throw SavefileError("no header line in " + path);
}

Words readWords(FILE* f, u32 n, const std::string& path) {
  Words w(n);
  if (fread(w.data(), sizeof(u32), n, f) != n) { throw SavefileError("short file " + path); }
  return w;
}

u32 readU32(FILE* f, const std::string& path) { return readWords(f, 1, path)[0]; }

void checkCrc(const Words& w, u32 crc, const std::string& path) {
  if (crc32(w.data(), w.size() * sizeof(u32)) != crc) { throw SavefileError("CRC mismatch in " + path); }
}

Words fullWords(Words w, u32 E) {
  if (w.size() > nWords(E)) { throw std::invalid_argument("more words than the exponent has"); }
  w.resize(nWords(E), 0);
  return w;
}

// header line, u32 CRC, words: written to path + ".new", then renamed over path
// (src/CycleFile.cpp:6-21)
void writeFile(const std::string& path, const std::string& header, const Words& words) {
  const std::string tmp = path + ".new";
  {
    FilePtr f{fopen(tmp.c_str(), "wb")};
    if (!f) { throw SavefileError("cannot write " + tmp); }
    const u32 crc = crc32(words.data(), words.size() * sizeof(u32));
    bool ok = fwrite(header.data(), 1, header.size(), f.get()) == header.size() &&
              fwrite(&crc, sizeof crc, 1, f.get()) == 1 &&
              fwrite(words.data(), sizeof(u32), words.size(), f.get()) == words.size() && fflush(f.get()) == 0;
    if (!ok) { throw SavefileError("write failed: " + tmp); }
  }
  std::error_code ec;
  fs::rename(tmp, path, ec);
  if (ec) {
    fs::remove(path, ec);
    fs::rename(tmp, path);
  }
}

std::string str9(u32 k) {
  char buf[16];
  snprintf(buf, sizeof buf, "%09u", k);
  return buf;
}

}  // namespace

u32 crc32(const void* data, size_t size) {
  static const CrcTable tab;
  u32 c = ~0u;
  const unsigned char* p = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < size; ++i) { c = tab.t[(c ^ p[i]) & 0xff] ^ (c >> 8); }
  return ~c;
}

PRPState readPRPState(const std::string& path) {
  FilePtr f = openRead(path);
  const std::string header = readLine(f.get(), path);
  PRPState s;
  u32 crc = 0;
  if (sscanf(header.c_str(), kPRP13Scan, &s.exponent, &s.k, &s.blockSize, &s.res64, &s.nErrors, &s.elapsed) == 6) {
    crc = readU32(f.get(), path);
  } else if (sscanf(header.c_str(), kPRP12Scan, &s.exponent, &s.k, &s.blockSize, &s.res64, &s.nErrors, &crc) == 6) {
    s.elapsed = 0;
  } else {
    throw SavefileError("bad PRP header in " + path + ": '" + header + "'");
  }
  if (!s.exponent || !s.blockSize) { throw SavefileError("bad PRP header values in " + path); }
  s.check = readWords(f.get(), nWords(s.exponent), path);
  checkCrc(s.check, crc, path);
  return s;
}

LLState readLLState(const std::string& path) {
  FilePtr f = openRead(path);
  const std::string header = readLine(f.get(), path);
  LLState s;
  u32 crc = 0;
  if (sscanf(header.c_str(), kLL13Scan, &s.exponent, &s.k, &s.elapsed) == 3) {
    crc = readU32(f.get(), path);
  } else if (sscanf(header.c_str(), kLL1Scan, &s.exponent, &s.k, &crc) == 3) {
    s.elapsed = 0;
  } else {
    throw SavefileError("bad LL header in " + path + ": '" + header + "'");
  }
  if (!s.exponent) { throw SavefileError("bad LL header values in " + path); }
  s.data = readWords(f.get(), nWords(s.exponent), path);
  checkCrc(s.data, crc, path);
  return s;
}

void writePRPState(const std::string& path, const PRPState& s) {
  char header[256];
  snprintf(header, sizeof header, kPRP13Print, s.exponent, s.k, s.blockSize, s.res64, s.nErrors, s.elapsed);
  writeFile(path, header, fullWords(s.check, s.exponent));
}

void writeLLState(const std::string& path, const LLState& s) {
  char header[256];
  snprintf(header, sizeof header, kLL13Print, s.exponent, s.k, s.elapsed);
  writeFile(path, header, fullWords(s.data, s.exponent));
}

// ---------------------------------------------------------------- Saver
namespace {
template<typename State> std::string baseDir(const std::string& dir, u32 E) {
  const std::string name =
      std::is_same_v<State, PRPState> ? std::to_string(E) : std::string(State::KIND) + "-" + std::to_string(E);
  return (fs::path(dir.empty() ? "." : dir) / name).string();
}
template<typename State> State readAny(const std::string& path);
template<> PRPState readAny<PRPState>(const std::string& path) { return readPRPState(path); }
template<> LLState readAny<LLState>(const std::string& path) { return readLLState(path); }
void writeAny(const std::string& path, const PRPState& s) { writePRPState(path, s); }
void writeAny(const std::string& path, const LLState& s) { writeLLState(path, s); }
}  // namespace

template<typename State>
Saver<State>::Saver(const std::string& dir, u32 exponent, u32 blockSize, u32 nSavefiles)
    : exponent_{exponent}, blockSize_{blockSize}, nSavefiles_{std::max<u32>(1, nSavefiles)},
      base_{baseDir<State>(dir, exponent)}, prefix_{std::to_string(exponent) + "-"} {
  if (!exponent) { throw std::invalid_argument("exponent"); }
  fs::create_directories(base_);
}

template<typename State>
std::vector<u32> Saver<State>::savedKs() const {
  std::vector<u32> ks;
  const std::string suffix = std::string(".") + State::KIND;
  for (const auto& e : fs::directory_iterator(base_)) {
    if (!e.is_regular_file()) { continue; }
    const std::string name = e.path().filename().string();
    if (name.size() <= prefix_.size() + suffix.size() || name.rfind(prefix_, 0) != 0 ||
        name.compare(name.size() - suffix.size(), suffix.size(), suffix) != 0) {
      continue;
    }
    const std::string id = name.substr(prefix_.size(), name.size() - prefix_.size() - suffix.size());
    if (id.empty() || id.size() > 10 || !std::all_of(id.begin(), id.end(), [](char c) { return c >= '0' && c <= '9'; })) {
      continue;  // "unverified" and foreign names
    }
    ks.push_back(u32(std::stoul(id)));
  }
  std::sort(ks.begin(), ks.end());
  return ks;
}

template<typename State>
std::string Saver<State>::mostRecent() const {
  if constexpr (std::is_same_v<State, PRPState>) {
    const std::string unverified = (fs::path(base_) / (prefix_ + "unverified.prp")).string();
    if (fs::is_regular_file(unverified)) { return unverified; }
  }
  std::vector<u32> ks = savedKs();
  if (ks.empty()) { return {}; }
  return (fs::path(base_) / (prefix_ + str9(ks.back()) + "." + State::KIND)).string();
}

template<>
PRPState Saver<PRPState>::initState() const {
  PRPState s;
  s.exponent = exponent_;
  s.blockSize = blockSize_;
  s.res64 = 3;
  s.check = Words(nWords(exponent_), 0);
  s.check[0] = 1;
  return s;
}

template<>
LLState Saver<LLState>::initState() const {
  LLState s;
  s.exponent = exponent_;
  s.data = Words(nWords(exponent_), 0);
  s.data[0] = 4;
  return s;
}

template<typename State>
State Saver<State>::load() {
  for (int attempt = 0; attempt < 2; ++attempt) {
    const std::string path = mostRecent();
    if (path.empty()) { return initState(); }
    try {
      State s = readAny<State>(path);
      if (s.exponent == exponent_) { return s; }
    } catch (const SavefileError&) {
    }
    fs::rename(path, path + ".bad");  // out of the way; the next most recent is tried
  }
  throw SavefileError("no readable savefile in " + base_);
}

template<typename State>
void Saver<State>::save(const State& s) {
  if (s.exponent != exponent_) { throw std::invalid_argument("savefile of another exponent"); }
  writeAny((fs::path(base_) / (prefix_ + str9(s.k) + "." + State::KIND)).string(), s);
  trim();
  if constexpr (std::is_same_v<State, PRPState>) {
    std::error_code ec;
    fs::remove(fs::path(base_) / (prefix_ + "unverified.prp"), ec);
  }
}

template<typename State>
void Saver<State>::saveUnverified(const State& s) const {
  if constexpr (std::is_same_v<State, PRPState>) {
    writeAny((fs::path(base_) / (prefix_ + "unverified.prp")).string(), s);
  } else {
    throw std::logic_error("only PRP states are saved unverified");
  }
}

template<typename State>
void Saver<State>::dropMostRecent() {
  const std::string path = mostRecent();
  if (!path.empty()) { fs::rename(path, path + ".bad"); }
}

// Keeps the newest nSavefiles files. (The reference thins older files by their spacing
// and roundness of k, src/Saver.cpp:247-276; which old files remain does not affect
// resuming, which always takes the most recent.)
template<typename State>
void Saver<State>::trim() {
  std::vector<u32> ks = savedKs();
  for (size_t i = 0; i + nSavefiles_ < ks.size(); ++i) {
    std::error_code ec;
    fs::remove(fs::path(base_) / (prefix_ + str9(ks[i]) + "." + State::KIND), ec);
  }
}

template<typename State>
void Saver<State>::clear(const std::string& dir, u32 exponent) {
  std::error_code ec;
  fs::remove_all(baseDir<State>(dir, exponent), ec);
}

template class Saver<PRPState>;
template class Saver<LLState>;
