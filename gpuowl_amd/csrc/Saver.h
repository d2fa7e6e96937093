// PRP and LL savefiles in the reference's formats (src/Saver.cpp:17-30, 89-143,
// src/File.h:213-231, src/CycleFile.cpp:6-21), so that runs resume across this engine
// and the reference in either direction:
//   PRP v13: "OWL PRP 13 N=1*2^E-1 k=K block=B res64=XXXXXXXXXXXXXXXX err=N time=T\n",
//            then u32 CRC32 of the words, then the Gerbicz check as (E-1)/32+1 u32 words;
//   PRP v12: "OWL PRP 12 E K B XXXXXXXXXXXXXXXX N CRC\n" then the words (read only);
//   LL  v13: "OWL LL 13 N=1*2^E-1 k=K time=T\n", CRC32, then the data words;
//   LL  v1:  "OWL LL 1 E=E k=K CRC=CRC\n" then the words (read only).
// The saved PRP check is the Gerbicz product before it takes x_k (the value the check
// buffer holds when a check point is reached), so x_k = 3 * check^(2^B - 1); res64 is
// that of x_k and validates the file on load. Files live in <dir>/<E>/ (PRP) or
// <dir>/ll-<E>/ (LL) as <E>-<k, 9 digits>.prp|.ll and <E>-unverified.prp; every write
// goes to "<name>.new" first and is renamed over the target.
#pragma once

#include "state.h"

#include <stdexcept>
#include <string>
#include <vector>

// CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320, initial and final complement):
// the checksum the reference's savefiles carry (src/common.cpp:26-39).
u32 crc32(const void* data, size_t size);

struct PRPState {
  static constexpr const char* KIND = "prp";
  u32 exponent = 0;
  u32 k = 0;
  u32 blockSize = 0;
  u64 res64 = 0;
  Words check;
  u32 nErrors = 0;
  double elapsed = 0;
};

struct LLState {
  static constexpr const char* KIND = "ll";
  u32 exponent = 0;
  u32 k = 0;
  Words data;
  double elapsed = 0;
};

struct SavefileError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Single files. read* accept the current and the previous format and throw SavefileError
// on a bad header, a short file or a CRC mismatch; write* write the current format.
PRPState readPRPState(const std::string& path);
LLState readLLState(const std::string& path);
void writePRPState(const std::string& path, const PRPState& s);
void writeLLState(const std::string& path, const LLState& s);

// The savefile directory of one exponent (the reference's Saver<State>,
// src/Saver.cpp:160-298): load() returns the most recent readable state (the
// unverified PRP file first, then the highest k; unreadable files are renamed to
// "*.bad"), or the start state (PRP: k = 0, res64 3, check 1; LL: k = 0, data 4) when
// there is none; save() writes <E>-<k>.<kind>, keeps at most nSavefiles of them
// (dropping the one whose removal leaves the smallest gap, round k preferred to keep)
// and removes the unverified file.
template<typename State>
class Saver {
public:
  Saver(const std::string& dir, u32 exponent, u32 blockSize, u32 nSavefiles = 20);

  State load();
  void save(const State& s);
  void saveUnverified(const State& s) const;  // PRP only
  void dropMostRecent();
  const std::string& base() const { return base_; }
  static void clear(const std::string& dir, u32 exponent);

private:
  u32 exponent_, blockSize_, nSavefiles_;
  std::string base_, prefix_;
  std::string mostRecent() const;
  std::vector<u32> savedKs() const;
  void trim();
  State initState() const;
};
