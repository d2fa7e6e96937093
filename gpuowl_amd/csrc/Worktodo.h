// The serving front end's data formats, as the reference reads and writes them:
// worktodo lines (src/Worktodo.cpp:25-106) and the results JSON lines
// (src/Task.cpp:86-170), plus the PRP type-1 residue they report
// (x_E / 9 mod 2^E - 1, src/Gpu.cpp:1120-1148, 1495-1504).
#pragma once

#include "state.h"

#include <string>
#include <vector>

struct WorkTask {
  enum Kind { NONE = 0, PRP = 1, LL = 2, CERT = 3 };
  Kind kind = NONE;
  u32 exponent = 0;
  std::string aid;       // 32 hex digits, or empty
  u32 squarings = 0;     // CERT only
};

// One worktodo line: "PRP=" / "PRPDC=" (PRP), "Test=" / "DoubleCheck=" (LL) with an optional
// AID ("N/A" or empty skipped) and either the "1,2,E,-1,..." form or E first; "Cert=AID,1,2,E,-1,S".
// Comments, unknown kinds and exponents <= 1000 give kind NONE.
WorkTask parseWorktodoLine(const std::string& line);
// The reference's choice among lines: the smallest CERT, else the smallest exponent; -1 if none.
int bestWorktodoLine(const std::vector<std::string>& lines);

// x * 9^-1 mod 2^E - 1 for x < 2^E - 1 given as words (low first): the PRP-3 type-1 residue
// 3^(M-1) mod M from x_E = 3^(2^E).
Words div9(const Words& x, u32 E);
// the 64 low words of w as 512 hex digits, most significant first (src/Gpu.cpp:364-376)
std::string res2048Hex(const Words& w);

struct ResultInfo {
  WorkTask::Kind kind = WorkTask::PRP;
  u32 exponent = 0;
  bool isPrime = false;
  u64 res64 = 0;             // PRP: of the type-1 residue
  std::string res2048;       // PRP only
  u32 fftLength = 0;
  u32 nErrors = 0;           // PRP: Gerbicz errors
  u32 proofPower = 0;        // PRP: power of the proof made by the run, 0 = none
  std::string proofMd5;      // PRP: MD5 of that proof file (32 hex digits)
  std::string sha3Hash;      // CERT: SHA3-256 of the final residue, 64 hex digits
  u32 squarings = 0;         // CERT
  std::string aid, user, uid;
  std::string programName = "gpuowl_amd", programVersion = "1.0";
};
// One results line with the reference's keys and order: status, exponent, worktype, the
// kind's fields (PRP: with the proof object {version 1, power, hashsize 64, md5} when the run
// made a proof), program {name, version, port, os {os, version, architecture}}, user, aid,
// uid (empty ones omitted) and a UTC timestamp. A CERT line has no status and starts with
// worktype "Cert", exponent, sha3-hash, squarings, fft-length, shift-count, error-code
// (src/Task.cpp:182-197).
std::string resultJson(const ResultInfo& r);
