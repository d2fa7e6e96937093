// PRP proofs as the reference makes and checks them (src/Proof.cpp:20-297, src/Proof.h:43):
// residues of the PRP-3 run are kept at 2^power points (ProofSet), the proof is the final
// residue B and `power` middles M_i, each step's exponent h_i being the first 64 bits
// (little-endian) of a SHA3-256 hash chain over B and the middles; the file is
// "PRP PROOF\nVERSION=2\nHASHSIZE=64\nPOWER=p\nNUMBER=M<E>\n" then B and the middles as
// (E-1)/8+1 little-endian bytes each.
#pragma once

#include "state.h"

#include <array>
#include <string>
#include <vector>

// SHA3-256 (FIPS 202) of a byte stream.
class Sha3_256 {
public:
  Sha3_256();
  Sha3_256& update(const void* data, size_t n);
  std::array<u64, 4> finish();  // the 32-byte digest as four little-endian u64

private:
  u64 st_[25];
  unsigned char buf_[136];
  size_t fill_ = 0;
  void absorbBlock(const unsigned char* p);
};

// MD5 (RFC 1321) as 32 lowercase hex digits; of a whole file for the "proof" object of a PRP
// results line (proof::fileHash, src/Proof.cpp:32-39; MD5Hash::finish, src/MD5.h:21-31)
std::string md5Hex(const void* data, size_t n);
std::string md5File(const std::string& path);

// hash of the first (E-1)/8+1 bytes of the words, optionally after a previous digest
std::array<u64, 4> proofHash(u32 E, const Words& w);
std::array<u64, 4> proofHash(u32 E, const std::array<u64, 4>& prefix, const Words& w);

// The iterations at which a PRP run with proof power p keeps its residue
// (ProofSet::ProofSet, src/Proof.cpp:130-164): start from {0}; for each level with span
// s = ceil(E/2), ceil(s/2), ... add s to every point so far; 0 becomes E. Sorted, 2^p points.
std::vector<u32> proofPoints(u32 E, u32 power);

// The residues a PRP run keeps for its proof (ProofSet::save / load, src/Proof.cpp:230-243):
// one file per point, dir/<k>, holding the CRC-32 of the words and the (E-1)/32+1 words.
class ProofStore {
public:
  ProofStore(const std::string& dir, u32 E, u32 power);
  const std::vector<u32>& points() const { return points_; }
  u32 power() const { return power_; }
  u32 next(u32 k) const;  // the first point > k, or ~0u
  void save(u32 k, const Words& w) const;
  Words load(u32 k) const;

private:
  std::string dir_;
  u32 E_, power_;
  std::vector<u32> points_;
};

struct PrpProof {
  u32 E = 0;
  Words B;                     // the final residue x_E
  std::vector<Words> middles;  // one per level
  void save(const std::string& path) const;
  static PrpProof load(const std::string& path);
};
