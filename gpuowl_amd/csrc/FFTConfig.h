// FFT shape and per-exponent FFT selection.
// Mirrors the reference interface src/FFTConfig.h:24-87 (FFTShape, FFTConfig,
// bestFit, maxBpw, carry32BPW, needsLargeCarry, isFavoredShape) so that the
// FFT length chosen for an exponent is the reference's, except at the top of the
// shapes where this engine's measured round-off calls for a lower limit.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

using u32 = uint32_t;
using u64 = uint64_t;

// Format n with a K or M suffix when a multiple of 1024 (src/FFTConfig.cpp:266-282).
std::string numberK(u32 n);

class FFTShape {
public:
  static constexpr float MIN_BPW = 3;
  static constexpr int NUM_BPW = 6;  // variants 000, 101, 202, 010, 111, 212

  // All WIDTH x MIDDLE x HEIGHT shapes with size in [from, to], in the reference's
  // preference order (src/FFTConfig.cpp:77-100).
  static std::vector<FFTShape> allShapes(u32 from = 0, u32 to = u32(-1));
  // "W:M:H", a size "6.5M", a range "6M-7M", or a comma list of these (src/FFTConfig.cpp:41-75).
  static std::vector<FFTShape> multiSpec(const std::string& spec);

  u32 width = 0, middle = 0, height = 0;
  std::array<double, NUM_BPW> bpw{};

  FFTShape(u32 w = 1, u32 m = 1, u32 h = 1);
  explicit FFTShape(const std::string& spec);

  u32 size() const { return width * middle * height * 2; }
  // this engine's max bits per word for the shape where its kernels' round-off (Gumbel Z = 28)
  // needs less than the reference's table allows; +infinity elsewhere (engine_bpw.inc)
  double engineCap() const;
  u32 nW() const { return (width == 1024 || width == 256) ? 4 : 8; }
  u32 nH() const { return (height == 1024 || height == 256) ? 4 : 8; }
  double maxBpw() const;
  std::string spec() const { return numberK(width) + ':' + numberK(middle) + ':' + numberK(height); }

  double carry32BPW() const;
  bool needsLargeCarry(u32 E) const { return E / double(size()) > carry32BPW(); }
  bool isFavoredShape() const;
};

enum CarryKind : u32 { CARRY_32 = 0, CARRY_64 = 1, CARRY_AUTO = 2 };

// A variant is three digits v_W v_M v_H with v_W, v_H < 3 and v_M < 2.
constexpr u32 LAST_VARIANT = 212;

struct FFTConfig {
  // spec empty: the first FFT of the loaded tune file (FftTune) whose maxExp covers E,
  // else the first shape (variants 101 then 202) that does (src/FFTConfig.cpp:235-263).
  // Otherwise the spec's FFT. refLimits: the reference's limits (refMaxExp) instead of this
  // engine's (maxExp) and no tune file, i.e. the reference's own choice.
  static FFTConfig bestFit(u32 E, const std::string& spec, double overdrive = 1.0, bool refLimits = false);

  FFTShape shape{};
  u32 variant = LAST_VARIANT;
  u32 carry = CARRY_AUTO;

  FFTConfig() = default;
  explicit FFTConfig(const std::string& spec);
  FFTConfig(FFTShape shape, u32 variant, u32 carry);

  std::string spec() const;
  u32 size() const { return shape.size(); }
  // this engine's limits: the reference's, capped by FFTShape::engineCap()
  u32 maxExp() const { return u32(maxBpw() * shape.size()); }
  double maxBpw() const;
  // the reference's limits (src/FFTConfig.cpp:218-233)
  u32 refMaxExp() const { return u32(refMaxBpw() * shape.size()); }
  double refMaxBpw() const;
};

// The reference's tune.txt (TuneEntry, src/TuneEntry.cpp:7-84): FFTs ordered by cost with
// maxExp increasing, so that each entry is the cheapest FFT up to its maxExp. Lines are
// "cost spec" with an optional "# maxExp" comment; written as "%6.1f %14s # %u".
struct FftTuneEntry {
  double cost = 0;
  FFTConfig fft;
};
namespace FftTune {
// Inserts e unless an entry at most as costly handles at least its maxExp; removes the
// costlier entries that handle at most e's maxExp. Returns whether the list changed.
bool update(std::vector<FftTuneEntry>& list, const FftTuneEntry& e);
std::vector<FftTuneEntry> read(const std::string& path);  // throws on an unreadable or unordered file
void write(const std::string& path, const std::vector<FftTuneEntry>& list);
// the process-wide list bestFit consults (empty: no tune file)
void set(std::vector<FftTuneEntry> list);
std::vector<FftTuneEntry> get();
}  // namespace FftTune
