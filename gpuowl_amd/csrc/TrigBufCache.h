// Process-wide cache of the shape-dependent device twiddle tables, shared by every
// Gpu of the same FFT geometry (the reference's TrigBufCache, src/TrigBufCache.h:13-48).
#pragma once

#include "Buffer.h"

#include <hip/hip_vector_types.h>

#include <cstdint>
#include <memory>

using u32 = uint32_t;

struct TrigTables {
  Buffer<double2> row;    // W: exp(-2 pi i m / W)
  Buffer<double2> col;    // H: exp(-2 pi i m / H)
  Buffer<double2> midA;   // W*64: exp(-2 pi i a*lo / ND)
  Buffer<double2> midB;   // W*ceil(H/64): exp(-2 pi i a*64*hi / ND)
  Buffer<double2> pair;   // W: exp(-2 pi i a / ND)
  Buffer<double2> midS;   // W*16: exp(-2 pi i a*k*(H/16) / ND)
};

class TrigBufCache {
public:
  // Tables for a W x H transform, built once per process (thread-safe) on queue q.
  static std::shared_ptr<const TrigTables> get(u32 W, u32 H, owl_queue q);
};
