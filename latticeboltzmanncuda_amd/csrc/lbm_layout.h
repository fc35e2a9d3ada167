// Device layout of the populations: blocked structure of arrays.
//
// Nodes are grouped in tiles of 64 consecutive indices n (one wavefront);
// inside a tile each direction i owns a contiguous run of 64 values:
//
//   element (i, n) -> (n / 64) * (Q * 64) + i * 64 + (n % 64)
//
// A wavefront therefore writes one contiguous Q*64-element chunk per step
// and reads at most a few neighbouring chunks, instead of touching Q planes
// that are N elements apart (measured on MI355X: ~5 % faster than plain
// per-direction planes, ~12 % faster than unpadded planes).  Indices are
// 32-bit; lbm_create rejects lattices whose index range does not fit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbm_global.h"

namespace lbm {

constexpr uint32_t kTile = 64;
constexpr uint32_t kTileShift = 6;

// Element indices are below 2^32 (lbm_create checks), so the tile number
// n >> 6 is below 2^32 / (64 Q) < 2^24 and the product fits the GPU's
// full-rate 24-bit multiply.
template <int Q>
__host__ __device__ __forceinline__ uint32_t pidx(uint32_t n, int i) {
#ifdef __HIP_DEVICE_COMPILE__
  const uint32_t tile = __umul24(n >> kTileShift, (uint32_t)(Q * kTile));
#else
  const uint32_t tile = (n >> kTileShift) * (uint32_t)(Q * kTile);
#endif
  return tile + (uint32_t)i * kTile + (n & (kTile - 1));
}

// Elements of one population array for N nodes.
inline uint64_t population_elems(int q, uint64_t n) {
  return (n + kTile - 1) / kTile * kTile * (uint64_t)q;
}

// Node classes of the 2-bit class map used by the standard step.
constexpr uint32_t kClassInterior = 0;  // fluid, every link streams
constexpr uint32_t kClassBoundary = 1;  // fluid with bounce-back / lid links
constexpr uint32_t kClassSolid = 2;     // not updated; holds the rest state

// D2Q9 lattices (other than FP64) keep a 16-bit copy of their link masks
// after the 32-bit ones (their mask array holds 6 bytes per node):
// bits 0-13 as in the 32-bit mask, bit 14 next to the moving wall (the
// 32-bit mask's bit 30), bit 15 solid (bit 31).
constexpr uint32_t kMask16Lid = 1u << 14, kMask16Solid = 1u << 15;

__device__ __forceinline__ uint32_t node_class(const uint32_t* cls,
                                               uint32_t n) {
  return (gload(cls + (n >> 4)) >> (2u * (n & 15u))) & 3u;
}

}  // namespace lbm
