// Population storage types.  F32/F64 store f_i itself; F16/BF16 store the
// shifted deviation f_i - w_i (the rest-state value w_i carries most of the
// magnitude, so the 16-bit mantissa spends its bits on the deviation).
// Conversions to 16 bits round to nearest, ties to even.
//
// store_nt writes with the non-temporal hint: the line goes to HBM without
// being kept in L2 / the Infinity Cache.  For a lattice whose two
// population copies exceed the 256 MB Infinity Cache, the step kernel is
// ~7 % faster with it on MI355X (D3Q19 FP32 256x128x128: 113.6 -> 106.2 us);
// for smaller, cache-resident lattices it is slower (128^3: 26.3 -> 29.5 us),
// so the launchers pick it by size (streaming_stores).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/lbm.h"
#include "lbm_global.h"
#include "lbm_internal.h"

namespace lbm {

template <int S>
struct Store;

template <>
struct Store<LBM_STORE_F32> {
  using T = float;  // stored
  using C = float;  // computed
  static constexpr bool kShifted = false;
  __device__ static C cvt(T v) { return v; }
  __device__ static T rest(int, double w) { return (T)w; }
  __device__ static C load(const T* p) { return gload(p); }
  __device__ static void store(T* p, C v) { gstore(p, v); }
  __device__ static void store_nt(T* p, C v) { gstore_nt(p, v); }
};

template <>
struct Store<LBM_STORE_F64> {
  using T = double;
  using C = double;
  static constexpr bool kShifted = false;
  __device__ static C cvt(T v) { return v; }
  __device__ static T rest(int, double w) { return w; }
  __device__ static C load(const T* p) { return gload(p); }
  __device__ static void store(T* p, C v) { gstore(p, v); }
  __device__ static void store_nt(T* p, C v) { gstore_nt(p, v); }
};

template <>
struct Store<LBM_STORE_F16> {
  using T = _Float16;
  using C = float;
  static constexpr bool kShifted = true;
  __device__ static C cvt(T v) { return (float)v; }
  __device__ static T rest(int, double) { return (T)0; }
  __device__ static C load(const T* p) { return (float)gload(p); }
  __device__ static void store(T* p, C v) { gstore(p, (_Float16)v); }
  __device__ static void store_nt(T* p, C v) { gstore_nt(p, (_Float16)v); }
};

template <>
struct Store<LBM_STORE_BF16> {
  using T = uint16_t;
  using C = float;
  static constexpr bool kShifted = true;
  __device__ static C cvt(T v) { return __uint_as_float(((uint32_t)v) << 16); }
  __device__ static T rest(int, double) { return 0; }
  __device__ static C load(const T* p) { return cvt(gload(p)); }
  __device__ static T round(C v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
  }
  __device__ static void store(T* p, C v) { gstore(p, round(v)); }
  __device__ static void store_nt(T* p, C v) { gstore_nt(p, round(v)); }
};

inline size_t storage_bytes(int storage) {
  switch (storage) {
    case LBM_STORE_F32:
      return 4;
    case LBM_STORE_F64:
      return 8;
    case LBM_STORE_F16:
    case LBM_STORE_BF16:
      return 2;
    default:
      return 0;
  }
}

// Element idx of a population array.  Arrays hold fewer than 2^31 elements
// (lbm_create checks), so for 16-bit storage the byte offset fits 32 bits
// and the address is base (scalar) + 32-bit offset: one vector instruction
// instead of a 64-bit add.
template <class T>
__device__ __forceinline__ const T* elem(const T* base, uint32_t idx) {
  if constexpr (sizeof(T) == 2)
    return (const T*)((const char*)base + (uint32_t)(idx * 2u));
  else
    return base + idx;
}
template <class T>
__device__ __forceinline__ T* elem(T* base, uint32_t idx) {
  if constexpr (sizeof(T) == 2)
    return (T*)((char*)base + (uint32_t)(idx * 2u));
  else
    return base + idx;
}

// Non-temporal stores pay off once both population copies (2 x bytes of
// one copy) no longer fit the Infinity Cache.
constexpr uint64_t kStreamingStoreBytes = 256ull << 20;

inline bool streaming_stores(uint64_t both_copies_bytes) {
  return both_copies_bytes >= kStreamingStoreBytes;
}

// Calls fn(QS<Q, S>{}) for the runtime pair (q, storage), so that a
// generic lambda can instantiate kernels for compile-time Q and storage.
template <int Qv, int Sv>
struct QS {
  static constexpr int Q = Qv;
  static constexpr int S = Sv;
};

template <class Fn>
int dispatch_qs(int q, int storage, Fn&& fn) {
  switch (q * 8 + storage) {
    case 9 * 8 + LBM_STORE_F32: fn(QS<9, LBM_STORE_F32>{}); return LBM_OK;
    case 9 * 8 + LBM_STORE_F64: fn(QS<9, LBM_STORE_F64>{}); return LBM_OK;
    case 9 * 8 + LBM_STORE_F16: fn(QS<9, LBM_STORE_F16>{}); return LBM_OK;
    case 9 * 8 + LBM_STORE_BF16: fn(QS<9, LBM_STORE_BF16>{}); return LBM_OK;
    case 19 * 8 + LBM_STORE_F32: fn(QS<19, LBM_STORE_F32>{}); return LBM_OK;
    case 19 * 8 + LBM_STORE_F64: fn(QS<19, LBM_STORE_F64>{}); return LBM_OK;
    case 19 * 8 + LBM_STORE_F16: fn(QS<19, LBM_STORE_F16>{}); return LBM_OK;
    case 19 * 8 + LBM_STORE_BF16: fn(QS<19, LBM_STORE_BF16>{}); return LBM_OK;
    case 27 * 8 + LBM_STORE_F32: fn(QS<27, LBM_STORE_F32>{}); return LBM_OK;
    case 27 * 8 + LBM_STORE_F64: fn(QS<27, LBM_STORE_F64>{}); return LBM_OK;
    case 27 * 8 + LBM_STORE_F16: fn(QS<27, LBM_STORE_F16>{}); return LBM_OK;
    case 27 * 8 + LBM_STORE_BF16: fn(QS<27, LBM_STORE_BF16>{}); return LBM_OK;
    default:
      return fail(LBM_ERR_UNSUPPORTED, "lattice/storage combination not built");
  }
}

}  // namespace lbm
