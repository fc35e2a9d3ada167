// Population storage types.  F32/F64 store f_i itself; F16/BF16 store the
// shifted deviation f_i - w_i (the rest-state value w_i carries most of the
// magnitude, so the 16-bit mantissa spends its bits on the deviation).  F16
// holds the deviation times 2^12 (exact both ways), so deviations down to
// 1.5e-8 are normal FP16 numbers; |f_i - w_i| < 65504 / 2^12 ~ 15.99 fits
// (a larger deviation is stored as +-Inf, which the residual and
// conservation checks then see).
//
// Conversions to 16 bits round to nearest, ties to even -- or, through
// store_sr, stochastically: a flow driven by a small body force changes its
// populations by less than half a 16-bit unit in the last place per step,
// so rounding to nearest would return the old value every step and the
// flow would not accelerate.  Stochastic rounding rounds up with the
// probability given by the dropped bits, so the increments survive on
// average.  Its random bits are a hash of the value, the node and the
// direction (sr_bits), so a lattice's result depends only on its own state.
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

// Random bits for stochastic rounding of the float with bit pattern b, the
// population of direction i of a node: one 24-bit multiply-add, the low 24
// bits of the value times the direction's odd multiplier (sr_mul) plus the
// node hash (sr_node_hash, the low 24 bits of the node's element index
// times an odd constant).  For a given value the hash, uniform over nodes,
// makes the bits uniform; the multipliers decorrelate a node's directions;
// a node's bits change from step to step with its values.  Callers take
// the top bits (taking low ones would let the compiler drop the 24-bit
// operand mask and fall back to the quarter-rate 32-bit multiply).
struct SrKey {
  uint32_t mul;   // odd, below 2^24
  uint32_t hash;  // node hash
};
constexpr uint32_t sr_mul(int i) { return ((0x9E3779B9u * (uint32_t)(i + 1)) >> 8) | 1u; }
// (the node hash's constant differs from every direction's multiplier, or
// the compiler would merge the two products into one 32-bit multiply)
__device__ __forceinline__ uint32_t sr_node_hash(uint32_t own) { return __umul24(own, 0x5BD1E9u); }
__device__ __forceinline__ SrKey sr_key(uint32_t node_hash, int i) { return {sr_mul(i), node_hash}; }
__device__ __forceinline__ uint32_t sr_bits(uint32_t b, SrKey k) { return __umul24(b, k.mul) + k.hash; }

// Each storage type also exposes its values in stored units (cvt_u,
// round_u, round_sr_u): the stored number itself, whose real value is
// kUnit times it.  Only FP16 has kUnit != 1 (2^-12).  The collision works
// in stored units, so the FP16 scale costs no multiply per population;
// scaling by a power of two commutes with every rounding, so the results
// are the same bits as in real units.
template <int S>
struct Store;

// Types whose stored units are real units.
#define LBM_UNIT_IS_VALUE                                                 \
  static constexpr C kUnit = 1;                                            \
  __device__ static C cvt_u(T v) { return cvt(v); }                        \
  __device__ static T round_u(C v) { return round(v); }                    \
  __device__ static T round_sr_u(C v, SrKey key) { return round_sr(v, key); }

template <>
struct Store<LBM_STORE_F32> {
  using T = float;  // stored
  using C = float;  // computed
  static constexpr bool kShifted = false;
  __device__ static C cvt(T v) { return v; }
  __device__ static T rest(int, double w) { return (T)w; }
  __device__ static C load(const T* p) { return gload(p); }
  __device__ static T round(C v) { return v; }
  // full-precision storage rounds to nearest
  __device__ static T round_sr(C v, SrKey) { return v; }
  __device__ static void store(T* p, C v) { gstore(p, v); }
  __device__ static void store_nt(T* p, C v) { gstore_nt(p, v); }
  __device__ static void store_sr(T* p, C v, SrKey) { gstore(p, v); }
  __device__ static void store_sr_nt(T* p, C v, SrKey) { gstore_nt(p, v); }
  LBM_UNIT_IS_VALUE
};

// Standard-mode FP32 storage (internal code; the public storage type is
// LBM_STORE_F32): the shifted deviation f_i - w_i in float, collided in
// the shifted form like the 16-bit types.  Increments far below f_i (a
// small body force) then keep their precision.  Reference-mode lattices
// store f_i itself (Store<LBM_STORE_F32>), as the reference does.
constexpr int kStoreF32Shifted = 4;

template <>
struct Store<kStoreF32Shifted> {
  using T = float;
  using C = float;
  static constexpr bool kShifted = true;
  __device__ static C cvt(T v) { return v; }
  __device__ static T rest(int, double) { return 0.0f; }
  __device__ static C load(const T* p) { return gload(p); }
  __device__ static T round(C v) { return v; }
  __device__ static T round_sr(C v, SrKey) { return v; }
  __device__ static void store(T* p, C v) { gstore(p, v); }
  __device__ static void store_nt(T* p, C v) { gstore_nt(p, v); }
  __device__ static void store_sr(T* p, C v, SrKey) { gstore(p, v); }
  __device__ static void store_sr_nt(T* p, C v, SrKey) { gstore_nt(p, v); }
  LBM_UNIT_IS_VALUE
};

// The storage code the kernels use for a lattice of this mode and public
// storage type.
inline int device_storage(int mode, int storage) {
  return mode == LBM_MODE_STANDARD && storage == LBM_STORE_F32 ? kStoreF32Shifted : storage;
}

template <>
struct Store<LBM_STORE_F64> {
  using T = double;
  using C = double;
  static constexpr bool kShifted = false;
  __device__ static C cvt(T v) { return v; }
  __device__ static T rest(int, double w) { return w; }
  __device__ static C load(const T* p) { return gload(p); }
  __device__ static T round(C v) { return v; }
  // full-precision storage rounds to nearest
  __device__ static T round_sr(C v, SrKey) { return v; }
  __device__ static void store(T* p, C v) { gstore(p, v); }
  __device__ static void store_nt(T* p, C v) { gstore_nt(p, v); }
  __device__ static void store_sr(T* p, C v, SrKey) { gstore(p, v); }
  __device__ static void store_sr_nt(T* p, C v, SrKey) { gstore_nt(p, v); }
  LBM_UNIT_IS_VALUE
};

template <>
struct Store<LBM_STORE_F16> {
  using T = _Float16;
  using C = float;
  static constexpr bool kShifted = true;
  static constexpr float kScale = 4096.0f, kUnscale = 1.0f / 4096.0f;
  __device__ static C cvt(T v) { return (float)v * kUnscale; }
  __device__ static T rest(int, double) { return (T)0; }
  __device__ static C load(const T* p) { return cvt(gload(p)); }
  __device__ static T round(C v) { return (_Float16)(v * kScale); }
  __device__ static T round_sr(C v, SrKey key) { return round_sr_u(v * kScale, key); }
  __device__ static void store(T* p, C v) { gstore(p, round(v)); }
  __device__ static void store_nt(T* p, C v) { gstore_nt(p, round(v)); }
  __device__ static void store_sr(T* p, C v, SrKey key) { gstore(p, round_sr(v, key)); }
  __device__ static void store_sr_nt(T* p, C v, SrKey key) { gstore_nt(p, round_sr(v, key)); }

  static constexpr C kUnit = kUnscale;
  __device__ static C cvt_u(T v) { return (float)v; }
  // where v comes straight from an fma (the collision's last step), the
  // compiler folds the fma into this conversion (v_fma_mixlo_f16, under
  // HIP's default contraction): the exact result is rounded once to FP16,
  // not first to FP32
  __device__ static T round_u(C v) { return (_Float16)v; }
  // gfx950's stochastically rounding conversion (v_cvt_sr_f16_f32) with
  // the hash as its random word.  Its rule, established on MI355X over
  // 65,536 (value, word) pairs: with r = the word's top 13 bits, a value
  // of FP16's normal range is converted as (its bit pattern + r) with the
  // 13 bits FP16 drops cut (exact; overflow gives Inf, NaN stays NaN); a
  // value below 2^-14 becomes floor(|v| 2^24 + r / 2^13) units of 2^-24
  // with v's sign -- stochastic rounding in the subnormal range too.  The
  // random word hashes the low 24 bits of the pattern, which scaling by
  // 2^12 leaves alone (mantissa, and the exponent's parity).
  __device__ static T round_sr_u(C v, SrKey key) {
    // the conversion writes the low half of its destination and keeps the
    // high half, which nothing reads: v's own register (dead afterwards)
    // serves, so no register is cleared for it
    typedef _Float16 H2 __attribute__((ext_vector_type(2)));
    const H2 h = __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(H2, v), v,
                                                  sr_bits(__float_as_uint(v), key), false);
    return h[0];
  }
};

template <>
struct Store<LBM_STORE_BF16> {
  using T = uint16_t;
  using C = float;
  static constexpr bool kShifted = true;
  __device__ static C cvt(T v) { return __uint_as_float(((uint32_t)v) << 16); }
  __device__ static T rest(int, double) { return 0; }
  __device__ static C load(const T* p) { return cvt(gload(p)); }
  // gfx950's conversions: v_cvt_pk_bf16_f32 rounds to nearest, ties to
  // even; v_cvt_sr_bf16_f32 rounds stochastically, adding its random word's
  // top 16 bits to the pattern and cutting the low 16 (established on
  // MI355X over 65,536 (value, word) pairs, FP32 subnormals included;
  // overflow gives Inf, NaN stays NaN)
  __device__ static T round(C v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
  __device__ static T round_sr(C v, SrKey key) {
    // the conversion keeps its destination's high half, which nothing
    // reads: v's own register (dead afterwards) serves
    typedef __bf16 B2 __attribute__((ext_vector_type(2)));
    const B2 h = __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(B2, v), v,
                                                   sr_bits(__float_as_uint(v), key), false);
    return __builtin_bit_cast(uint16_t, h[0]);
  }
  __device__ static void store(T* p, C v) { gstore(p, round(v)); }
  __device__ static void store_nt(T* p, C v) { gstore_nt(p, round(v)); }
  __device__ static void store_sr(T* p, C v, SrKey key) { gstore(p, round_sr(v, key)); }
  __device__ static void store_sr_nt(T* p, C v, SrKey key) { gstore_nt(p, round_sr(v, key)); }
  LBM_UNIT_IS_VALUE
};

#undef LBM_UNIT_IS_VALUE

inline size_t storage_bytes(int storage) {
  switch (storage) {
    case LBM_STORE_F32:
    case kStoreF32Shifted:
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
    case 9 * 8 + kStoreF32Shifted: fn(QS<9, kStoreF32Shifted>{}); return LBM_OK;
    case 19 * 8 + LBM_STORE_F32: fn(QS<19, LBM_STORE_F32>{}); return LBM_OK;
    case 19 * 8 + LBM_STORE_F64: fn(QS<19, LBM_STORE_F64>{}); return LBM_OK;
    case 19 * 8 + LBM_STORE_F16: fn(QS<19, LBM_STORE_F16>{}); return LBM_OK;
    case 19 * 8 + LBM_STORE_BF16: fn(QS<19, LBM_STORE_BF16>{}); return LBM_OK;
    case 19 * 8 + kStoreF32Shifted: fn(QS<19, kStoreF32Shifted>{}); return LBM_OK;
    case 27 * 8 + LBM_STORE_F32: fn(QS<27, LBM_STORE_F32>{}); return LBM_OK;
    case 27 * 8 + LBM_STORE_F64: fn(QS<27, LBM_STORE_F64>{}); return LBM_OK;
    case 27 * 8 + LBM_STORE_F16: fn(QS<27, LBM_STORE_F16>{}); return LBM_OK;
    case 27 * 8 + LBM_STORE_BF16: fn(QS<27, LBM_STORE_BF16>{}); return LBM_OK;
    case 27 * 8 + kStoreF32Shifted: fn(QS<27, kStoreF32Shifted>{}); return LBM_OK;
    default:
      return fail(LBM_ERR_UNSUPPORTED, "lattice/storage combination not built");
  }
}

}  // namespace lbm
