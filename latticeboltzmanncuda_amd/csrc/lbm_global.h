// Loads and stores through the global address space.
//
// Every buffer the kernels touch is device memory from hipMalloc.  When a
// kernel reaches a buffer through a pointer read from memory (the batch
// kernels' lattice descriptors) instead of a kernel argument, the compiler
// cannot tell it is global and emits flat instructions, which count against
// both the vector-memory and the LDS wait counters and so serialise the
// loads.  Casting to address space 1 restores global_load / global_store.
#pragma once

#include <hip/hip_runtime.h>

namespace lbm {

template <class T>
__device__ __forceinline__ T gload(const T* p) {
  using G = const __attribute__((address_space(1))) T;
  return *(G*)p;
}

template <class T>
__device__ __forceinline__ void gstore(T* p, T v) {
  using G = __attribute__((address_space(1))) T;
  *(G*)p = v;
}

// non-temporal store: the line is written to HBM without being kept in the
// caches
template <class T>
__device__ __forceinline__ void gstore_nt(T* p, T v) {
  using G = __attribute__((address_space(1))) T;
  __builtin_nontemporal_store(v, (G*)p);
}

// A buffer resource over `bytes` bytes from base (raw, untyped access).
// A buffer load takes a 32-bit byte offset whose constant part the
// compiler moves into the instruction -- a global load can do that only
// where it proves the 32-bit sum does not wrap -- and an offset at or past
// `bytes` reads 0 instead of touching memory.  Dword 3 is the value for
// raw buffer access on the gfx9 family (32-bit data format).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}

// 16-bit load through a buffer resource, at byte offset off
template <class T>
__device__ __forceinline__ T bload16(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  static_assert(sizeof(T) == 2, "16-bit loads only");
  return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0));
}

// 16-bit store through a buffer resource at byte offset off; NT: with the
// non-temporal hint (as gstore_nt)
template <bool NT, class T>
__device__ __forceinline__ void bstore16(__amdgpu_buffer_rsrc_t r, uint32_t off, T v) {
  static_assert(sizeof(T) == 2, "16-bit stores only");
  __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, (int)off, 0,
                                        NT ? 2 : 0);
}

// a * b + c with one rounding
__device__ __forceinline__ float fmadd(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fmadd(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

}  // namespace lbm
