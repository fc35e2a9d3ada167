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

// a * b + c with one rounding
__device__ __forceinline__ float fmadd(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fmadd(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

}  // namespace lbm
