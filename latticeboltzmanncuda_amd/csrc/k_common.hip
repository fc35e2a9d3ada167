// Kernels shared by both modes: conversion between the reference's 144-byte
// LatticePoint host records (simulation/include/LatticePoint.h:3-17) and the
// device layout (population planes + flag bytes), and the coordinate-mapping
// check that replaces the reference's test_coordinates kernel
// (test/CudaTestHelper.cu:12-31).
#include <hip/hip_runtime.h>

#include "lbm_internal.h"
#include "lbm_lattice.h"
#include "lbm_layout.h"
#include "lbm_store.h"

namespace lbm {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void decode(const Grid& g, uint32_t n, uint32_t& x,
                                       uint32_t& y, uint32_t& z) {
  const uint32_t t = g.dz.div(n);
  z = n - t * g.nz;
  x = g.dy.div(t);
  y = t - x * g.ny;
}

template <int Q, int S>
__global__ void __launch_bounds__(kBlock)
    k_points_in(Grid g, const lbm_lattice_point* __restrict__ pts,
                typename Store<S>::T* __restrict__ f,
                uint8_t* __restrict__ flags) {
  using L = Lat<Q>;
  using St = Store<S>;
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  const lbm_lattice_point* p = pts + n;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    float v = p->particle_distribution[i];
    if (St::kShifted) v -= (float)L::W[i];
    St::store(f + pidx<Q>(n, i), (typename St::C)v);
  }
  if (flags)
    flags[n] = (uint8_t)((p->isInternal ? LBM_FLAG_INTERNAL : 0u) |
                         (p->isReflected ? LBM_FLAG_REFLECTED : 0u));
}

template <int Q, int S>
__global__ void __launch_bounds__(kBlock)
    k_points_out(Grid g, const typename Store<S>::T* __restrict__ f,
                 const uint8_t* __restrict__ flags,
                 lbm_lattice_point* __restrict__ pts) {
  using L = Lat<Q>;
  using St = Store<S>;
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  uint32_t x, y, z;
  decode(g, n, x, y, z);
  lbm_lattice_point* p = pts + n;
  p->x = (int32_t)x;
  p->y = (int32_t)y;
  p->z = (int32_t)z;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    float v = 0.0f;
    if (i < Q) {
      v = (float)St::load(f + pidx<Q>(n, i));
      if (St::kShifted) v += (float)L::W[i < Q ? i : 0];
    }
    p->particle_distribution[i] = v;
  }
  const uint8_t fl = flags ? flags[n] : 0;
  p->isReflected = (fl & LBM_FLAG_REFLECTED) ? 1 : 0;
  p->isInternal = (fl & LBM_FLAG_INTERNAL) ? 1 : 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) p->pad_[k] = 0;
  p->reflections = nullptr;
  p->d_reflections = nullptr;
}

// Launched over a 3-D grid (x fastest in threads, as the reference's
// kernels map threads); each thread computes its node index, checks that
// decoding it gives back (x, y, z) and counts the visit.
__global__ void k_test_coords_visit(Grid g, uint32_t* __restrict__ counts,
                                    int* __restrict__ ok) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t y = blockIdx.y * blockDim.y + threadIdx.y;
  const uint32_t z = blockIdx.z * blockDim.z + threadIdx.z;
  if (x >= g.nx || y >= g.ny || z >= g.nz) return;
  const uint32_t n = z + g.nz * (y + g.ny * x);
  uint32_t dx, dy, dz;
  decode(g, n, dx, dy, dz);
  if (dx != x || dy != y || dz != z) atomicAnd(ok, 0);
  atomicAdd(counts + n, 1u);
}

__global__ void k_test_coords_check(Grid g, const uint32_t* __restrict__ counts,
                                    int* __restrict__ ok) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  if (counts[n] != 1u) atomicAnd(ok, 0);
}

// The bandwidth reference of this GPU: a 16-byte-per-lane copy with
// non-temporal loads and stores, the fastest of the copy forms measured on
// MI355X (318.8 MB each way: 6.66 TB/s; ordinary loads and stores 6.07,
// ordinary loads + non-temporal stores 6.36).
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(kBlock)
    k_copy_stream(const f32x4* __restrict__ a, f32x4* __restrict__ b, uint64_t n4) {
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(&a[i]), &b[i]);
}

inline unsigned blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

int launch_copy_stream(const void* src, void* dst, uint64_t bytes, hipStream_t s) {
  const uint64_t n4 = bytes / 16;
  const uint64_t nb = (n4 + kBlock - 1) / kBlock;
  if (nb == 0 || nb > 0x7fffffffull) return fail(LBM_ERR_INVALID, "copy size out of range");
  hipLaunchKernelGGL(k_copy_stream, dim3((unsigned)nb), dim3(kBlock), 0, s,
                     (const f32x4*)src, (f32x4*)dst, n4);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}


int launch_points_in(int q, int storage, const Grid& g,
                     const lbm_lattice_point* pts, void* f, uint8_t* flags,
                     hipStream_t s) {
  const int drc = dispatch_qs(q, storage, [&](auto tag) {
    using Tag = decltype(tag);
    using T = typename Store<Tag::S>::T;
    hipLaunchKernelGGL((k_points_in<Tag::Q, Tag::S>), dim3(blocks_for(g.n)),
                       dim3(kBlock), 0, s, g, pts, (T*)f, flags);
  });
  if (drc) return drc;
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_points_out(int q, int storage, const Grid& g, const void* f,
                      const uint8_t* flags, lbm_lattice_point* pts,
                      hipStream_t s) {
  const int drc = dispatch_qs(q, storage, [&](auto tag) {
    using Tag = decltype(tag);
    using T = typename Store<Tag::S>::T;
    hipLaunchKernelGGL((k_points_out<Tag::Q, Tag::S>), dim3(blocks_for(g.n)),
                       dim3(kBlock), 0, s, g, (const T*)f, flags, pts);
  });
  if (drc) return drc;
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_test_coordinates(const Grid& g, uint32_t* counts, int* ok,
                            hipStream_t s) {
  const dim3 blk(8, 8, 4);
  const dim3 grd((g.nx + 7) / 8, (g.ny + 7) / 8, (g.nz + 3) / 4);
  hipLaunchKernelGGL(k_test_coords_visit, grd, blk, 0, s, g, counts, ok);
  LBM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_test_coords_check, dim3(blocks_for(g.n)), dim3(kBlock),
                     0, s, g, counts, ok);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

}  // namespace lbm
