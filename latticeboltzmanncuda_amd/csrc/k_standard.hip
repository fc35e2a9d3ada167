// Standard-mode kernels: the fused BGK pull collide+stream step, the link
// mask and node-class builders, equilibrium initialisation, moments and
// FP64 conservation sums.  Populations use the blocked layout of
// lbm_layout.h; node index n = z + nz*(y + ny*x).
//
// One thread per node over a flat 1-D grid (coalesced along n, whatever the
// lattice shape); (x, y, z) come from two magic-number divisions.
//
// The per-node update is std_node_step (lbm_step_node.h); large D2Q9
// lattices with 16-bit storage take std_group_step instead, 2 or 4 nodes
// per thread (k_std_stepv, chosen by group_step_width).  Lattices whose
// two population copies exceed the Infinity Cache use non-temporal stores
// and walk the tiles in opposite orders on consecutive steps; smaller ones
// use ordinary stores and a fixed order (both measured fastest on MI355X).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "lbm_internal.h"
#include "lbm_lattice.h"
#include "lbm_layout.h"
#include "lbm_step_node.h"
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

// D2Q9 lattices with 16-bit storage hand each XCD runs of kXcdRun
// consecutive blocks (workgroup w runs on XCD w % 8; the grid is rounded
// up to a multiple of 8 kXcdRun, and the map below is a permutation of
// it), so a block's neighbouring rows are mostly read through the same
// L2: 4096^2 one node per thread 2-5 %, four nodes per thread 1-1.5 %
// faster.  The 3-D lattices keep the hardware's round-robin order (the
// D3Q19 16-bit Taylor-Green step was 4 % slower in runs).
constexpr uint32_t kXcdRun = 16;
template <int Q, int S>
constexpr bool kXcdRuns = Q == 9 && sizeof(typename Store<S>::T) == 2;
inline unsigned xcd_run_grid(unsigned nb) {
  return (nb + 8u * kXcdRun - 1u) / (8u * kXcdRun) * (8u * kXcdRun);
}
__device__ __forceinline__ uint32_t xcd_run_block(uint32_t w) {
  const uint32_t j = w / 8u;
  return (j / kXcdRun) * (8u * kXcdRun) + (w % 8u) * kXcdRun + j % kXcdRun;
}

template <int Q, int S, bool LID, bool NT>
__global__ void __launch_bounds__(kBlock)
    k_std_step(Grid g, const typename Store<S>::T* __restrict__ A,
               typename Store<S>::T* __restrict__ B,
               const uint32_t* __restrict__ mask,
               const uint32_t* __restrict__ cls, StdParams p, int rev) {
  const uint32_t w = kXcdRuns<Q, S> ? xcd_run_block(blockIdx.x) : blockIdx.x;
  const uint32_t blk = rev ? gridDim.x - 1 - w : w;
  const uint32_t n = blk * kBlock + threadIdx.x;
  if (n >= g.n) return;
  std_node_step<Q, S, LID, NT>(g, A, B, mask, cls, p, n);
}

// Groups of NV nodes per thread (std_group_step).
template <int Q, int S, int NV, bool LID, bool NT>
__global__ void __launch_bounds__(kBlock)
    k_std_stepv(Grid g, const typename Store<S>::T* __restrict__ A,
                typename Store<S>::T* __restrict__ B, const uint32_t* __restrict__ mask,
                const uint32_t* __restrict__ cls, StdParams p, int rev) {
  const uint32_t w = xcd_run_block(blockIdx.x);  // D2Q9 16-bit only
  const uint32_t blk = rev ? gridDim.x - 1 - w : w;
  const uint32_t n0 = (blk * kBlock + threadIdx.x) * (uint32_t)NV;
  if (n0 >= g.n) return;  // the rows of active groups are whole
  std_group_step<Q, S, NV, LID, NT>(g, A, B, mask, cls, p, n0);
}

template <int Q>
__global__ void __launch_bounds__(kBlock)
    k_build_mask(Grid g, const uint8_t* __restrict__ solid, int px, int py,
                 int pz, int lid_face, uint32_t* __restrict__ mask) {
  using L = Lat<Q>;
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  if (solid && solid[n]) {
    mask[n] = LBM_MASK_SOLID;
    return;
  }
  uint32_t x, y, z;
  decode(g, n, x, y, z);
  const int c[3] = {(int)x, (int)y, (int)z};
  const int dims[3] = {(int)g.nx, (int)g.ny, (int)g.nz};
  const int per[3] = {px, py, pz};
  uint32_t m = 0;
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    int s[3];
    bool out = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      s[a] = c[a] - L::E[i][a];
      if (s[a] < 0 || s[a] >= dims[a]) {
        if (per[a])
          s[a] = s[a] < 0 ? s[a] + dims[a] : s[a] - dims[a];
        else
          out = true;
      }
    }
    if (!out && solid) {
      const uint32_t sn = (uint32_t)s[2] + g.nz * ((uint32_t)s[1] + g.ny * (uint32_t)s[0]);
      out = solid[sn] != 0;
    }
    if (out) m |= 1u << i;
  }
  if (lid_face >= 0) {
    const int ax = lid_face / 2, side = lid_face % 2;
    if ((side == 1 && c[ax] == dims[ax] - 1) || (side == 0 && c[ax] == 0))
      m |= LBM_MASK_LID;
  }
  mask[n] = m;
}

// The 16-bit copy of D2Q9 link masks, after the 32-bit ones (layout in
// lbm_layout.h).
__global__ void __launch_bounds__(kBlock) k_mask16(Grid g, uint32_t* __restrict__ mask) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  const uint32_t m = mask[n];
  reinterpret_cast<uint16_t*>(mask + g.n)[n] =
      (uint16_t)((m & (kMask16Lid - 1u)) | ((m & LBM_MASK_LID) ? kMask16Lid : 0u) |
                 ((m & LBM_MASK_SOLID) ? kMask16Solid : 0u));
}

// Packs the 2-bit node classes of 16 nodes per word from the link masks.
__global__ void __launch_bounds__(kBlock)
    k_build_class(Grid g, const uint32_t* __restrict__ mask,
                  uint32_t* __restrict__ cls) {
  const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= (g.n + 15) / 16) return;
  uint32_t out = 0;
  for (uint32_t k = 0; k < 16; ++k) {
    const uint32_t n = w * 16 + k;
    uint32_t c = kClassInterior;
    if (n < g.n) {
      const uint32_t m = mask[n];
      c = (m & LBM_MASK_SOLID) ? kClassSolid
                               : (m ? kClassBoundary : kClassInterior);
    }
    out |= c << (2 * k);
  }
  cls[w] = out;
}

template <int Q, int S>
__global__ void __launch_bounds__(kBlock)
    k_init_eq(Grid g, const double* __restrict__ rho,
              const double* __restrict__ u, typename Store<S>::T* __restrict__ f) {
  using L = Lat<Q>;
  using St = Store<S>;
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  const double r = rho ? rho[n] : 1.0;
  const double ux = u ? u[n] : 0.0, uy = u ? u[g.n + n] : 0.0,
               uz = u ? u[2 * (uint64_t)g.n + n] : 0.0;
  const double usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const double eu = L::E[i][0] * ux + L::E[i][1] * uy + L::E[i][2] * uz;
    double v = L::W[i] * r * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq);
    if (St::kShifted) v -= L::W[i];
    St::store(f + pidx<Q>(n, i), (typename St::C)v);
  }
}

template <int Q, int S>
__global__ void __launch_bounds__(kBlock)
    k_std_moments(Grid g, const typename Store<S>::T* __restrict__ f,
                  const uint32_t* __restrict__ mask, float fx, float fy,
                  float fz, float* __restrict__ rho, float* __restrict__ u) {
  using L = Lat<Q>;
  using St = Store<S>;
  using C = typename St::C;
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  float r_out = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
  if (!(mask && (mask[n] & LBM_MASK_SOLID))) {
    C r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      C v = St::load(f + pidx<Q>(n, i));
      if (St::kShifted) v += (C)L::W[i];
      r += v;
      jx += (C)L::E[i][0] * v;
      jy += (C)L::E[i][1] * v;
      jz += (C)L::E[i][2] * v;
    }
    r_out = (float)r;
    ux = (float)((jx - (C)0.5 * (C)fx) / r);
    uy = (float)((jy - (C)0.5 * (C)fy) / r);
    uz = (float)((jz - (C)0.5 * (C)fz) / r);
  }
  if (rho) rho[n] = r_out;
  if (u) {
    u[n] = ux;
    u[g.n + n] = uy;
    u[2 * (uint64_t)g.n + n] = uz;
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Per-block FP64 sums of mass and momentum over non-solid nodes; the sum
// order is fixed (grid-stride per thread, then a fixed tree).
template <int Q, int S>
__global__ void __launch_bounds__(kBlock)
    k_conservation(Grid g, const typename Store<S>::T* __restrict__ f,
                   const uint32_t* __restrict__ mask,
                   double* __restrict__ partials) {
  using L = Lat<Q>;
  using St = Store<S>;
  double acc[4] = {0, 0, 0, 0};
  for (uint32_t n = blockIdx.x * kBlock + threadIdx.x; n < g.n;
       n += gridDim.x * kBlock) {
    if (mask && (mask[n] & LBM_MASK_SOLID)) continue;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      double v = (double)St::load(f + pidx<Q>(n, i));
      if (St::kShifted) v += L::W[i];
      acc[0] += v;
      acc[1] += L::E[i][0] * v;
      acc[2] += L::E[i][1] * v;
      acc[3] += L::E[i][2] * v;
    }
  }
  __shared__ double red[4][kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) red[k][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0;
    for (int w = 0; w < kBlock / 64; ++w) s += red[threadIdx.x][w];
    partials[4 * blockIdx.x + threadIdx.x] = s;
  }
}

// Sums the per-block FP64 partials (4 per block) in a fixed order: thread
// t takes blocks t, t + 1024, ..., then a fixed tree -- the same result on
// every run, and 1024 loads in flight instead of one chain.
constexpr int kSumThreads = 1024;

__global__ void __launch_bounds__(kSumThreads)
    k_sum_partials(const double* __restrict__ partials, int nblocks,
                   double* __restrict__ out) {
  __shared__ double red[4][kSumThreads];
  const int t = threadIdx.x;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = t; b < nblocks; b += kSumThreads)
    for (int k = 0; k < 4; ++k) a[k] += partials[4 * b + k];
  for (int k = 0; k < 4; ++k) red[k][t] = a[k];
  __syncthreads();
  for (int h = kSumThreads / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int k = 0; k < 4; ++k) red[k][t] += red[k][t + h];
    __syncthreads();
  }
  if (t < 4) out[t] = red[t][0];
}

template <int Q, int S>
__global__ void __launch_bounds__(kBlock)
    k_convert_in(Grid g, const void* __restrict__ src, int is_f64,
                 typename Store<S>::T* __restrict__ f) {
  using L = Lat<Q>;
  using St = Store<S>;
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const uint64_t k = (uint64_t)i * g.n + n;
    double v = is_f64 ? ((const double*)src)[k] : (double)((const float*)src)[k];
    if (St::kShifted) {
      // f32 inputs are shifted in f32 so that f32 -> store -> f32 is exact
      // wherever the storage can hold the deviation exactly
      if (is_f64)
        v -= L::W[i];
      else
        v = (double)(((const float*)src)[k] - (float)L::W[i]);
    }
    St::store(f + pidx<Q>(n, i), (typename St::C)v);
  }
}

template <int Q, int S>
__global__ void __launch_bounds__(kBlock)
    k_convert_out(Grid g, const typename Store<S>::T* __restrict__ f,
                  void* __restrict__ dst, int is_f64) {
  using L = Lat<Q>;
  using St = Store<S>;
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const uint64_t k = (uint64_t)i * g.n + n;
    const typename St::C v = St::load(f + pidx<Q>(n, i));
    if (is_f64) {
      double d = (double)v;
      if (St::kShifted) d += L::W[i];
      ((double*)dst)[k] = d;
    } else {
      float d = (float)v;
      if (St::kShifted) d += (float)L::W[i];
      ((float*)dst)[k] = d;
    }
  }
}

inline unsigned blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }



}  // namespace

int launch_std_step(int q, int storage, const Grid& g, const void* A, void* B,
                    const uint32_t* mask, const uint32_t* cls,
                    const StdParams& p, int rev, hipStream_t s) {
  const unsigned nb = blocks_for(g.n);
  const bool lid = p.lid_axis >= 0;
  const bool nt = streaming_stores(2 * population_elems(q, g.n) * storage_bytes(storage));
  // measured on MI355X: with streaming stores, alternating the tile order
  // between steps is fastest; without them, a fixed order is
  const int r = nt ? rev : 0;
  const dim3 block(kBlock);
  if (const int width = group_step_width(q, storage, g, nt)) {
    const dim3 gridv(xcd_run_grid((g.n / (uint32_t)width + kBlock - 1) / kBlock));
    auto go = [&](auto tag, auto nv) {
      using Tag = decltype(tag);
      using T = typename Store<Tag::S>::T;
      constexpr int V = decltype(nv)::value;
      const T* a = (const T*)A;
      T* b = (T*)B;
      if (lid && nt)
        hipLaunchKernelGGL((k_std_stepv<9, Tag::S, V, true, true>), gridv, block, 0, s, g, a, b,
                           mask, cls, p, r);
      else if (lid)
        hipLaunchKernelGGL((k_std_stepv<9, Tag::S, V, true, false>), gridv, block, 0, s, g, a, b,
                           mask, cls, p, r);
      else if (nt)
        hipLaunchKernelGGL((k_std_stepv<9, Tag::S, V, false, true>), gridv, block, 0, s, g, a, b,
                           mask, cls, p, r);
      else
        hipLaunchKernelGGL((k_std_stepv<9, Tag::S, V, false, false>), gridv, block, 0, s, g, a,
                           b, mask, cls, p, r);
    };
    using Two = std::integral_constant<int, 2>;
    using Four = std::integral_constant<int, 4>;
    if (storage == LBM_STORE_F16)
      width == 4 ? go(QS<9, LBM_STORE_F16>{}, Four{}) : go(QS<9, LBM_STORE_F16>{}, Two{});
    else
      width == 4 ? go(QS<9, LBM_STORE_BF16>{}, Four{}) : go(QS<9, LBM_STORE_BF16>{}, Two{});
    LBM_HIP(hipGetLastError());
    return LBM_OK;
  }
  const int drc = dispatch_qs(q, storage, [&](auto tag) {
    using Tag = decltype(tag);
    using T = typename Store<Tag::S>::T;
    const T* a = (const T*)A;
    T* b = (T*)B;
    const dim3 grid(kXcdRuns<Tag::Q, Tag::S> ? xcd_run_grid(nb) : nb);
    if (lid && nt)
      hipLaunchKernelGGL((k_std_step<Tag::Q, Tag::S, true, true>), grid, block, 0, s, g, a, b,
                         mask, cls, p, r);
    else if (lid)
      hipLaunchKernelGGL((k_std_step<Tag::Q, Tag::S, true, false>), grid, block, 0, s, g, a, b,
                         mask, cls, p, r);
    else if (nt)
      hipLaunchKernelGGL((k_std_step<Tag::Q, Tag::S, false, true>), grid, block, 0, s, g, a, b,
                         mask, cls, p, r);
    else
      hipLaunchKernelGGL((k_std_step<Tag::Q, Tag::S, false, false>), grid, block, 0, s, g, a, b,
                         mask, cls, p, r);
  });
  if (drc) return drc;
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_build_mask(int q, const Grid& g, const uint8_t* solid,
                      const int periodic[3], int lid_face, uint32_t* mask,
                      uint32_t* cls, hipStream_t s) {
  const unsigned nb = blocks_for(g.n);
  switch (q) {
    case 9:
      hipLaunchKernelGGL(k_build_mask<9>, dim3(nb), dim3(kBlock), 0, s, g,
                         solid, periodic[0], periodic[1], periodic[2],
                         lid_face, mask);
      break;
    case 19:
      hipLaunchKernelGGL(k_build_mask<19>, dim3(nb), dim3(kBlock), 0, s, g,
                         solid, periodic[0], periodic[1], periodic[2],
                         lid_face, mask);
      break;
    case 27:
      hipLaunchKernelGGL(k_build_mask<27>, dim3(nb), dim3(kBlock), 0, s, g,
                         solid, periodic[0], periodic[1], periodic[2],
                         lid_face, mask);
      break;
    default:
      return fail(LBM_ERR_UNSUPPORTED, "lattice not built");
  }
  LBM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_build_class, dim3(blocks_for((g.n + 15) / 16)),
                     dim3(kBlock), 0, s, g, mask, cls);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_mask16(const Grid& g, uint32_t* mask, hipStream_t s) {
  hipLaunchKernelGGL(k_mask16, dim3(blocks_for(g.n)), dim3(kBlock), 0, s, g, mask);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_init_eq(int q, int storage, const Grid& g, const double* rho,
                   const double* u, void* f, hipStream_t s) {
  const unsigned nb = blocks_for(g.n);
  const int drc = dispatch_qs(q, storage, [&](auto tag) {
    using Tag = decltype(tag);
    using T = typename Store<Tag::S>::T;
    hipLaunchKernelGGL((k_init_eq<Tag::Q, Tag::S>), dim3(nb), dim3(kBlock), 0, s, g,
                       rho, u, (T*)f);
  });
  if (drc) return drc;
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_std_moments(int q, int storage, const Grid& g, const void* f,
                       const uint32_t* mask, const float force[3], float* rho,
                       float* u, hipStream_t s) {
  const unsigned nb = blocks_for(g.n);
  const int drc = dispatch_qs(q, storage, [&](auto tag) {
    using Tag = decltype(tag);
    using T = typename Store<Tag::S>::T;
    hipLaunchKernelGGL((k_std_moments<Tag::Q, Tag::S>), dim3(nb), dim3(kBlock), 0, s,
                       g, (const T*)f, mask, force[0], force[1], force[2],
                       rho, u);
  });
  if (drc) return drc;
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_conservation(int q, int storage, const Grid& g, const void* f,
                        const uint32_t* mask, double* partials, int nblocks,
                        double* out4, hipStream_t s) {
  const int drc = dispatch_qs(q, storage, [&](auto tag) {
    using Tag = decltype(tag);
    using T = typename Store<Tag::S>::T;
    hipLaunchKernelGGL((k_conservation<Tag::Q, Tag::S>), dim3(nblocks), dim3(kBlock),
                       0, s, g, (const T*)f, mask, partials);
  });
  if (drc) return drc;
  LBM_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kSumThreads), 0, s, partials,
                     nblocks, out4);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_convert_in(int q, int storage, const Grid& g, const void* src,
                      int src_is_f64, void* f, hipStream_t s) {
  const unsigned nb = blocks_for(g.n);
  const int drc = dispatch_qs(q, storage, [&](auto tag) {
    using Tag = decltype(tag);
    using T = typename Store<Tag::S>::T;
    hipLaunchKernelGGL((k_convert_in<Tag::Q, Tag::S>), dim3(nb), dim3(kBlock), 0, s,
                       g, src, src_is_f64, (T*)f);
  });
  if (drc) return drc;
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_convert_out(int q, int storage, const Grid& g, const void* f,
                       void* dst, int dst_is_f64, hipStream_t s) {
  const unsigned nb = blocks_for(g.n);
  const int drc = dispatch_qs(q, storage, [&](auto tag) {
    using Tag = decltype(tag);
    using T = typename Store<Tag::S>::T;
    hipLaunchKernelGGL((k_convert_out<Tag::Q, Tag::S>), dim3(nb), dim3(kBlock), 0, s,
                       g, (const T*)f, dst, dst_is_f64);
  });
  if (drc) return drc;
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

}  // namespace lbm
