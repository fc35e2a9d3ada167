// Batched kernels: per step, one persistent launch per kernel class
// (lattice type) updates every resident lattice of that class
// (a batch mixes D2Q9 / D3Q19 / D3Q27 with one storage type), and a
// residual kernel measures |u - u_prev| / |u| per lattice.
#include <hip/hip_runtime.h>

#include "lbm_batch.h"
#include "lbm_lattice.h"
#include "lbm_layout.h"
#include "lbm_step_node.h"
#include "lbm_store.h"

namespace lbm {
namespace {

// Workgroups walk their class's tile table with a grid stride; a tile is
// blockDim.x consecutive nodes of one lattice, so the descriptor lookup is
// uniform across the workgroup.
constexpr int kMaxBatchBlock = 512;
constexpr uint32_t kXcdRun = 8;

// Eight waves per SIMD: the compiler then keeps the kernel within 80
// scalar registers (spilling a few to vector-register lanes), which on
// MI355X is what lets eight waves stay resident -- the D3Q19 kernels used
// 105-106 and ran a D3Q19 FP16 batch 3 % slower.  Not for FP64 storage,
// whose D3Q19 / D3Q27 kernels would spill vector registers to memory.
template <int Q, int S, bool NT>
__global__ void __launch_bounds__(kMaxBatchBlock)
    __attribute__((amdgpu_waves_per_eu(S == LBM_STORE_F64 ? 1 : 8)))
    k_batch_step(const BatchLattice* __restrict__ lats,
                 const BatchTile* __restrict__ tiles,
                 const uint32_t* __restrict__ ntiles, int parity) {
  using T = typename Store<S>::T;
  const uint32_t nt = *ntiles;
  // Workgroup w runs on XCD w % 8; the walk hands each XCD runs of kXcdRun
  // consecutive tiles (a permutation of the table padded to a multiple of
  // 8 kXcdRun), so a tile's row neighbours are mostly read through the
  // same L2: serving batches 2-2.5 % faster than tiles dealt round-robin.
  const uint32_t ntr = (nt + 8u * kXcdRun - 1u) / (8u * kXcdRun) * (8u * kXcdRun);
  for (uint32_t w = blockIdx.x; w < ntr; w += gridDim.x) {
    const uint32_t j = w / 8u;
    const uint32_t t = (j / kXcdRun) * (8u * kXcdRun) + (w % 8u) * kXcdRun + j % kXcdRun;
    if (t >= nt) continue;
    const BatchTile tile = tiles[t];
    const BatchLattice& L = lats[tile.lat];
    const uint32_t n = tile.first + threadIdx.x;
    if (n < L.g.n) {
      const int cur = (L.off + parity) & 1;
      // the moving-wall code only for lattices that have one (uniform per
      // tile)
      if (L.p.lid_axis >= 0)
        std_node_step<Q, S, true, NT>(L.g, (const T*)L.f[cur], (T*)L.f[cur ^ 1], L.mask, L.cls,
                                      L.p, n);
      else
        std_node_step<Q, S, false, NT>(L.g, (const T*)L.f[cur], (T*)L.f[cur ^ 1], L.mask, L.cls,
                                       L.p, n);
    }
  }
}

__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  __syncthreads();
  return s;
}

// Maximum of non-negative values over the workgroup (thread 0's result).
__device__ __forceinline__ double block_max(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s = fmax(s, red[w]);
  __syncthreads();
  return s;
}

// Velocity of the current state (u = (sum f e - F/2)/rho, fluid nodes),
// compared with the snapshot, which it then replaces.  Gives the node's
// |u - u_prev|^2 and |u|^2 for the per-chunk sums; a non-finite velocity
// makes both non-finite, so the sums carry it.
template <int Q, int S>
__device__ __forceinline__ void residual_node(const BatchLattice& L, uint32_t n,
                                              double& d2, double& u2) {
  using Lt = Lat<Q>;
  using St = Store<S>;
  using C = typename St::C;
  if (node_class(L.cls, n) == kClassSolid) return;
  const typename St::T* f = (const typename St::T*)L.f[L.off & 1];
  C r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    C v = St::load(f + pidx<Q>(n, i));
    if (St::kShifted) v += (C)Lt::W[i];
    r += v;
    jx += (C)Lt::E[i][0] * v;
    jy += (C)Lt::E[i][1] * v;
    jz += (C)Lt::E[i][2] * v;
  }
  const double ux = (double)((jx - (C)0.5 * (C)L.p.force[0]) / r);
  const double uy = (double)((jy - (C)0.5 * (C)L.p.force[1]) / r);
  const double uz = (double)((jz - (C)0.5 * (C)L.p.force[2]) / r);
  float* up = L.u_prev;
  const uint64_t N = L.g.n;
  const double dx = ux - gload(up + n), dy = uy - gload(up + N + n),
               dz = uz - gload(up + 2 * N + n);
  d2 = dx * dx + dy * dy + dz * dz;
  u2 = ux * ux + uy * uy + uz * uz;
  gstore(up + n, (float)ux);
  gstore(up + N + n, (float)uy);
  gstore(up + 2 * N + n, (float)uz);
}

template <int S>
__global__ void __launch_bounds__(kResidualChunk)
    k_batch_residual(const BatchLattice* __restrict__ lats,
                     const BatchTile* __restrict__ chunks, uint32_t nchunks,
                     double* __restrict__ partials) {
  __shared__ double red[kResidualChunk / 64];
  const uint32_t c = blockIdx.x;
  if (c >= nchunks) return;
  const BatchTile ch = chunks[c];
  const BatchLattice& L = lats[ch.lat];
  const uint32_t n = ch.first + threadIdx.x;
  double d2 = 0.0, u2 = 0.0;
  if (n < L.g.n && L.u_prev) {
    switch (L.q) {
      case 9:
        residual_node<9, S>(L, n, d2, u2);
        break;
      case 19:
        residual_node<19, S>(L, n, d2, u2);
        break;
      default:
        residual_node<27, S>(L, n, d2, u2);
        break;
    }
  }
  // largest |u|^2 of the chunk, +inf for a non-finite velocity (fmax
  // would drop a NaN)
  const double um = u2 <= 1e300 ? u2 : (double)INFINITY;
  const double sd = block_sum(d2, red);
  const double su = block_sum(u2, red);
  const double sm = block_max(um, red);
  if (threadIdx.x == 0) {
    partials[kResidualPartials * c] = sd;
    partials[kResidualPartials * c + 1] = su;
    partials[kResidualPartials * c + 2] = sm;
  }
}

template <int Q, int S>
int launch_step_q(bool nt, int block, int grid, const BatchLattice* lats, const BatchTile* tiles,
                  const uint32_t* ntiles, int parity, hipStream_t s) {
  const dim3 g((unsigned)grid), b((unsigned)block);
  if (nt)
    hipLaunchKernelGGL((k_batch_step<Q, S, true>), g, b, 0, s, lats, tiles, ntiles, parity);
  else
    hipLaunchKernelGGL((k_batch_step<Q, S, false>), g, b, 0, s, lats, tiles, ntiles, parity);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

}  // namespace

int launch_batch_step(int storage, int cls, bool nt, int block, int grid,
                      const BatchLattice* lats, const BatchTile* tiles,
                      const uint32_t* ntiles, int parity, hipStream_t s) {
  if (grid <= 0) return fail(LBM_ERR_INVALID, "batch grid must be positive");
  if (block != 64 && block != 128 && block != 256 && block != kMaxBatchBlock)
    return fail(LBM_ERR_INVALID, "batch block size must be 64, 128, 256 or 512");
  if (cls < 0 || cls >= kBatchClasses) return fail(LBM_ERR_INVALID, "bad kernel class");
  int rc = LBM_OK;
  const int drc = dispatch_qs(class_q(cls), storage, [&](auto tag) {
    using Tag = decltype(tag);
    rc = launch_step_q<Tag::Q, Tag::S>(nt, block, grid, lats, tiles, ntiles, parity, s);
  });
  return drc ? drc : rc;
}

int launch_batch_residual(int storage, const BatchLattice* lats,
                          const BatchTile* chunks, uint32_t nchunks,
                          double* partials, hipStream_t s) {
  if (nchunks == 0) return LBM_OK;
  const dim3 g(nchunks), b(kResidualChunk);
  switch (storage) {
    case LBM_STORE_F32:
      hipLaunchKernelGGL(k_batch_residual<LBM_STORE_F32>, g, b, 0, s, lats, chunks, nchunks, partials);
      break;
    case LBM_STORE_F64:
      hipLaunchKernelGGL(k_batch_residual<LBM_STORE_F64>, g, b, 0, s, lats, chunks, nchunks, partials);
      break;
    case LBM_STORE_F16:
      hipLaunchKernelGGL(k_batch_residual<LBM_STORE_F16>, g, b, 0, s, lats, chunks, nchunks, partials);
      break;
    case LBM_STORE_BF16:
      hipLaunchKernelGGL(k_batch_residual<LBM_STORE_BF16>, g, b, 0, s, lats, chunks, nchunks, partials);
      break;
    case kStoreF32Shifted:
      hipLaunchKernelGGL(k_batch_residual<kStoreF32Shifted>, g, b, 0, s, lats, chunks, nchunks,
                         partials);
      break;
    default:
      return fail(LBM_ERR_UNSUPPORTED, "unknown storage");
  }
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

}  // namespace lbm
