// Batched kernels: one persistent launch updates every resident lattice of
// a batch (any mix of D2Q9 / D3Q19 / D3Q27 with one storage type), and a
// residual kernel measures |u - u_prev| / |u| per lattice.
#include <hip/hip_runtime.h>

#include "lbm_batch.h"
#include "lbm_lattice.h"
#include "lbm_layout.h"
#include "lbm_step_node.h"
#include "lbm_store.h"

namespace lbm {
namespace {

// Workgroups walk the tile table with a grid stride; every tile belongs to
// one lattice, so the descriptor lookup is uniform across the workgroup.
template <int S, int BS, bool NT>
__global__ void __launch_bounds__(BS)
    k_batch_step(const BatchLattice* __restrict__ lats,
                 const BatchTile* __restrict__ tiles,
                 const uint32_t* __restrict__ ntiles, int parity) {
  using T = typename Store<S>::T;
  const uint32_t nt = *ntiles;
  for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
    const BatchTile tile = tiles[t];
    const BatchLattice& L = lats[tile.lat];
    const uint32_t n = tile.first + threadIdx.x;
    if (n < L.g.n) {
      const int cur = (L.off + parity) & 1;
      const T* A = (const T*)L.f[cur];
      T* B = (T*)L.f[cur ^ 1];
      switch (L.q) {
        case 9:
          std_node_step<9, S, true, NT>(L.g, A, B, L.mask, L.cls, L.p, n);
          break;
        case 19:
          std_node_step<19, S, true, NT>(L.g, A, B, L.mask, L.cls, L.p, n);
          break;
        default:
          std_node_step<27, S, true, NT>(L.g, A, B, L.mask, L.cls, L.p, n);
          break;
      }
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

// Velocity of the current state (u = (sum f e - F/2)/rho, fluid nodes),
// compared with the snapshot, which it then replaces.  Writes per-chunk
// sums of |u - u_prev|^2 and |u|^2.
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
  const double dx = ux - up[n], dy = uy - up[N + n], dz = uz - up[2 * N + n];
  d2 = dx * dx + dy * dy + dz * dz;
  u2 = ux * ux + uy * uy + uz * uz;
  up[n] = (float)ux;
  up[N + n] = (float)uy;
  up[2 * N + n] = (float)uz;
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
  const double sd = block_sum(d2, red);
  const double su = block_sum(u2, red);
  if (threadIdx.x == 0) {
    partials[2 * c] = sd;
    partials[2 * c + 1] = su;
  }
}

template <int S, bool NT>
int launch_step_s(int block, int grid, const BatchLattice* lats,
                  const BatchTile* tiles, const uint32_t* ntiles, int parity,
                  hipStream_t s) {
  const dim3 g((unsigned)grid);
  switch (block) {
    case 64:
      hipLaunchKernelGGL((k_batch_step<S, 64, NT>), g, dim3(64), 0, s, lats, tiles, ntiles, parity);
      break;
    case 128:
      hipLaunchKernelGGL((k_batch_step<S, 128, NT>), g, dim3(128), 0, s, lats, tiles, ntiles, parity);
      break;
    case 256:
      hipLaunchKernelGGL((k_batch_step<S, 256, NT>), g, dim3(256), 0, s, lats, tiles, ntiles, parity);
      break;
    case 512:
      hipLaunchKernelGGL((k_batch_step<S, 512, NT>), g, dim3(512), 0, s, lats, tiles, ntiles, parity);
      break;
    default:
      return fail(LBM_ERR_INVALID, "batch block size must be 64, 128, 256 or 512");
  }
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

template <int S>
int launch_step_nt(bool nt, int block, int grid, const BatchLattice* lats,
                   const BatchTile* tiles, const uint32_t* ntiles, int parity,
                   hipStream_t s) {
  return nt ? launch_step_s<S, true>(block, grid, lats, tiles, ntiles, parity, s)
            : launch_step_s<S, false>(block, grid, lats, tiles, ntiles, parity, s);
}

}  // namespace

int launch_batch_step(int storage, bool nt, int block, int grid,
                      const BatchLattice* lats, const BatchTile* tiles,
                      const uint32_t* ntiles, int parity, hipStream_t s) {
  if (grid <= 0) return fail(LBM_ERR_INVALID, "batch grid must be positive");
  switch (storage) {
    case LBM_STORE_F32:
      return launch_step_nt<LBM_STORE_F32>(nt, block, grid, lats, tiles, ntiles, parity, s);
    case LBM_STORE_F64:
      return launch_step_nt<LBM_STORE_F64>(nt, block, grid, lats, tiles, ntiles, parity, s);
    case LBM_STORE_F16:
      return launch_step_nt<LBM_STORE_F16>(nt, block, grid, lats, tiles, ntiles, parity, s);
    case LBM_STORE_BF16:
      return launch_step_nt<LBM_STORE_BF16>(nt, block, grid, lats, tiles, ntiles, parity, s);
    default:
      return fail(LBM_ERR_UNSUPPORTED, "unknown storage");
  }
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
    default:
      return fail(LBM_ERR_UNSUPPORTED, "unknown storage");
  }
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

}  // namespace lbm
