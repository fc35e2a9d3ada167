// The standard-mode update of one node, shared by the single-lattice step
// kernel (k_standard.hip) and the batched persistent kernel (k_batch.hip).
//
// For node n of lattice g:
//   1. read its class from the 2-bit class map: interior fluid, boundary
//      fluid, or solid;
//   2. fluid nodes pull f_i = A_i(x - e_i) with periodic wrap; boundary
//      nodes first read their link mask and, for each bounced link, load
//      the half-way bounce-back value A_inv(i)(x) instead (+ the moving-wall
//      term 6 w_i rho_w e_i.u_w when the source lies beyond the moving
//      face) -- one load per link, the address is selected, so no solid
//      data is fetched; solid nodes load nothing;
//   3. collide (BGK + Guo forcing) in FP32 (FP64 for F64 storage);
//   4. store every direction with all lanes active; solid nodes store the
//      rest state w_i.  Skipping the stores of solid lanes leaves partly
//      written cache lines that cost ~1.7x in time on MI355X.  NT selects
//      non-temporal stores (lattices larger than the Infinity Cache).
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_internal.h"
#include "lbm_lattice.h"
#include "lbm_layout.h"
#include "lbm_store.h"

namespace lbm {

__device__ __forceinline__ void decode_node(const Grid& g, uint32_t n,
                                            uint32_t& x, uint32_t& y,
                                            uint32_t& z) {
  const uint32_t t = g.dz.div(n);
  z = n - t * g.nz;
  x = g.dy.div(t);
  y = t - x * g.ny;
}

template <int Q, int S, bool LID, bool NT = false>
__device__ __forceinline__ void std_node_step(
    const Grid& g, const typename Store<S>::T* __restrict__ A,
    typename Store<S>::T* __restrict__ B, const uint32_t* __restrict__ mask,
    const uint32_t* __restrict__ cls, const StdParams& p, uint32_t n) {
  using L = Lat<Q>;
  using St = Store<S>;
  using C = typename St::C;
  const uint32_t c = node_class(cls, n);
  uint32_t x, y, z;
  decode_node(g, n, x, y, z);
  const uint32_t plane = g.ny * g.nz;
  // source coordinate c - e for e = -1, 0, +1 (periodic wrap; a wall link
  // is flagged in the mask and its wrapped value is replaced)
  const uint32_t sx[3] = {(x + 1 == g.nx ? 0u : x + 1) * plane, x * plane,
                          (x == 0 ? g.nx : x) * plane - plane};
  const uint32_t sy[3] = {(y + 1 == g.ny ? 0u : y + 1) * g.nz, y * g.nz,
                          (y == 0 ? g.ny : y) * g.nz - g.nz};
  const uint32_t sz[3] = {z + 1 == g.nz ? 0u : z + 1, z,
                          (z == 0 ? g.nz : z) - 1};

  const bool solid = c == kClassSolid;
  const uint32_t m = c == kClassBoundary ? mask[n] : 0u;
  C f[Q];
  if (!solid) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const uint32_t src = sx[L::E[i][0] + 1] + sy[L::E[i][1] + 1] + sz[L::E[i][2] + 1];
      const bool bb = (m >> i) & 1u;
      C v = St::load(A + (bb ? pidx<Q>(n, L::INV[i]) : pidx<Q>(src, i)));
      if (LID) {
        const int ex = L::E[i][0], ey = L::E[i][1], ez = L::E[i][2];
        const int el = p.lid_axis == 0 ? ex : (p.lid_axis == 1 ? ey : ez);
        if (bb && (m & LBM_MASK_LID) && el == -p.lid_sign) {
          const C eu = (C)ex * (C)p.lid_u[0] + (C)ey * (C)p.lid_u[1] +
                       (C)ez * (C)p.lid_u[2];
          // shifted storage: the wall term is added to f, not to f - w
          v += (C)6 * (C)L::W[i] * (C)p.wall_rho * eu;
        }
      }
      f[i] = v;
    }
  } else {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = St::kShifted ? (C)0 : (C)L::W[i];
  }
  if (St::kShifted) {
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] += (C)L::W[i];
  }

  C rho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    rho += f[i];
    if (L::E[i][0]) jx += L::E[i][0] > 0 ? f[i] : -f[i];
    if (L::E[i][1]) jy += L::E[i][1] > 0 ? f[i] : -f[i];
    if (L::E[i][2]) jz += L::E[i][2] > 0 ? f[i] : -f[i];
  }
  const C omega = (C)1 / (C)p.tau;
  const C guo = (C)1 - (C)0.5 * omega;
  const C Fx = (C)p.force[0], Fy = (C)p.force[1], Fz = (C)p.force[2];
  const C ir = (C)1 / rho;
  const C ux = (jx + (C)0.5 * Fx) * ir;
  const C uy = (jy + (C)0.5 * Fy) * ir;
  const C uz = (jz + (C)0.5 * Fz) * ir;
  const C usq = ux * ux + uy * uy + uz * uz;
  const C uF = ux * Fx + uy * Fy + uz * Fz;
  const C base = (C)1 - (C)1.5 * usq;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const C ex = (C)L::E[i][0], ey = (C)L::E[i][1], ez = (C)L::E[i][2];
    const C eu = ex * ux + ey * uy + ez * uz;
    const C eF = ex * Fx + ey * Fy + ez * Fz;
    const C w = (C)L::W[i];
    const C feq = w * rho * (base + (C)3 * eu + (C)4.5 * eu * eu);
    const C src = guo * w * ((C)3 * (eF - uF) + (C)9 * eu * eF);
    C out = f[i] + omega * (feq - f[i]) + src;
    out = solid ? w : out;
    if (St::kShifted) out -= w;
    if (NT)
      St::store_nt(B + pidx<Q>(n, i), out);
    else
      St::store(B + pidx<Q>(n, i), out);
  }
}

}  // namespace lbm
