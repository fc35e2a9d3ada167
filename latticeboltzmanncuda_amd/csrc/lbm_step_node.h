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
//      face) -- one load per link, the address is selected without a
//      branch, so all Q loads are in flight together; solid nodes load
//      nothing;
//   3. collide (BGK + Guo forcing) in FP32 (FP64 for F64 storage);
//   4. store every direction with all lanes active; solid nodes store the
//      rest state (w_i, or 0 for shifted storage).  Skipping the stores of
//      solid lanes leaves partly written cache lines that cost ~1.7x in
//      time on MI355X.  NT selects non-temporal stores (lattices larger
//      than the Infinity Cache).
//
// Shifted storage (FP32, FP16, BF16 in standard mode) holds
// g_i = f_i - w_i and is collided in that form from small terms only
// (drho = sum g_i, j = sum e_i g_i; see std_relax), so deviations and a
// body force's small increments are not cancelled against w_i or 1.
//
// The FP16 / BF16 paths are bound by vector-instruction issue, not by HBM,
// so the update is written to spend few instructions per node: per-lattice
// constants come precomputed (StepCoef), addresses need no 32-bit integer
// multiplications, the body-force terms run only for lattices that have
// one, and solid lanes get no values.
#pragma once

#include <hip/hip_runtime.h>

#include "lbm_global.h"
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
  y = t - __umul24(x, g.ny);  // x < 2^24 and ny < 2^24 (lbm_create checks)
}

template <class C>
__device__ __forceinline__ const StepCoef<C>& step_coef(const StdParams& p);
template <>
__device__ __forceinline__ const StepCoef<float>& step_coef<float>(const StdParams& p) {
  return p.cf;
}
template <>
__device__ __forceinline__ const StepCoef<double>& step_coef<double>(const StdParams& p) {
  return p.cd;
}

// 1 / r: IEEE division in FP64; in FP32 the hardware reciprocal refined by
// one Newton-Raphson step (error below one unit in the last place)
__device__ __forceinline__ double recip(double r) { return 1.0 / r; }
__device__ __forceinline__ float recip(float r) {
  const float y = __builtin_amdgcn_rcpf(r);
  return fmadd(fmadd(-r, y, 1.0f), y, y);
}

// Loads of one node: the Q pulled (or bounced) stored values, the link
// mask and the node's own element index.
template <int Q, int S>
struct NodeLoad {
  typename Store<S>::T raw[Q];
  uint32_t m;    // link mask (0 for interior and solid nodes)
  uint32_t own;  // element index of (n, direction 0)
  bool solid;
};

// Phase 1: issue every load of node n (no value is used here, so the loads
// of several nodes can be in flight together).
template <int Q, int S>
__device__ __forceinline__ void std_gather(const Grid& g,
                                           const typename Store<S>::T* __restrict__ A,
                                           const uint32_t* __restrict__ mask,
                                           const uint32_t* __restrict__ cls, uint32_t n,
                                           NodeLoad<Q, S>& ld) {
  using L = Lat<Q>;
  const uint32_t c = node_class(cls, n);
  uint32_t x, y, z;
  if (Q == 9) {  // D2Q9 lattices have nz == 1
    z = 0;
    x = g.dy.div(n);
    y = n - __umul24(x, g.ny);
  } else {
    decode_node(g, n, x, y, z);
  }
  const uint32_t plane = g.ny * g.nz;
  // offsets (mod 2^32) from n to the source node n - e for e = -1, 0, +1
  // along each axis, with periodic wrap (a wall link is flagged in the mask
  // and its wrapped value is replaced)
  const uint32_t ox[3] = {x + 1 == g.nx ? plane - g.n : plane, 0u,
                          x == 0 ? g.n - plane : 0u - plane};
  const uint32_t oy[3] = {y + 1 == g.ny ? g.nz - plane : g.nz, 0u,
                          y == 0 ? plane - g.nz : 0u - g.nz};
  const uint32_t oz[3] = {z + 1 == g.nz ? 1u - g.nz : 1u, 0u,
                          z == 0 ? g.nz - 1u : 0u - 1u};

  ld.solid = c == kClassSolid;
  ld.m = c == kClassBoundary ? gload(mask + n) : 0u;
  ld.own = pidx<Q>(n, 0);
  // solid lanes load nothing: their values are left unspecified (no
  // instruction spent on them), the collision replaces their outputs with
  // the rest state
#pragma unroll
  for (int i = 0; i < Q; ++i) ld.raw[i] = __builtin_nondeterministic_value(ld.raw[i]);
  if (!ld.solid) {
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const uint32_t src = n + ox[L::E[i][0] + 1] + oy[L::E[i][1] + 1] + oz[L::E[i][2] + 1];
      // pidx(src, i) as one 24-bit multiply-add: lane | i * kTile < Q * kTile
      const uint32_t from =
          __umul24(src / kTile, (uint32_t)(Q * kTile)) + ((src % kTile) | ((uint32_t)i * kTile));
      const uint32_t back = ld.own + (uint32_t)L::INV[i] * kTile;
      // bounced link: read the node's own opposite population instead
      const uint32_t sel = 0u - ((ld.m >> i) & 1u);
      ld.raw[i] = gload(elem(A, from ^ ((from ^ back) & sel)));
    }
  }
}

// BGK relaxation towards the equilibrium of (rho, u), plus the Guo source
// term when FORCED, and the stores of the Q outputs (solid nodes: the rest
// state).  rs is rho (full storage) or drho = rho - 1 (shifted storage).
// Shifted storage forms feq - f = w rho t - w - g as
//   w ((t - 1) + drho + drho (t - 1)) - g
// from small terms only: no intermediate is near 1 or near w, so the
// rounding error of the collision scales with the deviations and a body
// force's small per-step increments are not lost.
template <int Q, int S, bool NT, bool FORCED>
__device__ __forceinline__ void std_relax(const StepCoef<typename Store<S>::C>& k,
                                          const typename Store<S>::C (&gv)[Q],
                                          typename Store<S>::C rs, typename Store<S>::C ux,
                                          typename Store<S>::C uy, typename Store<S>::C uz,
                                          typename Store<S>::C m3uF, bool solid,
                                          typename Store<S>::T* out0, uint32_t node_hash) {
#pragma clang fp contract(on)
  using L = Lat<Q>;
  using St = Store<S>;
  using C = typename St::C;
  constexpr bool kSh = St::kShifted;
  const C usq = ux * ux + uy * uy + uz * uz;
  // t - 1 (shifted) or t = 1 - 1.5 u.u + 3 e.u + 4.5 (e.u)^2 (full)
  const C base = kSh ? (C)-1.5 * usq : (C)1 - (C)1.5 * usq;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    // e.u from the nonzero components only (-0 + v == v exactly)
    C eu = (C)-0.0;
    if (L::E[i][0]) eu = L::E[i][0] > 0 ? eu + ux : eu - ux;
    if (L::E[i][1]) eu = L::E[i][1] > 0 ? eu + uy : eu - uy;
    if (L::E[i][2]) eu = L::E[i][2] > 0 ? eu + uz : eu - uz;
    const C w = (C)L::W[i];
    const C t = fmadd(fmadd((C)4.5, eu, (C)3), eu, base);
    // feq - f in the stored form: w ((t-1) + drho + drho (t-1)) - g
    // (shifted, t here is t - 1), w rho t - f (full)
    const C d = kSh ? fmadd(w, fmadd(rs, t, t + rs), -gv[i]) : (w * rs) * t - gv[i];
    C out = fmadd(k.omega, d, gv[i]);
    if (FORCED) {
      // Guo source (1 - omega/2) w (3 (e.F - u.F) + 9 (e.u)(e.F)), added
      // term by term (each step takes one per-lattice constant)
      out = out + k.gf3[i];
      out = fmadd(k.gw[i], m3uF, out);
      out = fmadd(k.gf9[i], eu, out);
    }
    out = solid ? (kSh ? (C)0 : w) : out;
    // a body force moves 16-bit populations by less than their rounding
    // step: those lattices round stochastically (no effect at full precision)
    if (FORCED && NT)
      St::store_sr_nt(out0 + i * kTile, out, sr_key(node_hash, i));
    else if (FORCED)
      St::store_sr(out0 + i * kTile, out, sr_key(node_hash, i));
    else if (NT)
      St::store_nt(out0 + i * kTile, out);
    else
      St::store(out0 + i * kTile, out);
  }
}

// Phase 2: collide the loaded node and store its Q post-collision values.
template <int Q, int S, bool LID, bool NT>
__device__ __forceinline__ void std_collide(const StdParams& p,
                                            typename Store<S>::T* __restrict__ B,
                                            const NodeLoad<Q, S>& ld) {
  // a * b + c is fused only where one source expression has that form, so
  // every instantiation (single-lattice or batch, any block size, with or
  // without a moving wall) rounds identically
#pragma clang fp contract(on)
  using L = Lat<Q>;
  using St = Store<S>;
  using C = typename St::C;
  constexpr bool kSh = St::kShifted;
  const StepCoef<C>& k = step_coef<C>(p);
  const uint32_t m = ld.m;
  // g[i]: the stored quantity (f_i, or f_i - w_i for shifted storage), plus
  // the moving-wall term of bounced links from beyond the moving face
  C gv[Q];
  const bool lidnode = LID && (m & LBM_MASK_LID);
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    C v = St::cvt(ld.raw[i]);
    if (LID && i > 0) {  // the rest direction never crosses a face
      const bool wall = lidnode && ((m >> i) & 1u);
      v += wall ? k.lid[i] : (C)0;
    }
    gv[i] = v;
  }

  // moments; with shifted storage sum w_i = 1 and sum w_i e_i = 0, so the
  // sum of the stored values is rho itself (full storage) or the density
  // deviation drho = rho - 1 (shifted storage)
  C rs = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    rs += gv[i];
    if (L::E[i][0]) jx += L::E[i][0] > 0 ? gv[i] : -gv[i];
    if (L::E[i][1]) jy += L::E[i][1] > 0 ? gv[i] : -gv[i];
    if (L::E[i][2]) jz += L::E[i][2] > 0 ? gv[i] : -gv[i];
  }
  const C ir = recip(kSh ? (C)1 + rs : rs);
  // the node's Q outputs lie kTile elements apart: one 64-bit address, the
  // direction offsets become constant offsets of the stores
  typename St::T* const out0 = B + ld.own;
  // Guo forcing only where the lattice has a body force (uniform per
  // lattice); without one, u = j / rho and the source term is zero
  if (k.forced) {
    const C ux = (jx + k.half_force[0]) * ir;
    const C uy = (jy + k.half_force[1]) * ir;
    const C uz = (jz + k.half_force[2]) * ir;
    const C m3uF = (C)-3 * (ux * k.force[0] + uy * k.force[1] + uz * k.force[2]);
    std_relax<Q, S, NT, true>(k, gv, rs, ux, uy, uz, m3uF, ld.solid, out0,
                              kSh ? sr_node_hash(ld.own) : 0u);
  } else {
    std_relax<Q, S, NT, false>(k, gv, rs, jx * ir, jy * ir, jz * ir, (C)0, ld.solid, out0, 0u);
  }
}

template <int Q, int S, bool LID, bool NT = false>
__device__ __forceinline__ void std_node_step(
    const Grid& g, const typename Store<S>::T* __restrict__ A,
    typename Store<S>::T* __restrict__ B, const uint32_t* __restrict__ mask,
    const uint32_t* __restrict__ cls, const StdParams& p, uint32_t n) {
  NodeLoad<Q, S> ld;
  std_gather<Q, S>(g, A, mask, cls, n, ld);
  std_collide<Q, S, LID, NT>(p, B, ld);
}

}  // namespace lbm
