// The standard-mode update of one node, shared by the single-lattice step
// kernel (k_standard.hip) and the batched persistent kernel (k_batch.hip),
// and of a group of nodes per thread (single lattices).
//
// For node n of lattice g:
//   1. read its class from the 2-bit class map: interior fluid, boundary
//      fluid, or solid (D2Q9 other than FP64, and 3-D lattices with solid
//      nodes: its link mask, which marks solid nodes too);
//   2. fluid nodes pull f_i = A_i(x - e_i) with periodic wrap; boundary
//      nodes first read their link mask and, for each bounced link, load
//      the half-way bounce-back value A_inv(i)(x) instead (+ the moving-wall
//      term 6 w_i rho_w e_i.u_w when the source lies beyond the moving
//      face) -- one load per link, the address is selected without a
//      branch, so all Q loads are in flight together; with FP32 / FP64
//      storage solid nodes load nothing, with 16-bit storage they gather
//      like fluid nodes (fewer instructions; see std_gather_pops);
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
// constants come precomputed (StepCoef), in FP16's stored units; addresses
// need no 32-bit integer multiplications and go through buffer resources
// (constant offsets in the instructions); the body-force terms run only
// for lattices that have one; wavefronts skip the moving-wall term, the
// bounced-link select and (D2Q9) the rest-state select where none of
// their nodes needs them; 16-bit stores of forced lattices round with the
// hardware's stochastic conversion.  Large D2Q9 lattices with 16-bit
// storage take the group step (std_group_step, below): two or four nodes
// per thread share their loads, stores and address arithmetic, with the
// same per-node arithmetic, so every width gives the same bytes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>

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

// The step constants of storage type S, in its stored units.
template <int S>
__device__ __forceinline__ const StepCoef<typename Store<S>::C>& step_coef(const StdParams& p) {
  if constexpr (S == LBM_STORE_F64)
    return p.cd;
  else if constexpr (S == LBM_STORE_F16)
    return p.ch;
  else
    return p.cf;
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
  uint32_t m;          // link mask (0 for interior and solid nodes)
  uint32_t own;        // element index of (n, direction 0)
  uint32_t pop_bytes;  // bytes of one population array (16-bit storage)
  bool solid;
};

// b where bit i of m is set, else a: bit i spread over a word by a signed
// bit-field extract, then a bit-field insert -- two instructions where the
// compiler's compare-and-select takes three (the insert is written out:
// the compiler turns the equivalent and/or expression back into the select)
__device__ __forceinline__ uint32_t select_bit(uint32_t m, int i, uint32_t b, uint32_t a) {
  const uint32_t sel = (uint32_t)__builtin_amdgcn_sbfe((int)m, i, 1);
  uint32_t r;
  __asm__("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(sel), "v"(b), "v"(a));
  return r;
}

// Phase 1: issue every load of node n (no value is used here, so the loads
// of several nodes can be in flight together).
template <int Q, int S>
__device__ __forceinline__ void std_gather(const Grid& g,
                                           const typename Store<S>::T* __restrict__ A,
                                           const uint32_t* __restrict__ mask,
                                           const uint32_t* __restrict__ cls, uint32_t n,
                                           NodeLoad<Q, S>& ld) {
  using L = Lat<Q>;
  // the node's 2-bit class (read only on the branch below that uses it)
  const uint32_t c = node_class(cls, n);
  const uint32_t plane = g.ny * g.nz;
  // FP32 / FP64 storage: offsets (mod 2^32) from n to the source node
  // n - e for e = -1, 0, +1 along each axis, with periodic wrap along every
  // axis -- a wall link is flagged in the mask and its wrapped address
  // replaced -- so that every address formed stays within the lattice
  // (these loads have no range check)
  uint32_t ox[3] = {}, oy[3] = {}, oz[3] = {};
  if constexpr (sizeof(typename Store<S>::T) != 2) {
    uint32_t x, y, z;
    if (Q == 9) {  // D2Q9 lattices have nz == 1
      z = 0;
      x = g.dy.div(n);
      y = n - __umul24(x, g.ny);
    } else {
      decode_node(g, n, x, y, z);
    }
    ox[0] = x + 1 == g.nx ? plane - g.n : plane;
    ox[2] = x == 0 ? g.n - plane : 0u - plane;
    oy[0] = y + 1 == g.ny ? g.nz - plane : g.nz;
    oy[2] = y == 0 ? plane - g.nz : 0u - g.nz;
    oz[0] = z + 1 == g.nz ? 1u - g.nz : 1u;
    oz[2] = z == 0 ? g.nz - 1u : 0u - 1u;
  }

  using T = typename Store<S>::T;
  if constexpr (sizeof(T) != 8 && Q == 9) {
    // D2Q9 with FP32 or 16-bit storage: every lane reads its link mask
    // from the 16-bit copy that follows the 32-bit masks (bit 15 marks a
    // solid node, bit 14 a node next to the moving wall), instead of the
    // class word and then, for boundary nodes, the mask -- one load
    // latency fewer before the gather, for under 2 bytes per node more
    // read (C4 batches: FP16 3.5 %, FP32 2.4 % faster)
    const uint32_t mw =
        bload16<uint16_t>(buffer_rsrc(mask + g.n, 2u * g.n), 2u * n);
    ld.solid = (mw & kMask16Solid) != 0u;
    ld.m = mw & ~kMask16Solid;
  } else if (sizeof(T) != 8 && (g.flags & kReadMasks)) {
    // 3-D lattices with solid nodes (many boundary nodes): every lane
    // reads its link mask, whose top bit marks a solid node -- one load
    // latency fewer before the gather, for 3-4 bytes per node more read
    // (D3Q19 FP32 channel with 20 % spheres 1.2 % faster, D3Q19 FP16
    // batches of sphere channels 1-3 %); lattices without solids (few or
    // no boundary nodes: the Taylor-Green vortex 2 % faster this way)
    // read the 2-bit class and the mask only where needed
    const uint32_t mw = gload(mask + n);
    ld.solid = (mw & LBM_MASK_SOLID) != 0u;
    ld.m = mw & ~LBM_MASK_SOLID;
  } else {
    ld.solid = c == kClassSolid;
    ld.m = c == kClassBoundary ? gload(mask + n) : 0u;
  }
  ld.own = pidx<Q>(n, 0);
  ld.pop_bytes = sizeof(T) == 2 ? (g.n + kTile - 1u) / kTile * (uint32_t)(Q * kTile * 2) : 0u;
  // solid lanes' values are unspecified (the collision replaces their
  // outputs with the rest state).  With FP32 / FP64 storage solid lanes
  // load nothing.  With 16-bit storage (bound by instruction issue) every
  // lane loads: the lines solid lanes read are mostly those their fluid
  // neighbours read, and skipping them would cost instructions to give
  // the skipped registers a value on every wavefront
#pragma unroll
  for (int i = 0; i < Q; ++i) ld.raw[i] = __builtin_nondeterministic_value(ld.raw[i]);
  if (sizeof(T) == 2 || !ld.solid) {
    if constexpr (sizeof(T) == 2) {
      // 16-bit storage, in byte offsets (below 2^32: lbm_create keeps
      // element indices below 2^31) through a buffer resource over the
      // population array: element (m, i) lies at
      //   2 pidx(m, i) = 2 m + (2 m / 128) (Q - 1) 128 + 128 i,
      // formed from the doubled index 2 m (the offsets to the neighbours
      // doubled once per node) with a shift and one 24-bit multiply-add
      // per source node.  The part of the offset
      // that the pulled address (direction i of the source node) and the
      // bounced one (direction INV[i] of the node itself) share,
      // 128 min(i, INV[i]), is a constant of the load instruction.
      constexpr uint32_t kRow = (uint32_t)kTile * 2u;  // bytes of one direction's run
      const auto rsrc = buffer_rsrc(A, ld.pop_bytes);
      const uint32_t own_b = 2u * ld.own;
      // the doubled source indices 2 (n - e_i) mod 2^32
      const uint32_t p2 = 2u * plane, n2 = 2u * g.n, z2 = 2u * g.nz;
      uint32_t s2[Q];                                // D2Q9
      uint32_t ox2[3] = {}, oy2[3] = {}, oz2[3] = {};  // 3-D
      if constexpr (Q == 9) {
        // D2Q9 (nz == 1; batches and single lattices bound by instruction
        // issue): one add of a scalar per direction, without wrap.  Along
        // a walled axis every link that leaves the box is bounced (flagged
        // in the mask, its pulled address replaced below), and any other
        // address a lane forms there -- a solid node's -- is read within
        // the buffer resource's range or as 0, so these serve.  A periodic
        // axis then wraps the links of nodes on its end faces, in place and
        // only in wavefronts that hold such a node: sources beyond the
        // high end (e = -1) move back by the axis's period, those beyond
        // the low end (e = +1) forward.  The x ends are found by comparing
        // n with the row length; y is decoded only for lattices periodic
        // along y.
#pragma unroll
        for (int i = 0; i < Q; ++i) {
          uint32_t d = 0u;
          if (L::E[i][0]) d += L::E[i][0] > 0 ? 0u - p2 : p2;
          if (L::E[i][1]) d += L::E[i][1] > 0 ? 0u - 2u : 2u;
          s2[i] = 2u * n + d;
        }
        auto wrap_axis = [&](int a, bool at_hi, bool at_lo, uint32_t period2) {
          if (__builtin_amdgcn_ballot_w64(at_hi || at_lo)) {
            const uint32_t fh = at_hi ? period2 : 0u, fl = at_lo ? period2 : 0u;
#pragma unroll
            for (int i = 0; i < Q; ++i) {
              if (L::E[i][a] < 0) s2[i] -= fh;
              if (L::E[i][a] > 0) s2[i] += fl;
            }
          }
        };
        if (g.flags & kWrapX) wrap_axis(0, n + plane >= g.n, n < plane, n2);
        if (g.flags & kWrapY) {
          const uint32_t x = g.dy.div(n);
          const uint32_t y = n - __umul24(x, g.ny);  // x, ny < 2^24
          wrap_axis(1, y + 1 == g.ny, y == 0, p2);
        }
      } else {
        // 3-D lattices (bound by HBM bandwidth; the form above measured
        // slower for them): doubled offsets per axis, with periodic wrap,
        // selected from doubled (scalar) values by the decoded coordinates
        // and summed per direction where its load needs them (formed all
        // ahead, they hold more registers: the D3Q27 batch kernels spill)
        uint32_t x, y, z;
        decode_node(g, n, x, y, z);
        ox2[0] = x + 1 == g.nx ? p2 - n2 : p2;
        ox2[2] = x == 0 ? n2 - p2 : 0u - p2;
        oy2[0] = y + 1 == g.ny ? z2 - p2 : z2;
        oy2[2] = y == 0 ? p2 - z2 : 0u - z2;
        oz2[0] = z + 1 == g.nz ? 2u - z2 : 2u;
        oz2[2] = z == 0 ? z2 - 2u : 0u - 2u;
      }
      auto src2 = [&](int i) {
        if constexpr (Q == 9)
          return s2[i];
        else
          return 2u * n + ox2[L::E[i][0] + 1] + oy2[L::E[i][1] + 1] + oz2[L::E[i][2] + 1];
      };
      // (Q - 1) kRow passed through an empty asm: for D2Q9 it is a power of
      // two, which the compiler would turn into a shift and a mask -- one
      // instruction more than the multiply-add
      uint32_t run_bytes = (uint32_t)(Q - 1) * kRow;
      asm("" : "+s"(run_bytes));
      auto pull_b = [&](int i) {
        const uint32_t v = src2(i);
        return __umul24(v / kRow, run_bytes) + v;
      };
      // a wavefront without a bounced link among its nodes (a third of
      // them in the batching workload, more in large lattices) skips the
      // per-link selection: the branch is uniform
      if (!__any(ld.m != 0u)) {
#pragma unroll
        for (int i = 0; i < Q; ++i) ld.raw[i] = bload16<T>(rsrc, pull_b(i) + (uint32_t)i * kRow);
      } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
          const uint32_t lo = (uint32_t)(i < L::INV[i] ? i : L::INV[i]) * kRow;
          const uint32_t a = pull_b(i) + ((uint32_t)i * kRow - lo);
          const uint32_t b = own_b + ((uint32_t)L::INV[i] * kRow - lo);
          // bounced link: read the node's own opposite population instead
          ld.raw[i] = bload16<T>(rsrc, select_bit(ld.m, i, b, a) + lo);
        }
      }
    } else {
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
}

template <bool B>
struct Forced {
  static constexpr bool value = B;
};

// BGK relaxation towards the equilibrium of (rho, u), plus the Guo source
// term when FORCED; emit(Forced<FORCED>{}, i, out) receives the Q outputs
// in the stored form and stored units (solid nodes: the rest state;
// SOLID false: the caller's wavefront has no solid node).  gv
// holds the inputs in stored units, k the constants in stored units; rs is
// rho (full storage) or drho = rho - 1 (shifted storage), in real units.
// Shifted storage forms feq - f = w rho t - w - g as
//   w ((t - 1) + drho + drho (t - 1)) - g
// from small terms only: no intermediate is near 1 or near w, so the
// rounding error of the collision scales with the deviations and a body
// force's small per-step increments are not lost.
template <int Q, int S, bool FORCED, bool SOLID, class Emit>
__device__ __forceinline__ void std_relax(const StepCoef<typename Store<S>::C>& k,
                                          const typename Store<S>::C (&gv)[Q],
                                          typename Store<S>::C rs, typename Store<S>::C ux,
                                          typename Store<S>::C uy, typename Store<S>::C uz,
                                          typename Store<S>::C m3uF, bool solid, Emit&& emit) {
#pragma clang fp contract(on)
  using L = Lat<Q>;
  using St = Store<S>;
  using C = typename St::C;
  constexpr bool kSh = St::kShifted;
  const C usq = L::D == 2 ? ux * ux + uy * uy : ux * ux + uy * uy + uz * uz;
  // t - 1 (shifted) or t = 1 - 1.5 u.u + 3 e.u + 4.5 (e.u)^2 (full)
  const C base = kSh ? (C)-1.5 * usq : (C)1 - (C)1.5 * usq;
  // 16-bit storage forms (t-1) rho + drho with rho = 1 + drho rounded once
  // per node: one fma per direction, an error of 6e-8 relative to the
  // term, far below the 16-bit stored value's resolution (FP32 storage
  // keeps the exact small-term form below)
  constexpr bool k16 = sizeof(typename St::T) == 2;
  const C rho = (C)1 + rs;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    // e.u from the nonzero components only (-0 + v == v exactly)
    C eu = (C)-0.0;
    if (L::E[i][0]) eu = L::E[i][0] > 0 ? eu + ux : eu - ux;
    if (L::E[i][1]) eu = L::E[i][1] > 0 ? eu + uy : eu - uy;
    if (L::E[i][2]) eu = L::E[i][2] > 0 ? eu + uz : eu - uz;
    const C w = (C)L::W[i];
    const C wu = w / St::kUnit;  // w in stored units (exact: kUnit is 2^k)
    const C t = fmadd(fmadd((C)4.5, eu, (C)3), eu, base);
    // feq - f in the stored form: w ((t-1) + drho + drho (t-1)) - g
    // (shifted, t here is t - 1), w rho t - f (full)
    const C d = k16 ? fmadd(wu, fmadd(t, rho, rs), -gv[i])
                : kSh ? fmadd(wu, fmadd(rs, t, t + rs), -gv[i])
                      : (w * rs) * t - gv[i];
    C out = fmadd(k.omega, d, gv[i]);
    if (FORCED) {
      // Guo source (1 - omega/2) w (3 (e.F - u.F) + 9 (e.u)(e.F)), added
      // term by term (each step takes one per-lattice constant); the rest
      // direction has e = 0, so only its -3 u.F term is nonzero
      const bool kRest = L::E[i][0] == 0 && L::E[i][1] == 0 && L::E[i][2] == 0;
      if (!kRest) out = out + k.gf3[i];
      out = fmadd(k.gw[i], m3uF, out);
      if (!kRest) out = fmadd(k.gf9[i], eu, out);
    }
    if (SOLID) out = solid ? (kSh ? (C)0 : w) : out;
    emit(Forced<FORCED>{}, i, out);
  }
}

// The collision of one node from its Q pulled (or bounced) values in the
// stored form and stored units, gv[i] (f_i, or f_i - w_i for shifted
// storage), its link mask m and solid flag; emit receives the outputs (see
// std_relax).
template <int Q, int S, bool LID, class Emit>
__device__ __forceinline__ void std_collide_node(const StdParams& p,
                                                 typename Store<S>::C (&gv)[Q], uint32_t m,
                                                 bool solid, Emit&& emit) {
  // a * b + c is fused only where one source expression has that form, so
  // every instantiation (single-lattice or batch, one or several nodes per
  // thread, with or without a moving wall) rounds identically
#pragma clang fp contract(on)
  using L = Lat<Q>;
  using St = Store<S>;
  using C = typename St::C;
  constexpr bool kSh = St::kShifted;
  const StepCoef<C>& k = step_coef<S>(p);
  // the moving-wall term of bounced links from beyond the moving face:
  // only nodes next to the moving face have such links, so a wavefront
  // without one skips the selection (the branch is uniform)
  if (LID) {
    // links of a node next to it (the 32-bit mask's flag, or for D2Q9 the
    // 16-bit copy's, which the 32-bit D2Q9 masks never set)
    constexpr uint32_t kLidBits = Q == 9 ? LBM_MASK_LID | kMask16Lid : LBM_MASK_LID;
    const uint32_t ml = (m & kLidBits) ? m : 0u;
    if (__builtin_amdgcn_ballot_w64(ml != 0u) != 0) {
#pragma unroll
      for (int i = 1; i < Q; ++i) {  // the rest direction never crosses a face
        if constexpr (sizeof(C) == 4) {
          // the term or +0: its bits masked by bit i spread over a word
          const uint32_t sel = (uint32_t)__builtin_amdgcn_sbfe((int)ml, i, 1);
          gv[i] += __uint_as_float(sel & __float_as_uint(k.lid[i]));
        } else {
          gv[i] += ((ml >> i) & 1u) ? k.lid[i] : (C)0;
        }
      }
    }
  }

  // moments; with shifted storage sum w_i = 1 and sum w_i e_i = 0, so the
  // sum of the stored values is rho itself (full storage) or the density
  // deviation drho = rho - 1 (shifted storage)
  // (the sum starts from the rest direction's value, not from 0 + it: the
  // two differ only for all inputs -0, whose outputs are the same)
  C rsu = gv[0], jx = 0, jy = 0, jz = 0;  // stored units
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (i) rsu += gv[i];
    if (L::E[i][0]) jx += L::E[i][0] > 0 ? gv[i] : -gv[i];
    if (L::E[i][1]) jy += L::E[i][1] > 0 ? gv[i] : -gv[i];
    if (L::E[i][2]) jz += L::E[i][2] > 0 ? gv[i] : -gv[i];
  }
  const C rs = rsu * St::kUnit;
  // 1 / rho, times kUnit: j (stored units) times it is u
  const C ir = recip(kSh ? (C)1 + rs : rs) * St::kUnit;
  // Guo forcing only where the lattice has a body force (uniform per
  // lattice); without one, u = j / rho and the source term is zero
  // a wavefront without a solid node skips the rest-state selection (the
  // branch is uniform).  Only for D2Q9 with 16-bit storage, which is bound
  // by instruction issue: the second copy of the relaxation costs the 3-D
  // lattices' batch kernels registers (their eight waves per SIMD then
  // spill to memory)
  constexpr bool kSkipSolid = Q == 9 && sizeof(typename St::T) == 2;
  const bool any_solid = !kSkipSolid || __builtin_amdgcn_ballot_w64(solid) != 0;
  // D2Q9 has no z velocity (its z force is 0: lbm_create checks)
  constexpr bool k2D = L::D == 2;
  if (k.forced) {
    const C ux = (jx + k.half_force[0]) * ir;
    const C uy = (jy + k.half_force[1]) * ir;
    const C uz = k2D ? (C)0 : (jz + k.half_force[2]) * ir;
    const C m3uF = k2D ? (C)-3 * (ux * k.force[0] + uy * k.force[1])
                       : (C)-3 * (ux * k.force[0] + uy * k.force[1] + uz * k.force[2]);
    if (any_solid)
      std_relax<Q, S, true, true>(k, gv, rs, ux, uy, uz, m3uF, solid, emit);
    else
      std_relax<Q, S, true, false>(k, gv, rs, ux, uy, uz, m3uF, solid, emit);
  } else {
    const C ux = jx * ir, uy = jy * ir, uz = k2D ? (C)0 : jz * ir;
    if (any_solid)
      std_relax<Q, S, false, true>(k, gv, rs, ux, uy, uz, (C)0, solid, emit);
    else
      std_relax<Q, S, false, false>(k, gv, rs, ux, uy, uz, (C)0, solid, emit);
  }
}

// Phase 2: collide the loaded node and store its Q post-collision values.
template <int Q, int S, bool LID, bool NT>
__device__ __forceinline__ void std_collide(const StdParams& p,
                                            typename Store<S>::T* __restrict__ B,
                                            const NodeLoad<Q, S>& ld) {
  using St = Store<S>;
  using C = typename St::C;
  C gv[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) gv[i] = St::cvt_u(ld.raw[i]);
  // the node's Q outputs lie kTile elements apart: one address (16-bit
  // storage: a byte offset into a buffer resource), the direction offsets
  // become constant offsets of the stores
  using T = typename St::T;
  T* const out0 = B + ld.own;
  const auto rsrc = buffer_rsrc(B, ld.pop_bytes);
  std_collide_node<Q, S, LID>(p, gv, ld.m, ld.solid, [&](auto forced, int i, C out) {
    // a body force moves 16-bit populations by less than their rounding
    // step: those lattices round stochastically (no effect at full
    // precision)
    T v;
    if constexpr (decltype(forced)::value)
      v = St::round_sr_u(out, sr_key(St::kShifted ? sr_node_hash(ld.own) : 0u, i));
    else
      v = St::round_u(out);
    if constexpr (sizeof(T) == 2)
      bstore16<NT>(rsrc, 2u * ld.own + (uint32_t)(i * kTile * 2), v);
    else if (NT)
      gstore_nt(out0 + i * kTile, v);
    else
      gstore(out0 + i * kTile, v);
  });
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

// ------------------------------------------------------ groups of NV nodes
// Nodes per thread of the group step for a lattice: 4 (rows along y of a
// multiple of 4 nodes), 2 (even rows), or 0 = the one-node step.  It is
// taken by D2Q9 lattices with 16-bit storage (their one-node update is
// bound by instruction issue) that stream from HBM (nt: both population
// copies exceed the Infinity Cache).  Measured on MI355X, FP16, per step:
// 4096^2 cavity 131.7 us one node per thread, 110.6 us two, 104.8 us four
// (eight: 110.0); at 2048^2 and below (cache-resident) the group step is
// 0-17 % slower than one node per thread.  The environment variable
// LBM_GROUP_STEP, read whenever a kernel is chosen, overrides the size
// rule: 1 = wherever the row length allows, 0 = nowhere (every width gives
// the same bytes).
inline int group_step_width(int q, int storage, const Grid& g, bool nt) {
  if (!(q == 9 && (storage == LBM_STORE_F16 || storage == LBM_STORE_BF16) && g.nz == 1))
    return 0;
  const int width = g.ny % 4 == 0 ? 4 : (g.ny % 2 == 0 ? 2 : 0);
  const char* env = std::getenv("LBM_GROUP_STEP");
  if (env && env[0] == '0') return 0;
  if (env && env[0] == '1') return width;
  return nt ? width : 0;
}

// A thread updates the NV consecutive nodes r0..r0+NV-1 (r0 % NV == 0)
// of one row along the fastest axis (y for D2Q9, z otherwise), so every
// load and store of a direction moves NV values per lane (consecutive
// elements of one tile run, aligned to NV elements), and the address
// arithmetic is spent once per group instead of once per node.
//
// Gathering direction i reads the aligned group of the source row (the
// row shifted by the other axes' components of -e_i) at r0..r0+NV-1.
// Directions without a component along the row need exactly that group.
// With +1 along the row the nodes need r0-1..r0+NV-2: NV-1 come from the
// group, r0-1 is the last value of the previous group of the same source
// row -- held by the thread of the previous group of this row, one lane
// down -- fetched with a lane permute; -1 likewise takes r0+NV from the
// next group's thread.  At the row ends the row wraps round (periodic;
// with walls the link is bounced and the value replaced).  Where the
// neighbouring group's thread sits in another wavefront (rows that do not
// tile a wavefront) or loaded nothing (a group of solid nodes; the source
// row need not be solid), the value is loaded instead.  Bounced links take
// the node's own opposite population: one load of the group's own run per
// direction that any of its nodes bounces, issued with the group loads.
//
// Callers: consecutive lanes of a wavefront take consecutive groups
// (n0 = NV x (a multiple of 64 + lane)), n0 < N, and every thread whose
// group lies in a row that any lane of the wavefront updates calls in
// together (rows are whole, so only threads past the lattice end drop out).
template <class T, int NV>
struct VecOf {
  typedef T V __attribute__((ext_vector_type(NV)));
};

// value v of lane `lane` of this wavefront (any lane index is harmless:
// the permute uses it modulo 64); 16-bit values travel zero-extended
template <class T>
__device__ __forceinline__ T lane_value(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(
        T, __builtin_amdgcn_ds_bpermute(lane << 2, __builtin_bit_cast(int, v)));
  } else {
    const int w = (int)__builtin_bit_cast(uint16_t, v);
    return __builtin_bit_cast(T, (uint16_t)__builtin_amdgcn_ds_bpermute(lane << 2, w));
  }
}

template <int Q, int S, int NV, bool LID, bool NT>
__device__ __forceinline__ void std_group_step(
    const Grid& g, const typename Store<S>::T* __restrict__ A,
    typename Store<S>::T* __restrict__ B, const uint32_t* __restrict__ mask,
    const uint32_t* __restrict__ cls, const StdParams& p, uint32_t n0) {
  using L = Lat<Q>;
  using St = Store<S>;
  using T = typename St::T;
  using C = typename St::C;
  using F = typename VecOf<T, NV>::V;
  using U = typename VecOf<uint32_t, NV>::V;
  constexpr int RA = Q == 9 ? 1 : 2;                  // the row axis
  constexpr uint32_t kBits = (1u << (2 * NV)) - 1u;  // the group's class bits
  const int lane = (int)(threadIdx.x & 63u);
  uint32_t x, y, z;
  if (Q == 9) {  // D2Q9 lattices have nz == 1
    z = 0;
    x = g.dy.div(n0);
    y = n0 - __umul24(x, g.ny);
  } else {
    decode_node(g, n0, x, y, z);
  }
  const uint32_t plane = g.ny * g.nz;
  // offsets from n0 to the source row for e = -1, 0, +1 along x and (3-D)
  // y: multiples of the row length, so the source group is aligned too
  const uint32_t ox[3] = {x + 1 == g.nx ? plane - g.n : plane, 0u,
                          x == 0 ? g.n - plane : 0u - plane};
  const uint32_t oy[3] = {y + 1 == g.ny ? g.nz - plane : g.nz, 0u,
                          y == 0 ? plane - g.nz : 0u - g.nz};
  auto source_row = [&](int i) -> uint32_t {
    return Q == 9 ? n0 + ox[L::E[i][0] + 1] : n0 + ox[L::E[i][0] + 1] + oy[L::E[i][1] + 1];
  };
  const uint32_t r0 = Q == 9 ? y : z, nr = Q == 9 ? g.ny : g.nz;

  // the group's 2-bit classes (2 NV bits of one class word), link masks
  const uint32_t cb = (gload(cls + (n0 >> 4)) >> (2u * (n0 & 15u))) & kBits;
  const bool all_solid = cb == (0xAAAAAAAAu & kBits);
  U mv = {};
  if (cb & 0x55555555u) mv = gload((const U*)(mask + n0));
  bool solid[NV];
  uint32_t m[NV];
  uint32_t mor = 0;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    solid[k] = ((cb >> (2 * k)) & 3u) == kClassSolid;
    m[k] = solid[k] ? 0u : mv[k];
    mor |= m[k];
  }

  // the aligned source groups (nothing for a group of solid nodes) and the
  // group's own opposite populations of bounced directions
  F G[Q], O[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) G[i] = __builtin_nondeterministic_value(G[i]);
  const uint32_t own = pidx<Q>(n0, 0);
  if (!all_solid) {
#pragma unroll
    for (int i = 0; i < Q; ++i) G[i] = gload((const F*)elem(A, pidx<Q>(source_row(i), i)));
  }
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    O[i] = __builtin_nondeterministic_value(O[i]);
    if ((mor >> i) & 1u) O[i] = gload((const F*)elem(A, own + (uint32_t)L::INV[i] * kTile));
  }

  // neighbouring groups of the row: their lanes; the value is loaded
  // instead where that lane is in another wavefront or loaded nothing;
  // node offsets from the source group to r0-1 and r0+NV.  Every permute
  // runs with all of the row's lanes active (a permute reads nothing from
  // an inactive lane).
  const int nq = (int)(nr / NV), rq = (int)(r0 / NV);
  const int lane_lo = rq == 0 ? lane + nq - 1 : lane - 1;
  const int lane_hi = rq + 1 == nq ? lane - nq + 1 : lane + 1;
  const int skipped = all_solid ? 1 : 0;
  const int lo_skipped = __builtin_amdgcn_ds_bpermute(lane_lo << 2, skipped);
  const int hi_skipped = __builtin_amdgcn_ds_bpermute(lane_hi << 2, skipped);
  const bool lo_far = !all_solid && ((unsigned)lane_lo >= 64u || lo_skipped != 0);
  const bool hi_far = !all_solid && ((unsigned)lane_hi >= 64u || hi_skipped != 0);
  const uint32_t d_lo = rq == 0 ? nr - 1u : 0u - 1u;
  const uint32_t d_hi = rq + 1 == nq ? (uint32_t)NV - nr : (uint32_t)NV;

  // V[i][k]: the stored value node r0+k pulls (or bounces) for direction i
  T V[Q][NV];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    if (L::E[i][RA] == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) V[i][k] = G[i][k];
    } else if (L::E[i][RA] > 0) {  // source r - 1
      T nb = lane_value(G[i][NV - 1], lane_lo);
      // (a bounced link's value is replaced below: nothing to load)
      if (lo_far && !((m[0] >> i) & 1u)) nb = gload(elem(A, pidx<Q>(source_row(i) + d_lo, i)));
      V[i][0] = nb;
#pragma unroll
      for (int k = 1; k < NV; ++k) V[i][k] = G[i][k - 1];
    } else {  // source r + 1
      T nb = lane_value(G[i][0], lane_hi);
      if (hi_far && !((m[NV - 1] >> i) & 1u))
        nb = gload(elem(A, pidx<Q>(source_row(i) + d_hi, i)));
#pragma unroll
      for (int k = 0; k + 1 < NV; ++k) V[i][k] = G[i][k + 1];
      V[i][NV - 1] = nb;
    }
  }
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    if ((mor >> i) & 1u) {
#pragma unroll
      for (int k = 0; k < NV; ++k) V[i][k] = ((m[k] >> i) & 1u) ? O[i][k] : V[i][k];
    }
  }

  // collide the group's nodes (in the stored type again: rounded as the
  // one-node step rounds), then store each direction's NV values
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    C gv[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) gv[i] = St::cvt_u(V[i][k]);
    std_collide_node<Q, S, LID>(p, gv, m[k], solid[k], [&](auto forced, int i, C out) {
      if constexpr (decltype(forced)::value)
        V[i][k] = St::round_sr_u(out, sr_key(St::kShifted ? sr_node_hash(own + k) : 0u, i));
      else
        V[i][k] = St::round_u(out);
    });
  }
  T* const out0 = B + own;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    F o;
#pragma unroll
    for (int k = 0; k < NV; ++k) o[k] = V[i][k];
    if (NT)
      gstore_nt((F*)elem(out0, (uint32_t)i * kTile), o);
    else
      gstore((F*)elem(out0, (uint32_t)i * kTile), o);
  }
}

}  // namespace lbm
