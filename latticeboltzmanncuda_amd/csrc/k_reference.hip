// Reference-mode kernels: the reference's D3Q27 stages re-designed for the
// blocked structure-of-arrays layout of lbm_layout.h (node index
// n = z + nz*(y + ny*x)), one thread per node over a flat grid.
//
//   k_ref_stream          SimCalcFuncs.cu:79-96   pull, zero outside the box
//   k_ref_collide         SimCalcFuncs.cu:98-156  SPLIT velocity, FP64
//                                                 equilibrium, FP32 relax;
//                         optionally followed by SimCalcFuncs.cu:200-215
//                         (flow injection) in the same pass
//   k_ref_stream_collide  Lattice.cpp:222-226     both of the above fused
//   k_ref_reflect         SimCalcFuncs.cu:158-198 reflection tables
//   k_ref_flow            SimCalcFuncs.cu:200-215 flow injection
//   k_ref_reflect_gather  the reflections of one step as a gather over
//   k_ref_step_full       (receiver, direction) groups, then the main
//                         loop's stream + reflect + collide + flow in one
//                         pass (LatticeBoltzmannCuda.cpp:85-93)
//
// The reference copies the whole lattice before streaming and before
// reflecting (Lattice.cpp:126-133,176-183).  Here streaming goes A -> B and
// every population is written, so no copy is needed; reflection runs in
// place: a reflecting (internal) node reads only its own populations, and
// the target (x + e_d, d) of a slot has exactly one possible source node
// x, so the scatter needs no atomics and has a fixed summation order.
//
// Floating-point contraction is off in this file so that every operation
// rounds as written in the reference source.
#include <hip/hip_runtime.h>

#include "lbm_internal.h"
#include "lbm_lattice.h"
#include "lbm_layout.h"
#include "lbm_store.h"

#pragma clang fp contract(off)

namespace lbm {
namespace {

constexpr int kBlock = 256;
using L27 = Lat<27>;

__device__ __forceinline__ void decode(const Grid& g, uint32_t n, uint32_t& x,
                                       uint32_t& y, uint32_t& z) {
  const uint32_t t = g.dz.div(n);
  z = n - t * g.nz;
  x = g.dy.div(t);
  y = t - x * g.ny;
}

__device__ __forceinline__ bool flow_match(const FlowSpec& fl, uint32_t n,
                                           int x, int y, int z) {
  switch (fl.kind) {
    case LBM_FLOW_POINT:
      return x == fl.params[0] && y == fl.params[1] && z == fl.params[2];
    case LBM_FLOW_PLANE: {
      const int c = fl.params[0] == 0 ? x : (fl.params[0] == 1 ? y : z);
      return c == fl.params[1];
    }
    case LBM_FLOW_BOX:
      return x >= fl.params[0] && y >= fl.params[1] && z >= fl.params[2] &&
             x <= fl.params[3] && y <= fl.params[4] && z <= fl.params[5];
    case LBM_FLOW_MASK:
      return fl.mask[n] != 0;
    default:
      return false;
  }
}

// Zero-fill pull of direction j into f[j]: f_j(x) = A_j(x - e_j).
__device__ __forceinline__ void pull27(const Grid& g, const float* __restrict__ A,
                                       uint32_t n, int x, int y, int z,
                                       float f[27]) {
  // per axis, indexed by e + 1 for e = -1, 0, +1: whether the source
  // coordinate x - e lies in the box, and the node offset of x - e (mod 2^32)
  const uint32_t plane = g.ny * g.nz;
  const bool bx[3] = {x + 1 < (int)g.nx, true, x > 0};
  const bool by[3] = {y + 1 < (int)g.ny, true, y > 0};
  const bool bz[3] = {z + 1 < (int)g.nz, true, z > 0};
  const uint32_t ox[3] = {plane, 0u, 0u - plane};
  const uint32_t oy[3] = {g.nz, 0u, 0u - g.nz};
  const uint32_t oz[3] = {1u, 0u, 0u - 1u};
#pragma unroll
  for (int j = 0; j < 27; ++j) {
    const int a = L27::E[j][0] + 1, b = L27::E[j][1] + 1, c = L27::E[j][2] + 1;
    const bool in = bx[a] && by[b] && bz[c];
    const uint32_t src = in ? n + ox[a] + oy[b] + oz[c] : n;
    const float v = A[pidx<27>(src, j)];
    f[j] = in ? v : 0.0f;
  }
}

// Byte offsets of the linear pull (lattices whose nz is a multiple of 64,
// so that n, n +- nz and n +- ny nz sit at the same place in their tiles):
// b[j] = 4 (pidx(n - e_j, j) - pidx(n, 0)) for a node whose z-source stays
// in its tile, computed on the host.
struct PullOffsets {
  int32_t b[27];
};

// pull27 for such lattices through a buffer resource over the population
// array: the source of direction j is the node's own offset plus b[j], plus
// one tile crossing for a z-source beyond the tile's first or last node,
// and a source outside the box gets an offset past the array, which reads
// 0.  Two instructions per direction instead of the index arithmetic.
// zero: every direction reads 0 (an internal node, zeroed by the
// reflection stage) without touching memory.
__device__ __forceinline__ void pull27_lin(const Grid& g, __amdgpu_buffer_rsrc_t r,
                                           const PullOffsets& po, uint32_t n, int x, int y,
                                           int z, bool zero, float f[27]) {
  // a source one node up z from a tile's last node is the next tile's
  // first (e_z = -1), one node down from its first node the previous
  // tile's last (e_z = +1): the linear offset misses by 26 rows of 64
  constexpr uint32_t kCross = (27u - 1u) * kTile * 4u;
  const uint32_t own = pidx<27>(n, 0) * 4u;
  const uint32_t l = n & (kTile - 1);
  const uint32_t base[3] = {own + (l == kTile - 1 ? kCross : 0u), own,
                            own - (l == 0 ? kCross : 0u)};
  const bool bx[3] = {x + 1 < (int)g.nx, true, x > 0};
  const bool by[3] = {y + 1 < (int)g.ny, true, y > 0};
  const bool bz[3] = {z + 1 < (int)g.nz, true, z > 0};
#pragma unroll
  for (int j = 0; j < 27; ++j) {
    const int a = L27::E[j][0] + 1, b = L27::E[j][1] + 1, c = L27::E[j][2] + 1;
    const bool in = !zero && bx[a] && by[b] && bz[c];
    const uint32_t off = in ? base[c] + (uint32_t)po.b[j] : 0x80000000u;
    f[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
  }
}

// The relaxation time and its reciprocal rounded to double (computed on
// the host); see div_tau.
struct Relax {
  float tau;
  double rinv;
};

inline Relax make_relax(float tau) { return Relax{tau, 1.0 / (double)tau}; }

// x / tau rounded to float, as the reference's IEEE division gives it.
// q = x * RN64(1/tau) is within 2^-52 (relative) of the exact quotient,
// while the quotient of two floats lies at least 2^-49 away from any point
// halfway between two floats (it can never be one: a halfway point needs
// 25 significant bits, and x = midpoint * tau would then need more than
// 24), so rounding q to float gives the correctly rounded quotient.  Below
// the float normal range the rounding position moves, so tiny quotients
// take the IEEE division itself.
__device__ __forceinline__ float div_tau(float x, const Relax& r) {
  const double q = (double)x * r.rinv;
  if (fabs(q) < 1.1754943508222875e-38) return x / r.tau;
  return (float)q;
}

// SimCalcFuncs.cu:102-155 on one node's 27 populations, bit-identical to
// evaluating the reference source as written, with fewer instructions:
//  - terms e_i,a * u_a with e_i,a = 0 are left out of the velocity sums and
//    of e.u: for finite values they only add a signed zero, and the sign of
//    a zero e.u or u_a cannot reach the result (u_a and e.u enter only
//    through squares and 1 + 3 e.u);
//  - 3.0 * ct, 4.5 * ct and (4.5 * ct) * ct, 1.5 * usq are exact in double
//    (ct, usq are floats: at most 2 + 24, 4 + 24, 28 + 24 and 2 + 24
//    significant bits), so each product-then-add is one FMA with the same
//    single rounding; the additions keep the source's order;
//  - opposite directions share e.u in double and 4.5 e.u, negated (exact);
//  - the division by tau is div_tau.
__device__ __forceinline__ void collide27(float f[27], const Relax& r) {
  float density = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    density += f[i];
    if (L27::E[i][0]) ux += (L27::E[i][0] > 0 ? f[i] : -f[i]) * kSplit27[i];
    if (L27::E[i][1]) uy += (L27::E[i][1] > 0 ? f[i] : -f[i]) * kSplit27[i];
    if (L27::E[i][2]) uz += (L27::E[i][2] > 0 ? f[i] : -f[i]) * kSplit27[i];
  }
  if (density != 0.0f) {
    ux /= density;
    uy /= density;
    uz /= density;
  }
  const float usq = (ux * ux) + (uy * uy) + (uz * uz);
  const double m15usq = -1.5 * (double)usq;
  // e.u of direction i in double, and 4.5 e.u: the opposite direction's
  // are their exact negatives (its sum has every term negated, and
  // rounding is symmetric), so a pair shares one conversion and one
  // product; each direction's arithmetic is unchanged
  auto relax = [&](int i, double c, double p) {
    const double w = i == 0 ? 8.0 / 27.0
                            : (i < 7 ? 2.0 / 27.0 : (i < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
    // ((1 + 3 ct) + (4.5 ct) ct) - 1.5 usq, each step rounded once
    const double t1 = fma(3.0, c, 1.0);
    const double t2 = fma(p, c, t1);
    const double t3 = t2 + m15usq;
    const float eq = (float)((w * (double)density) * t3);
    f[i] = f[i] + div_tau(eq - f[i], r);
  };
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    if (L27::INV[i] < i) continue;
    float ct = 0.0f;
    bool first = true;
    if (L27::E[i][0]) {
      ct = L27::E[i][0] > 0 ? ux : -ux;
      first = false;
    }
    if (L27::E[i][1]) {
      const float t = L27::E[i][1] > 0 ? uy : -uy;
      ct = first ? t : ct + t;
      first = false;
    }
    if (L27::E[i][2]) {
      const float t = L27::E[i][2] > 0 ? uz : -uz;
      ct = first ? t : ct + t;
    }
    const double c = (double)ct;
    const double p = 4.5 * c;
    relax(i, c, p);
    if (L27::INV[i] != i) relax(L27::INV[i], -c, -p);
  }
}

__global__ void __launch_bounds__(kBlock)
    k_ref_stream(Grid g, const float* __restrict__ A, float* __restrict__ B) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  uint32_t x, y, z;
  decode(g, n, x, y, z);
  float f[27];
  pull27(g, A, n, (int)x, (int)y, (int)z, f);
#pragma unroll
  for (int j = 0; j < 27; ++j) B[pidx<27>(n, j)] = f[j];
}

__global__ void __launch_bounds__(kBlock)
    k_ref_collide(Grid g, float* __restrict__ F, Relax tau, FlowSpec flow) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  float f[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) f[i] = F[pidx<27>(n, i)];
  collide27(f, tau);
  if (flow.kind != LBM_FLOW_NONE) {
    uint32_t x, y, z;
    decode(g, n, x, y, z);
    if (flow_match(flow, n, (int)x, (int)y, (int)z)) {
#pragma unroll
      for (int i = 0; i < 27; ++i) f[i] = f[i] + flow.induced[i];
    }
  }
#pragma unroll
  for (int i = 0; i < 27; ++i) F[pidx<27>(n, i)] = f[i];
}

__global__ void __launch_bounds__(kBlock)
    k_ref_stream_collide(Grid g, const float* __restrict__ A,
                         float* __restrict__ B, Relax tau) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  uint32_t x, y, z;
  decode(g, n, x, y, z);
  float f[27];
  pull27(g, A, n, (int)x, (int)y, (int)z, f);
  collide27(f, tau);
#pragma unroll
  for (int j = 0; j < 27; ++j) B[pidx<27>(n, j)] = f[j];
}

__global__ void __launch_bounds__(kBlock)
    k_ref_flow(Grid g, float* __restrict__ F, FlowSpec flow) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  uint32_t x, y, z;
  decode(g, n, x, y, z);
  if (!flow_match(flow, n, (int)x, (int)y, (int)z)) return;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    float* p = F + pidx<27>(n, i);
    *p = *p + flow.induced[i];
  }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Called by all kBlock threads of one workgroup: sums the partial forces
// (3 per entry) in FP64 in a fixed order -- thread t takes entries t,
// t + kBlock, ...; then a fixed tree -- and adds the total to the FP32
// force register: one rounding per reflection pass, the same result on
// every run.
__device__ void force_sum_block(const float* __restrict__ partials, int nblocks,
                                float* __restrict__ force) {
  __shared__ double red[3][kBlock];
  const int t = threadIdx.x;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int b = t; b < nblocks; b += kBlock) {
    s0 += (double)partials[3 * b];
    s1 += (double)partials[3 * b + 1];
    s2 += (double)partials[3 * b + 2];
  }
  red[0][t] = s0;
  red[1][t] = s1;
  red[2][t] = s2;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (t < h) {
      red[0][t] += red[0][t + h];
      red[1][t] += red[1][t + h];
      red[2][t] += red[2][t + h];
    }
    __syncthreads();
  }
  if (t < 3) force[t] = (float)((double)force[t] + red[t][0]);
}

// One thread per node with a nonzero flag byte.  Slot s of a record holds a
// 5-bit direction d in word s/6, bits 5*(s%6) (Lattice.cpp:450-477); its
// contribution f_{s/3}(x) * W[s] goes to (x + e_d, d).  The force term uses
// DIRECTIONS[s]/SPLIT[s] as SimCalcFuncs.cu:181-183 does; for s >= 27 that
// operand lies outside the reference's tables and is taken as 0 here.
__global__ void __launch_bounds__(kBlock)
    k_ref_reflect(Grid g, float* __restrict__ F, int32_t nspecial,
                  const int32_t* __restrict__ special,
                  const int32_t* __restrict__ special_rec,
                  const uint8_t* __restrict__ flags,
                  const uint32_t* __restrict__ rec_dirs,
                  const float* __restrict__ rec_w,
                  float* __restrict__ partials) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (k < nspecial) {
    const uint32_t n = (uint32_t)special[k];
    const uint8_t fl = flags[n];
    const int32_t r = special_rec[k];
    uint32_t x, y, z;
    decode(g, n, x, y, z);
    if ((fl & LBM_FLAG_REFLECTED) && r >= 0) {
      const uint32_t* words = rec_dirs + 14 * (uint64_t)r;
      const float* W = rec_w + 81 * (uint64_t)r;
      for (int s = 1; s < 81; ++s) {
        const int d = (int)((words[s / 6] >> (5 * (s % 6))) & 0x1Fu);
        if (d == 0 || d > 26) continue;
        const int tx = (int)x + L27::E[d][0], ty = (int)y + L27::E[d][1],
                  tz = (int)z + L27::E[d][2];
        if (tx < 0 || ty < 0 || tz < 0 || tx >= (int)g.nx ||
            ty >= (int)g.ny || tz >= (int)g.nz)
          continue;
        const float val = F[pidx<27>(n, s / 3)] * W[s];
        float* t = F + pidx<27>((uint32_t)tz + g.nz * ((uint32_t)ty + g.ny * (uint32_t)tx), d);
        *t = *t + val;
        const float ts0 = s < 27 ? (float)L27::E[s][0] * kSplit27[s] : 0.0f;
        const float ts1 = s < 27 ? (float)L27::E[s][1] * kSplit27[s] : 0.0f;
        const float ts2 = s < 27 ? (float)L27::E[s][2] * kSplit27[s] : 0.0f;
        fx += val * (((float)L27::E[d][0] * kSplit27[d]) - ts0);
        fy += val * (((float)L27::E[d][1] * kSplit27[d]) - ts1);
        fz += val * (((float)L27::E[d][2] * kSplit27[d]) - ts2);
      }
    }
    if (fl & LBM_FLAG_INTERNAL) {
#pragma unroll
      for (int i = 0; i < 27; ++i) F[pidx<27>(n, i)] = 0.0f;
    }
  }
  if (partials) {
    __shared__ float red[3][kBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const float sx = wave_sum(fx), sy = wave_sum(fy), sz = wave_sum(fz);
    if (lane == 0) {
      red[0][wv] = sx;
      red[1][wv] = sy;
      red[2][wv] = sz;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      float s = 0.0f;
      for (int w = 0; w < kBlock / 64; ++w) s += red[threadIdx.x][w];
      partials[3 * blockIdx.x + threadIdx.x] = s;
    }
  }
}

// e_s SPLIT_s of slot s per axis (0 for s >= 27, outside the reference's
// tables): the second factor of the force term of SimCalcFuncs.cu:181-183.
struct SlotTerms {
  float v[81][3];
};
constexpr SlotTerms make_slot_terms() {
  SlotTerms t{};
  for (int s = 0; s < 27; ++s)
    for (int a = 0; a < 3; ++a) t.v[s][a] = (float)L27::E[s][a] * kSplit27[s];
  return t;
}
__constant__ constexpr SlotTerms kSlotTerms = make_slot_terms();

// Gather pass of the fused reference step: one thread per group (t, j),
// the threads ordered by reflecting node x, then j (groups of one x read
// the same neighbours of x, so a wavefront touches fewer cache lines), each
// writing its sum to its group's slot gid of gsum.  The group's value is
// the staged pipeline's f_j(t) after reflection:
// the streamed value A_j(x) of the reflecting node x = t - e_j, then, in
// ascending slot order, + W[s] * F_k(x) with F_k(x) = A_k(x - e_k) (0
// outside the box) -- the same additions in the same order as the staged
// pass.  The force gets val * (e_j SPLIT_j - e_s SPLIT_s) per slot
// (SimCalcFuncs.cu:181-183), summed per wavefront.  Entries are taken four
// at a time: their loads are issued together, the sums still run in entry
// order.
__global__ void __launch_bounds__(kBlock)
    k_ref_reflect_gather(const float* __restrict__ A, uint32_t ngrp,
                         const uint32_t* __restrict__ gsrc,
                         const uint32_t* __restrict__ gbeg,
                         const uint32_t* __restrict__ gid, uint32_t nent,
                         const uint32_t* __restrict__ eidx,
                         const float* __restrict__ ew,
                         const uint8_t* __restrict__ es, float* __restrict__ gsum,
                         float* __restrict__ partials) {
  constexpr int kU = 4;
  const uint32_t gi = blockIdx.x * kBlock + threadIdx.x;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  if (gi < ngrp) {
    const uint32_t src = gsrc[gi];
    const uint32_t grp = gid[gi];
    float acc = A[src];
    const uint32_t end = gi + 1 < ngrp ? gbeg[gi + 1] : nent;
    // the group's direction j: gsrc is the element index of (x, j)
    const int j = (int)((src / kTile) % 27u);
    const float tj0 = (float)L27::E[j][0] * kSplit27[j];
    const float tj1 = (float)L27::E[j][1] * kSplit27[j];
    const float tj2 = (float)L27::E[j][2] * kSplit27[j];
    for (uint32_t e = gbeg[gi]; e < end; e += kU) {
      uint32_t el[kU], sl[kU];
      float w[kU], v[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const bool in = e + u < end;
        el[u] = in ? eidx[e + u] : ~0u;
        w[u] = in ? ew[e + u] : 0.0f;
        sl[u] = in ? es[e + u] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) v[u] = el[u] != ~0u ? A[el[u]] : 0.0f;
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        if (e + u < end) {
          const float val = v[u] * w[u];
          acc = acc + val;
          fx += val * (tj0 - kSlotTerms.v[sl[u]][0]);
          fy += val * (tj1 - kSlotTerms.v[sl[u]][1]);
          fz += val * (tj2 - kSlotTerms.v[sl[u]][2]);
        }
      }
    }
    gsum[grp] = acc;
  }
  if (partials) {
    const float sx = wave_sum(fx), sy = wave_sum(fy), sz = wave_sum(fz);
    if ((threadIdx.x & 63) == 0) {
      const uint32_t w = gi >> 6;
      partials[3 * w] = sx;
      partials[3 * w + 1] = sy;
      partials[3 * w + 2] = sz;
    }
  }
}

// One step of the reference's main loop in one pass (A -> B):
//   S  f_j(n) = A_j(n - e_j), 0 outside the box;
//   R  internal nodes are zeroed; for a receiving node n the gather pass
//      has already formed f_j(n) for every incoming direction j (bit j of
//      its in-mask; its groups ascend in j from rxbase[n]);
//   C  collide27; F flow injection.
// The populations are bit-identical to the staged passes.
template <bool GEOM, bool NT, bool LIN>
__device__ __forceinline__ void step_full_node(
    const Grid& g, const float* __restrict__ A, float* __restrict__ B, const Relax& tau,
    const FlowSpec& flow, const uint32_t* __restrict__ cls, const uint32_t* __restrict__ inmask,
    const int32_t* __restrict__ rxbase, const float* __restrict__ gsum, const PullOffsets& po,
    const uint32_t* __restrict__ keep, uint32_t n) {
  uint32_t ux, uy, uz;
  decode(g, n, ux, uy, uz);
  const int x = (int)ux, y = (int)uy, z = (int)uz;
  float f[27];
  if (LIN) {
    // the class word first: an internal node's loads read 0 without
    // touching memory, which is its zeroing
    const uint32_t c = GEOM ? (cls[n >> 4] >> (2u * (n & 15u))) & 3u : 0u;
    pull27_lin(g, buffer_rsrc(A, 27u * 4u * g.n), po, n, x, y, z, c == 2u, f);
    if (c == 1u) {
      const uint32_t m = inmask[n];
      uint32_t p = (uint32_t)rxbase[n];
#pragma unroll
      for (int j = 1; j < 27; ++j)
        if ((m >> j) & 1u) f[j] = gsum[p++];
    }
  } else {
    pull27(g, A, n, x, y, z, f);
  }
  if (GEOM && !LIN) {
    const uint32_t c = (cls[n >> 4] >> (2u * (n & 15u))) & 3u;
    if (c == 2u) {
#pragma unroll
      for (int i = 0; i < 27; ++i) f[i] = 0.0f;
    } else if (c == 1u) {
      const uint32_t m = inmask[n];
      uint32_t p = (uint32_t)rxbase[n];
#pragma unroll
      for (int j = 1; j < 27; ++j)
        if ((m >> j) & 1u) f[j] = gsum[p++];
    }
  }
  collide27(f, tau);
  if (flow.kind != LBM_FLOW_NONE && flow_match(flow, n, x, y, z)) {
#pragma unroll
    for (int i = 0; i < 27; ++i) f[i] = f[i] + flow.induced[i];
  }
  if constexpr (GEOM && NT && LIN) {
    // tiles around reflecting nodes (which the next step's gather and this
    // pass read) are stored with the ordinary policy, so that they stay in
    // the caches for them, the others non-temporally; the bit is a wavefront constant (a wavefront
    // is one tile).  Buffer stores carry the policy as an immediate (the
    // compiler would merge two branches of plain stores, hint and all).
    const auto rb = buffer_rsrc(B, 27u * 4u * g.n);
    const uint32_t own = pidx<27>(n, 0) * 4u;
    const uint32_t tile = __builtin_amdgcn_readfirstlane(n) >> kTileShift;
    if (((keep[tile >> 5] >> (tile & 31u)) & 1u) == 0u) {
#pragma unroll
      for (int i = 0; i < 27; ++i)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, f[i]), rb,
                                              (int)(own + (uint32_t)i * kTile * 4u), 0, 2);
    } else {
#pragma unroll
      for (int i = 0; i < 27; ++i)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, f[i]), rb,
                                              (int)(own + (uint32_t)i * kTile * 4u), 0, 0);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 27; ++i) {
      if (NT)
        __builtin_nontemporal_store(f[i], &B[pidx<27>(n, i)]);
      else
        B[pidx<27>(n, i)] = f[i];
    }
  }
}

template <bool GEOM, bool NT, bool LIN>
__global__ void __launch_bounds__(kBlock)
    k_ref_step_full(Grid g, const float* __restrict__ A, float* __restrict__ B,
                    Relax tau, FlowSpec flow, const uint32_t* __restrict__ cls,
                    const uint32_t* __restrict__ inmask,
                    const int32_t* __restrict__ rxbase,
                    const float* __restrict__ gsum, int rev,
                    const float* __restrict__ partials, int npart,
                    float* __restrict__ force, PullOffsets po,
                    const uint32_t* __restrict__ keep) {
  const uint32_t blk = rev ? gridDim.x - 1 - blockIdx.x : blockIdx.x;
  const uint32_t n = blk * kBlock + threadIdx.x;
  if (n < g.n)
    step_full_node<GEOM, NT, LIN>(g, A, B, tau, flow, cls, inmask, rxbase, gsum, po, keep, n);
  // the gather pass's force partials are complete when this kernel starts:
  // the first workgroup dispatched adds them to the force register (its
  // extra work overlaps the rest of the grid)
  if (GEOM && force && blockIdx.x == 0) force_sum_block(partials, npart, force);
}

// Adds the per-block partial forces of one reflection pass to the FP32
// force register; see force_sum_block.
__global__ void __launch_bounds__(kBlock)
    k_ref_force_sum(const float* __restrict__ partials, int nblocks,
                    float* __restrict__ force) {
  force_sum_block(partials, nblocks, force);
}

__global__ void __launch_bounds__(kBlock)
    k_ref_moments(Grid g, const float* __restrict__ F, float* __restrict__ rho,
                  float* __restrict__ u) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n) return;
  float d = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const float fi = F[pidx<27>(n, i)];
    d += fi;
    ux += fi * (float)L27::E[i][0] * kSplit27[i];
    uy += fi * (float)L27::E[i][1] * kSplit27[i];
    uz += fi * (float)L27::E[i][2] * kSplit27[i];
  }
  if (d != 0.0f) {
    ux /= d;
    uy /= d;
    uz /= d;
  }
  if (rho) rho[n] = d;
  if (u) {
    u[n] = ux;
    u[g.n + n] = uy;
    u[2 * (uint64_t)g.n + n] = uz;
  }
}

inline unsigned blocks_for(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

int launch_ref_stream(const Grid& g, const float* A, float* B, hipStream_t s) {
  hipLaunchKernelGGL(k_ref_stream, dim3(blocks_for(g.n)), dim3(kBlock), 0, s,
                     g, A, B);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_ref_collide(const Grid& g, float* f, float tau,
                       const FlowSpec* flow, hipStream_t s) {
  FlowSpec none{};
  none.kind = LBM_FLOW_NONE;
  hipLaunchKernelGGL(k_ref_collide, dim3(blocks_for(g.n)), dim3(kBlock), 0, s,
                     g, f, make_relax(tau), flow ? *flow : none);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_ref_stream_collide(const Grid& g, const float* A, float* B,
                              float tau, hipStream_t s) {
  hipLaunchKernelGGL(k_ref_stream_collide, dim3(blocks_for(g.n)),
                     dim3(kBlock), 0, s, g, A, B, make_relax(tau));
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int ref_reflect_blocks(int32_t nspecial) {
  return (int)((nspecial + kBlock - 1) / kBlock);
}

int launch_ref_reflect(const Grid& g, float* f, int32_t nspecial,
                       const int32_t* special, const int32_t* special_rec,
                       const uint8_t* flags, const uint32_t* rec_dirs,
                       const float* rec_w, float* force_partials,
                       int nblocks, float* force, hipStream_t s) {
  if (nspecial <= 0) return LBM_OK;
  hipLaunchKernelGGL(k_ref_reflect, dim3(nblocks), dim3(kBlock), 0, s, g, f,
                     nspecial, special, special_rec, flags, rec_dirs, rec_w,
                     force ? force_partials : nullptr);
  LBM_HIP(hipGetLastError());
  if (force) {
    hipLaunchKernelGGL(k_ref_force_sum, dim3(1), dim3(kBlock), 0, s,
                       force_partials, nblocks, force);
    LBM_HIP(hipGetLastError());
  }
  return LBM_OK;
}

int ref_gather_partials(int64_t ngrp) {
  return (int)(blocks_for((uint32_t)ngrp) * (kBlock / 64));
}

int launch_ref_step_full(const Grid& g, const float* A, float* B, float tau,
                         const FlowSpec& flow, const uint32_t* cls,
                         const uint32_t* inmask, const int32_t* rxbase,
                         int64_t ngrp, const uint32_t* gsrc, const uint32_t* gbeg,
                         const uint32_t* gid, float* gsum, int64_t nent, const uint32_t* eidx,
                         const uint32_t* keep,
                         const float* ew, const uint8_t* es, float* partials,
                         float* force, int rev, hipStream_t s) {
  const unsigned nb = blocks_for(g.n);
  // the force partials of the gather pass, summed by the fused pass
  const bool with_force = cls && ngrp > 0 && force && partials;
  if (cls && ngrp > 0) {
    hipLaunchKernelGGL(k_ref_reflect_gather, dim3(blocks_for((uint32_t)ngrp)), dim3(kBlock), 0, s,
                       A, (uint32_t)ngrp, gsrc, gbeg, gid, (uint32_t)nent, eidx, ew, es, gsum,
                       with_force ? partials : nullptr);
    LBM_HIP(hipGetLastError());
  }
  const int npart = with_force ? ref_gather_partials(ngrp) : 0;
  float* fsum = with_force ? force : nullptr;
  // non-temporal stores and alternating tile order once both population
  // copies exceed the Infinity Cache (as for the standard step)
  const bool nt = streaming_stores(2ull * 27 * sizeof(float) * g.n);
  const int r = nt ? rev : 0;
  const Relax rx = make_relax(tau);
  // the linear pull: nz a multiple of 64 (a node's y and x neighbours sit
  // at its place in their tiles) and population offsets below 2^31 (the
  // offset that reads 0 lies past the array)
  const bool lin = g.nz % kTile == 0 && 27ull * 4ull * g.n < (1ull << 31);
  PullOffsets po{};
  if (lin) {
    const int64_t plane = (int64_t)g.ny * g.nz;
    for (int j = 0; j < 27; ++j) {
      const int64_t dt = -((int64_t)L27::E[j][0] * plane + (int64_t)L27::E[j][1] * g.nz) / kTile;
      po.b[j] = (int32_t)(4 * (dt * 27 * kTile + (int64_t)j * kTile - L27::E[j][2]));
    }
  }
  auto go = [&](auto kernel, bool geom) {
    if (geom)
      hipLaunchKernelGGL(kernel, dim3(nb), dim3(kBlock), 0, s, g, A, B, rx, flow, cls, inmask,
                         rxbase, gsum, r, partials, npart, fsum, po, keep);
    else
      hipLaunchKernelGGL(kernel, dim3(nb), dim3(kBlock), 0, s, g, A, B, rx, flow, nullptr,
                         nullptr, nullptr, nullptr, r, nullptr, 0, nullptr, po, nullptr);
  };
  if (lin) {
    if (cls && nt) go(k_ref_step_full<true, true, true>, true);
    else if (cls) go(k_ref_step_full<true, false, true>, true);
    else if (nt) go(k_ref_step_full<false, true, true>, false);
    else go(k_ref_step_full<false, false, true>, false);
  } else {
    if (cls && nt) go(k_ref_step_full<true, true, false>, true);
    else if (cls) go(k_ref_step_full<true, false, false>, true);
    else if (nt) go(k_ref_step_full<false, true, false>, false);
    else go(k_ref_step_full<false, false, false>, false);
  }
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_ref_flow(const Grid& g, float* f, const FlowSpec& flow,
                    hipStream_t s) {
  if (flow.kind == LBM_FLOW_NONE) return LBM_OK;
  hipLaunchKernelGGL(k_ref_flow, dim3(blocks_for(g.n)), dim3(kBlock), 0, s, g,
                     f, flow);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

int launch_ref_moments(const Grid& g, const float* f, float* rho, float* u,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_ref_moments, dim3(blocks_for(g.n)), dim3(kBlock), 0, s,
                     g, f, rho, u);
  LBM_HIP(hipGetLastError());
  return LBM_OK;
}

}  // namespace lbm
