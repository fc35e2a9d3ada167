// Mesh geometry on the GPU.
//
// lbm_voxelize_mesh: which lattice nodes lie inside a closed triangle mesh.  It replaces the host loop of
// Lattice::insertModel (reference Lattice.cpp:228-260) over
// Model::isPointInsideModel (reference Model.cpp:51-63) and decides every
// node as the host Model does: a node on a face (coplanar up to 1e-12
// relative, inside its three edges) is inside; otherwise the generalized
// winding number W = sum of signed solid angles / 4 pi decides, inside
// when |W| > 1/2.  Same formulas, double precision, no contraction; only
// atan2 comes from the GPU's math library, and it cannot change a decision:
// W is within 1e-15 of 0.5 only for points closer to a face than the
// on-face tolerance.  Nodes outside the mesh's bounding box (with a 1e-9
// relative margin) are outside without a test: W is 0 there up to
// rounding and no face can touch them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "lbm_internal.h"
#include "lbm_lattice.h"

#pragma clang fp contract(off)

namespace lbm {
namespace {

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double norm(D3 a) { return sqrt(dot(a, a)); }

__device__ bool on_triangle(D3 p, D3 a, D3 b, D3 c) {
  const D3 n = cross(sub(b, a), sub(c, a));
  const double nn = norm(n);
  if (nn == 0.0) return false;
  const double off = dot(n, sub(p, a));
  const double scale = nn * (norm(sub(p, a)) + norm(sub(b, a)));
  if (fabs(off) > 1e-12 * scale) return false;
  const double tol = -1e-12 * nn * nn;
  return dot(cross(sub(b, a), sub(p, a)), n) >= tol && dot(cross(sub(c, b), sub(p, b)), n) >= tol &&
         dot(cross(sub(a, c), sub(p, c)), n) >= tol;
}

__device__ double solid_angle(D3 p, D3 a, D3 b, D3 c) {
  const D3 ra = sub(a, p), rb = sub(b, p), rc = sub(c, p);
  const double la = norm(ra), lb = norm(rb), lc = norm(rc);
  const double num = dot(ra, cross(rb, rc));
  const double den = la * lb * lc + dot(ra, rb) * lc + dot(rb, rc) * la + dot(rc, ra) * lb;
  return 2.0 * atan2(num, den);
}

constexpr int kBlock = 256;
constexpr int kFaceTile = 256;  // faces staged in LDS per pass

__device__ __forceinline__ D3 vert(const float* v, int i) {
  return {(double)v[3 * i], (double)v[3 * i + 1], (double)v[3 * i + 2]};
}

// One thread per node.  Faces are walked in the host's order (on-face pass
// first, then the winding sum in face order), staged through LDS.
__global__ void __launch_bounds__(kBlock)
    k_voxelize(Grid g, float spacing, const float* __restrict__ verts,
               const int32_t* __restrict__ faces, int32_t nf, float3 lo, float3 hi,
               uint8_t* __restrict__ inside) {
  __shared__ D3 tri[kFaceTile][3];
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  bool active = n < g.n;
  D3 p{0.0, 0.0, 0.0};
  if (active) {
    const uint32_t t = g.dz.div(n);
    const uint32_t z = n - t * g.nz, x = g.dy.div(t), y = t - x * g.ny;
    // the host forms the position in float, then widens it
    const float px = (float)x * spacing, py = (float)y * spacing, pz = (float)z * spacing;
    p = {(double)px, (double)py, (double)pz};
    if (px < lo.x || py < lo.y || pz < lo.z || px > hi.x || py > hi.y || pz > hi.z) {
      inside[n] = 0;
      active = false;
    }
  }
  const bool any = __syncthreads_or(active);
  if (!any) return;
  bool on = false;
  double omega = 0.0;
  // pass 0: on a face?  pass 1: winding number (only if on no face)
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1 && !__syncthreads_or(active && !on)) break;
    for (int32_t f0 = 0; f0 < nf; f0 += kFaceTile) {
      const int cnt = min(kFaceTile, nf - f0);
      __syncthreads();
      for (int k = threadIdx.x; k < cnt; k += kBlock) {
        tri[k][0] = vert(verts, faces[3 * (f0 + k)]);
        tri[k][1] = vert(verts, faces[3 * (f0 + k) + 1]);
        tri[k][2] = vert(verts, faces[3 * (f0 + k) + 2]);
      }
      __syncthreads();
      if (!active) continue;
      if (pass == 0) {
        for (int k = 0; k < cnt && !on; ++k) on = on_triangle(p, tri[k][0], tri[k][1], tri[k][2]);
      } else if (!on) {
        for (int k = 0; k < cnt; ++k) omega += solid_angle(p, tri[k][0], tri[k][1], tri[k][2]);
      }
    }
  }
  if (active) {
    const double winding = omega / (4.0 * M_PI);
    inside[n] = (on || fabs(winding) > 0.5) ? 1 : 0;
  }
}

// ---------------------------------------------------------------------
// lbm_build_reflections: the reflection tables of preProcessModel
// (reference Lattice.cpp:323-426, 450-477) on the GPU, one thread per
// internal node, with the host code's float / double operations in the
// same order (no contraction; sqrt and division correctly rounded); the
// sine and cosine of the half turn are computed on the host.
//
// For each link k of an internal node x whose source x + e_k is in the box
// and not internal: v = reflectionVector(h (x + e_k), h x), i.e. the unit
// vector from x to x + e_k rotated by pi about the normal of the first
// face hit by the ray (x + e_k) -> x (ties within 1e-9 of the nearest hit:
// lowest face index); rank the 27 directions by 2 - |v - e_j/|e_j|| (equal
// scores: higher index first); keep the first three whose target x + e_d is
// in the box and not internal, in slots INVERSE[k]*3 + rank; normalise the
// kept weights by their float sum (a kept rest direction, d = 0, keeps its
// weight as the host code does).

struct HalfTurn {
  float c, s, ci, si;  // cos, sin of +pi/2 and -pi/2 in float
};

__device__ double ray_hit(D3 o, D3 d, D3 a, D3 b, D3 c) {
  const D3 n = cross(sub(b, a), sub(c, a));
  const double denom = dot(n, d);
  const double scale = norm(n) * norm(d);
  if (scale == 0.0 || fabs(denom) <= 1e-14 * scale) return -1.0;
  const double t = dot(n, sub(a, o)) / denom;
  if (t < 0.0) return -1.0;
  const D3 q{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
  const double tol = -1e-12 * dot(n, n);
  if (dot(cross(sub(b, a), sub(q, a)), n) < tol || dot(cross(sub(c, b), sub(q, b)), n) < tol ||
      dot(cross(sub(a, c), sub(q, c)), n) < tol)
    return -1.0;
  return t;
}

struct F3 {
  float x, y, z;
};

// Model::reflectionVector(p1, p2)
__device__ F3 reflection_vector(F3 p1, F3 p2, const float* __restrict__ verts,
                                const int32_t* __restrict__ faces, int32_t nf,
                                const HalfTurn& ht) {
  F3 v{p1.x - p2.x, p1.y - p2.y, p1.z - p2.z};
  {
    const float nrm = sqrtf((v.x * v.x) + (v.y * v.y) + (v.z * v.z));
    v.x /= nrm;
    v.y /= nrm;
    v.z /= nrm;
  }
  const D3 o{(double)p1.x, (double)p1.y, (double)p1.z};
  const D3 d{(double)p2.x - p1.x, (double)p2.y - p1.y, (double)p2.z - p1.z};
  int32_t best = -1;
  double best_t = INFINITY;
  for (int32_t t = 0; t < nf; ++t) {
    const double h = ray_hit(o, d, vert(verts, faces[3 * t]), vert(verts, faces[3 * t + 1]),
                             vert(verts, faces[3 * t + 2]));
    if (h >= 0.0 && h < best_t) {
      best_t = h;
      best = t;
    }
  }
  if (best < 0) return F3{0.0f, 0.0f, 0.0f};
  for (int32_t t = 0; t < best; ++t) {
    const double h = ray_hit(o, d, vert(verts, faces[3 * t]), vert(verts, faces[3 * t + 1]),
                             vert(verts, faces[3 * t + 2]));
    if (h >= 0.0 && h <= best_t + 1e-9) {
      best = t;
      break;
    }
  }
  const float* A = verts + 3 * faces[3 * best];
  const float* B = verts + 3 * faces[3 * best + 1];
  const float* C = verts + 3 * faces[3 * best + 2];
  const float e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  const float e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
  F3 ax{(e1[1] * e2[2]) - (e1[2] * e2[1]), (e1[2] * e2[0]) - (e1[0] * e2[2]),
        (e1[0] * e2[1]) - (e1[1] * e2[0])};
  // rotateVectorAroundAxis(axis, v, pi) (Geometry.cpp:22-42)
  const float len = sqrtf((ax.x * ax.x) + (ax.y * ax.y) + (ax.z * ax.z));
  ax.x /= len;
  ax.y /= len;
  ax.z /= len;
  const float q[4] = {ht.c, ht.s * ax.x, ht.s * ax.y, ht.s * ax.z};
  const float qi[4] = {ht.ci, ht.si * ax.x, ht.si * ax.y, ht.si * ax.z};
  const float vv[4] = {0.0f, v.x, v.y, v.z};
  auto ham = [](const float* a, const float* b, float* r) {
    r[0] = (a[0] * b[0]) - (a[1] * b[1]) - (a[2] * b[2]) - (a[3] * b[3]);
    r[1] = (a[0] * b[1]) + (a[1] * b[0]) + (a[2] * b[3]) - (a[3] * b[2]);
    r[2] = (a[0] * b[2]) - (a[1] * b[3]) + (a[2] * b[0]) + (a[3] * b[1]);
    r[3] = (a[0] * b[3]) + (a[1] * b[2]) - (a[2] * b[1]) + (a[3] * b[0]);
  };
  float t1[4], r[4];
  ham(q, vv, t1);
  ham(t1, qi, r);
  return F3{r[1], r[2], r[3]};
}

__global__ void __launch_bounds__(kBlock)
    k_build_reflections(Grid g, float h, const uint8_t* __restrict__ flags,
                        const int32_t* __restrict__ internal_nodes, int32_t ninternal,
                        const float* __restrict__ verts, const int32_t* __restrict__ faces,
                        int32_t nf, HalfTurn ht, uint8_t* __restrict__ has,
                        uint32_t* __restrict__ dirs, float* __restrict__ weights) {
  using L = Lat<27>;
  const int32_t r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= ninternal) return;
  const uint32_t n = (uint32_t)internal_nodes[r];
  const uint32_t tq = g.dz.div(n);
  const int z = (int)(n - tq * g.nz), x = (int)g.dy.div(tq), y = (int)(tq - (uint32_t)x * g.ny);
  auto in_box = [&](int a, int b, int c) {
    return a >= 0 && b >= 0 && c >= 0 && a < (int)g.nx && b < (int)g.ny && c < (int)g.nz;
  };
  auto internal = [&](int a, int b, int c) {
    return (flags[(uint32_t)c + g.nz * ((uint32_t)b + g.ny * (uint32_t)a)] & LBM_FLAG_INTERNAL) != 0;
  };
  uint32_t* W = dirs + 14 * (size_t)r;
  float* Wt = weights + 81 * (size_t)r;
  for (int i = 0; i < 14; ++i) W[i] = 0u;
  for (int i = 0; i < 81; ++i) Wt[i] = 0.0f;
  bool any = false;
  for (int k = 0; k < 27; ++k) {
    const int sx = x + L::E[k][0], sy = y + L::E[k][1], sz = z + L::E[k][2];
    if (!in_box(sx, sy, sz) || internal(sx, sy, sz)) continue;
    const F3 v = reflection_vector(F3{h * (float)sx, h * (float)sy, h * (float)sz},
                                   F3{h * (float)x, h * (float)y, h * (float)z}, verts, faces,
                                   nf, ht);
    // distribute_vector: scores of the 27 directions
    float score[27];
    for (int j = 0; j < 27; ++j) {
      const int len2 = L::E[j][0] * L::E[j][0] + L::E[j][1] * L::E[j][1] + L::E[j][2] * L::E[j][2];
      const double inv = len2 == 0 ? 0.0
                                   : (len2 == 1 ? 1.0
                                                : (len2 == 2 ? 1.0 / 1.4142135623730951
                                                             : 1.0 / 1.7320508075688772));
      float unit[3];
      for (int a = 0; a < 3; ++a) unit[a] = L::E[j][a] == 0 ? 0.0f : (float)(L::E[j][a] * inv);
      const float dx = v.x - unit[0], dy = v.y - unit[1], dz = v.z - unit[2];
      score[j] = (float)(2.0 - (double)sqrtf((dx * dx) + (dy * dy) + (dz * dz)));
    }
    const int base = L::INV[k] * 3;
    int counter = 0;
    float total = 0.0f;
    uint32_t taken = 0u;
    // the ranked list in order: highest score first, equal scores by
    // higher index first
    for (int rank = 0; rank < 27 && counter < 3; ++rank) {
      int best = -1;
      for (int j = 26; j >= 0; --j) {
        if ((taken >> j) & 1u) continue;
        if (best < 0 || score[j] > score[best]) best = j;
      }
      taken |= 1u << best;
      const int tx = x + L::E[best][0], ty = y + L::E[best][1], tz = z + L::E[best][2];
      if (!in_box(tx, ty, tz) || internal(tx, ty, tz)) continue;
      const int slot = base + counter;
      const uint32_t shift = 5u * (uint32_t)(slot % 6);
      W[slot / 6] = (W[slot / 6] & ~(0x1Fu << shift)) | (((uint32_t)best & 0x1Fu) << shift);
      Wt[slot] = score[best];
      total += score[best];
      ++counter;
      any = true;
    }
    if (total != 0.0f)
      for (int i = 0; i < 3; ++i) {
        const int slot = base + i;
        if (((W[slot / 6] >> (5u * (uint32_t)(slot % 6))) & 0x1Fu) != 0u) Wt[slot] /= total;
      }
  }
  has[r] = any ? 1 : 0;
}

}  // namespace
}  // namespace lbm

// Entry points of include/lbm.h (exported by lbm_exports.hip).
namespace lbm {
namespace api {

int lbm_build_reflections(lbm_ctx* c, const float* vertices, int32_t nv,
                          const int32_t* faces, int32_t nf, float spacing,
                          int32_t max_records, int32_t* node, uint32_t* dirs,
                          float* weights, int32_t* count) {
  if (!c) return fail(LBM_ERR_INVALID, "null context");
  TraceRange tr("lbm_build_reflections");
  if (c->cfg.mode != LBM_MODE_REFERENCE)
    return fail(LBM_ERR_STATE, "lbm_build_reflections needs a reference-mode lattice");
  if (!count || nv < 0 || nf < 0 || max_records < 0 || (nf > 0 && (!vertices || !faces)) ||
      (max_records > 0 && (!node || !dirs || !weights)))
    return fail(LBM_ERR_INVALID, "bad arguments");
  for (int64_t k = 0; k < 3 * (int64_t)nf; ++k)
    if (faces[k] < 0 || faces[k] >= nv) return fail(LBM_ERR_INVALID, "face index out of range");
  *count = 0;
  std::vector<int32_t> internal_nodes;
  for (uint32_t n = 0; n < c->grid.n; ++n)
    if (c->host_flags[n] & LBM_FLAG_INTERNAL) internal_nodes.push_back((int32_t)n);
  const int32_t ni = (int32_t)internal_nodes.size();
  if (ni == 0) return LBM_OK;
  // the host's half turn: angle pi in float, halved in double, cos / sin
  // of the float overloads
  const float angle = (float)M_PI;
  const float half = (float)((double)angle / 2.0);
  const HalfTurn ht{std::cos(half), std::sin(half), std::cos(-half), std::sin(-half)};
  float* dv = nullptr;
  int32_t *df = nullptr, *dnodes = nullptr;
  uint8_t* dhas = nullptr;
  uint32_t* ddirs = nullptr;
  float* dw = nullptr;
  auto cleanup = [&]() {
    for (void* p : {(void*)dv, (void*)df, (void*)dnodes, (void*)dhas, (void*)ddirs, (void*)dw})
      if (p) (void)hipFree(p);
  };
  hipError_t e;
  if ((e = hipMalloc((void**)&dv, sizeof(float) * 3 * (size_t)std::max(nv, 1))) != hipSuccess ||
      (e = hipMalloc((void**)&df, sizeof(int32_t) * 3 * (size_t)std::max(nf, 1))) != hipSuccess ||
      (e = hipMalloc((void**)&dnodes, sizeof(int32_t) * (size_t)ni)) != hipSuccess ||
      (e = hipMalloc((void**)&dhas, (size_t)ni)) != hipSuccess ||
      (e = hipMalloc((void**)&ddirs, sizeof(uint32_t) * 14 * (size_t)ni)) != hipSuccess ||
      (e = hipMalloc((void**)&dw, sizeof(float) * 81 * (size_t)ni)) != hipSuccess) {
    cleanup();
    return hip_fail(e, "hipMalloc(reflection build)");
  }
  if ((nv > 0 && (e = hipMemcpyAsync(dv, vertices, sizeof(float) * 3 * (size_t)nv,
                                     hipMemcpyHostToDevice, c->stream)) != hipSuccess) ||
      (nf > 0 && (e = hipMemcpyAsync(df, faces, sizeof(int32_t) * 3 * (size_t)nf,
                                     hipMemcpyHostToDevice, c->stream)) != hipSuccess) ||
      (e = hipMemcpyAsync(dnodes, internal_nodes.data(), sizeof(int32_t) * (size_t)ni,
                          hipMemcpyHostToDevice, c->stream)) != hipSuccess) {
    cleanup();
    return hip_fail(e, "hipMemcpyAsync(reflection build)");
  }
  hipLaunchKernelGGL(k_build_reflections, dim3((ni + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     c->stream, c->grid, spacing, c->flags, dnodes, ni, dv, df, nf, ht, dhas,
                     ddirs, dw);
  std::vector<uint8_t> has((size_t)ni);
  std::vector<uint32_t> hd(14 * (size_t)ni);
  std::vector<float> hw(81 * (size_t)ni);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(has.data(), dhas, (size_t)ni, hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(hd.data(), ddirs, sizeof(uint32_t) * hd.size(), hipMemcpyDeviceToHost,
                          c->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(hw.data(), dw, sizeof(float) * hw.size(), hipMemcpyDeviceToHost,
                          c->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess) {
    cleanup();
    return hip_fail(e, "reflection build");
  }
  cleanup();
  int32_t nrec = 0;
  for (int32_t r = 0; r < ni; ++r) nrec += has[r];
  *count = nrec;
  if (nrec > max_records)
    return fail(LBM_ERR_INVALID, "max_records is smaller than the number of records");
  int32_t k = 0;
  for (int32_t r = 0; r < ni; ++r) {
    if (!has[r]) continue;
    node[k] = internal_nodes[r];
    std::copy(hd.begin() + 14 * (size_t)r, hd.begin() + 14 * (size_t)(r + 1), dirs + 14 * (size_t)k);
    std::copy(hw.begin() + 81 * (size_t)r, hw.begin() + 81 * (size_t)(r + 1), weights + 81 * (size_t)k);
    ++k;
  }
  return LBM_OK;
}

int lbm_voxelize_mesh(lbm_ctx* c, const float* vertices, int32_t nv,
                      const int32_t* faces, int32_t nf, float spacing,
                      uint8_t* inside, int64_t* count) {
  if (!c) return fail(LBM_ERR_INVALID, "null context");
  TraceRange tr("lbm_voxelize_mesh");
  if (!inside || nv < 0 || nf < 0 || (nf > 0 && (!vertices || !faces)))
    return fail(LBM_ERR_INVALID, "bad mesh arguments");
  for (int64_t k = 0; k < 3 * (int64_t)nf; ++k)
    if (faces[k] < 0 || faces[k] >= nv) return fail(LBM_ERR_INVALID, "face index out of range");
  const Grid& g = c->grid;
  float3 lo{0.f, 0.f, 0.f}, hi{-1.f, -1.f, -1.f};  // empty box: nothing inside
  if (nv > 0) {
    lo = {vertices[0], vertices[1], vertices[2]};
    hi = lo;
    for (int32_t v = 1; v < nv; ++v) {
      lo.x = std::fmin(lo.x, vertices[3 * v]);
      hi.x = std::fmax(hi.x, vertices[3 * v]);
      lo.y = std::fmin(lo.y, vertices[3 * v + 1]);
      hi.y = std::fmax(hi.y, vertices[3 * v + 1]);
      lo.z = std::fmin(lo.z, vertices[3 * v + 2]);
      hi.z = std::fmax(hi.z, vertices[3 * v + 2]);
    }
    const float m = 1e-9f * std::fmax(std::fmax(hi.x - lo.x, hi.y - lo.y),
                                      std::fmax(hi.z - lo.z, 1e-30f));
    lo = {lo.x - m, lo.y - m, lo.z - m};
    hi = {hi.x + m, hi.y + m, hi.z + m};
  }
  if (nf == 0) {
    for (uint32_t n = 0; n < g.n; ++n) inside[n] = 0;
    if (count) *count = 0;
    return LBM_OK;
  }
  float* dv = nullptr;
  int32_t* df = nullptr;
  uint8_t* dout = nullptr;
  auto cleanup = [&]() {
    if (dv) (void)hipFree(dv);
    if (df) (void)hipFree(df);
    if (dout) (void)hipFree(dout);
  };
  hipError_t e;
  if ((e = hipMalloc((void**)&dv, sizeof(float) * 3 * (size_t)nv)) != hipSuccess ||
      (e = hipMalloc((void**)&df, sizeof(int32_t) * 3 * (size_t)nf)) != hipSuccess ||
      (e = hipMalloc((void**)&dout, (size_t)g.n)) != hipSuccess) {
    cleanup();
    return hip_fail(e, "hipMalloc(voxelisation)");
  }
  if ((e = hipMemcpyAsync(dv, vertices, sizeof(float) * 3 * (size_t)nv, hipMemcpyHostToDevice,
                          c->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(df, faces, sizeof(int32_t) * 3 * (size_t)nf, hipMemcpyHostToDevice,
                          c->stream)) != hipSuccess) {
    cleanup();
    return hip_fail(e, "hipMemcpyAsync(mesh)");
  }
  hipLaunchKernelGGL(k_voxelize, dim3((g.n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, g,
                     spacing, dv, df, nf, lo, hi, dout);
  if ((e = hipGetLastError()) != hipSuccess ||
      (e = hipMemcpyAsync(inside, dout, (size_t)g.n, hipMemcpyDeviceToHost, c->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(c->stream)) != hipSuccess) {
    cleanup();
    return hip_fail(e, "voxelisation");
  }
  cleanup();
  if (count) {
    int64_t k = 0;
    for (uint32_t n = 0; n < g.n; ++n) k += inside[n];
    *count = k;
  }
  return LBM_OK;
}

}  // namespace api
}  // namespace lbm
