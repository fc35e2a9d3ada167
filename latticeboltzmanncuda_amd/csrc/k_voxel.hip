// Mesh voxelisation on the GPU (C-ABI lbm_voxelize_mesh): which lattice
// nodes lie inside a closed triangle mesh.  It replaces the host loop of
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

#include <cmath>
#include <vector>

#include "lbm_internal.h"

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

}  // namespace
}  // namespace lbm

using namespace lbm;

extern "C" int lbm_voxelize_mesh(lbm_ctx* c, const float* vertices, int32_t nv,
                                 const int32_t* faces, int32_t nf, float spacing,
                                 uint8_t* inside, int64_t* count) {
  if (!c) return fail(LBM_ERR_INVALID, "null context");
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
