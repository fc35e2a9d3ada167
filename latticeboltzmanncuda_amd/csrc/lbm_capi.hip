// C-ABI of liblbm_hip.so: context lifetime, host <-> device exchange,
// geometry upload, stage and step launches, readback.  Every call checks
// its arguments and HIP return codes and reports failures through an
// lbm_status plus a thread-local message (lbm_last_error).
//
// Host <-> device copies run on the context's stream and are followed by a
// synchronisation of that stream; step and stage launches are asynchronous.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "lbm_internal.h"
#include "lbm_lattice.h"
#include "lbm_layout.h"
#include "lbm_store.h"

namespace lbm {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? LBM_ERR_NOMEM : LBM_ERR_HIP;
}

int api_exception() noexcept {
  int code = LBM_ERR_INVALID;
  const char* msg = "internal error: unknown exception";
  try {
    throw;
  } catch (const std::bad_alloc&) {
    code = LBM_ERR_NOMEM;
    msg = "host allocation failed";
  } catch (const std::exception& e) {
    msg = e.what();
  } catch (...) {
  }
  try {
    g_last_error = msg;
  } catch (...) {
    // no memory for the message either: keep the previous one
  }
  return code;
}

namespace {

// Fills the per-direction constants of lattice L in precision C.  The
// products are formed in C from the configured float values, as the
// kernels formed them before they were hoisted.
template <class L, class C>
void fill_coef(const StdParams& p, StepCoef<C>& k) {
  const C omega = (C)1 / (C)p.tau;
  const C guo = (C)1 - (C)0.5 * omega;
  k.omega = omega;
  for (int a = 0; a < 3; ++a) {
    k.force[a] = (C)p.force[a];
    k.half_force[a] = (C)0.5 * (C)p.force[a];
  }
  k.forced = p.force[0] != 0.0f || p.force[1] != 0.0f || p.force[2] != 0.0f;
  for (int i = 0; i < L::Q; ++i) {
    const int* e = L::E[i];
    const C w = (C)L::W[i];
    const C eF = (C)e[0] * k.force[0] + (C)e[1] * k.force[1] + (C)e[2] * k.force[2];
    k.gw[i] = guo * w;
    k.gf3[i] = k.gw[i] * ((C)3 * eF);
    k.gf9[i] = k.gw[i] * ((C)9 * eF);
    k.lid[i] = (C)0;
    if (p.lid_axis >= 0 && e[p.lid_axis] == -p.lid_sign) {
      // a link of direction i whose source node lies beyond the moving face
      const C eu = (C)e[0] * (C)p.lid_u[0] + (C)e[1] * (C)p.lid_u[1] + (C)e[2] * (C)p.lid_u[2];
      k.lid[i] = (C)6 * w * (C)p.wall_rho * eu;
    }
  }
}

// cf in FP16 stored units: exact, as the scale is a power of two
template <class L>
void fill_coef_units(const StepCoef<float>& k, float scale, StepCoef<float>& u) {
  u = k;
  for (int a = 0; a < 3; ++a) u.half_force[a] = k.half_force[a] * scale;
  for (int i = 0; i < L::Q; ++i) {
    u.lid[i] = k.lid[i] * scale;
    u.gw[i] = k.gw[i] * scale;
    u.gf3[i] = k.gf3[i] * scale;
    u.gf9[i] = k.gf9[i] * scale;
  }
}

template <class L>
void fill_coefs(StdParams& p) {
  fill_coef<L, float>(p, p.cf);
  fill_coef<L, double>(p, p.cd);
  fill_coef_units<L>(p.cf, 1.0f / Store<LBM_STORE_F16>::kUnit, p.ch);
}

}  // namespace

StdParams make_std_params(const lbm_config& cfg) {
  StdParams p{};
  p.tau = cfg.tau;
  for (int a = 0; a < 3; ++a) {
    p.force[a] = cfg.force[a];
    p.lid_u[a] = cfg.lid_u[a];
  }
  p.lid_axis = cfg.lid_face >= 0 ? cfg.lid_face / 2 : -1;
  p.lid_sign = cfg.lid_face >= 0 ? (cfg.lid_face % 2 ? 1 : -1) : 0;
  p.wall_rho = cfg.wall_rho;
  switch (cfg.lattice) {
    case 9:
      fill_coefs<Lat<9>>(p);
      break;
    case 19:
      fill_coefs<Lat<19>>(p);
      break;
    default:
      fill_coefs<Lat<27>>(p);
      break;
  }
  return p;
}

double cell_reynolds(const lbm_config& cfg) {
  const double nu = ((double)cfg.tau - 0.5) / 3.0;
  const double uw = std::sqrt((double)cfg.lid_u[0] * cfg.lid_u[0] +
                              (double)cfg.lid_u[1] * cfg.lid_u[1] +
                              (double)cfg.lid_u[2] * cfg.lid_u[2]);
  const double F = std::sqrt((double)cfg.force[0] * cfg.force[0] +
                             (double)cfg.force[1] * cfg.force[1] +
                             (double)cfg.force[2] * cfg.force[2]);
  // narrowest walled extent (axes of one node carry no flow)
  const int32_t ext[3] = {cfg.nx, cfg.ny, cfg.nz};
  int32_t walled = 0, widest = 1;
  for (int a = 0; a < 3; ++a) {
    if (ext[a] < 2) continue;
    widest = std::max(widest, ext[a]);
    if (!cfg.periodic[a]) walled = walled ? std::min(walled, ext[a]) : ext[a];
  }
  const double W = walled ? walled : widest;
  const double U = std::max(cfg.lid_face >= 0 ? uw : 0.0, F * W * W / (8.0 * nu));
  return U / nu;
}

int64_t estimate_ctx_bytes(const lbm_config& cfg) {
  const uint64_t n = (uint64_t)cfg.nx * cfg.ny * cfg.nz;
  int64_t b = 2 * (int64_t)(population_elems(cfg.lattice, n) * storage_bytes(cfg.storage));
  if (cfg.mode == LBM_MODE_STANDARD)
    b += (has_mask16(cfg) ? 6 : 4) * (int64_t)n + 4 * (int64_t)((n + 15) / 16);
  else
    b += (int64_t)n;
  b += 4 * 1024 * 8 * 4 + 64;  // reduction scratch, force register
  return b;
}

namespace {

// Scoped device allocation for staging buffers.
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
      p = nullptr;
      return hip_fail(e, "hipMalloc(staging)");
    }
    return LBM_OK;
  }
};

int check_ctx(const lbm_ctx* c) {
  if (!c) return fail(LBM_ERR_INVALID, "null context");
  return LBM_OK;
}

int need_mode(const lbm_ctx* c, int mode, const char* what) {
  if (c->cfg.mode != mode)
    return fail(LBM_ERR_STATE,
                std::string(what) + " needs " +
                    (mode == LBM_MODE_REFERENCE ? "reference" : "standard") +
                    " mode");
  return LBM_OK;
}

int dev_alloc(lbm_ctx* c, void** p, size_t bytes, const char* what) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    *p = nullptr;
    return hip_fail(e, what);
  }
  c->device_bytes += (int64_t)bytes;
  return LBM_OK;
}

void dev_free(lbm_ctx* c, void** p, size_t bytes) {
  if (*p) {
    (void)hipFree(*p);
    c->device_bytes -= (int64_t)bytes;
  }
  *p = nullptr;
}

int sync(lbm_ctx* c) {
  LBM_HIP(hipStreamSynchronize(c->stream));
  return LBM_OK;
}

// Solid nodes are never written by a standard step, so after a host
// upload both population buffers are made equal.
int mirror_buffers(lbm_ctx* c) {
  const size_t fbytes = population_elems(c->q, c->grid.n) * c->elem;
  LBM_HIP(hipMemcpyAsync(c->f[c->cur ^ 1], c->f[c->cur], fbytes,
                         hipMemcpyDeviceToDevice, c->stream));
  return LBM_OK;
}

int h2d(lbm_ctx* c, void* dst, const void* src, size_t bytes) {
  LBM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return sync(c);
}

int d2h(lbm_ctx* c, void* dst, const void* src, size_t bytes) {
  LBM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

int validate_config(const lbm_config* cfg) {
  if (!cfg) return fail(LBM_ERR_INVALID, "null config");
  if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1)
    return fail(LBM_ERR_INVALID, "lattice extents must be >= 1");
  // the kernels multiply coordinates and extents with 24-bit multiplies
  if (cfg->nx >= (1 << 24) || cfg->ny >= (1 << 24) || cfg->nz >= (1 << 24))
    return fail(LBM_ERR_UNSUPPORTED, "lattice extents must be below 2^24");
  const uint64_t n = (uint64_t)cfg->nx * cfg->ny * cfg->nz;
  if (cfg->lattice != 9 && cfg->lattice != 19 && cfg->lattice != 27)
    return fail(LBM_ERR_INVALID, "lattice must be 9, 19 or 27");
  if (population_elems(cfg->lattice, n) >= (1ull << 31))
    return fail(LBM_ERR_UNSUPPORTED,
                "lattice too large for 32-bit population indices");
  if (cfg->lattice == 9 && cfg->nz != 1)
    return fail(LBM_ERR_INVALID, "D2Q9 needs nz == 1");
  if (storage_bytes(cfg->storage) == 0)
    return fail(LBM_ERR_INVALID, "unknown storage type");
  if (cfg->mode != LBM_MODE_STANDARD && cfg->mode != LBM_MODE_REFERENCE)
    return fail(LBM_ERR_INVALID, "unknown mode");
  if (cfg->mode == LBM_MODE_REFERENCE &&
      (cfg->lattice != 27 || cfg->storage != LBM_STORE_F32))
    return fail(LBM_ERR_UNSUPPORTED, "reference mode is D3Q27 with F32 storage");
  if (!(cfg->tau > 0.0f))
    return fail(LBM_ERR_INVALID, "tau must be positive");
  if (cfg->mode == LBM_MODE_STANDARD && !(cfg->tau > 0.5f))
    return fail(LBM_ERR_INVALID, "standard BGK needs tau > 1/2");
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(cfg->force[a]) || !std::isfinite(cfg->lid_u[a]))
      return fail(LBM_ERR_INVALID, "body force and wall velocity must be finite");
  if (!std::isfinite(cfg->wall_rho)) return fail(LBM_ERR_INVALID, "wall density must be finite");
  if (cfg->lattice == 9 && cfg->force[2] != 0.0f)
    return fail(LBM_ERR_INVALID, "D2Q9 has no z axis: the body force's z component must be 0");
  if (cfg->lid_face < -1 || cfg->lid_face > 5)
    return fail(LBM_ERR_INVALID, "lid_face must be -1..5");
  if (cfg->lid_face >= 0) {
    const int ax = cfg->lid_face / 2;
    if (cfg->periodic[ax])
      return fail(LBM_ERR_INVALID, "moving wall on a periodic axis");
    if (cfg->lattice == 9 && ax == 2)
      return fail(LBM_ERR_INVALID, "D2Q9 has no z faces");
  }
  return LBM_OK;
}

// Tables of the fused reference step.  Every slot s (1..80) of a
// reflecting node x with direction d (1..26) whose target t = x + e_d lies
// in the box becomes an entry of group (t, d); a group's entries keep
// ascending s, the order in which the staged reflection pass adds them, and
// groups are ordered by (t, d).  Targets must not be flagged: the staged
// pass would otherwise race between adding to a target and zeroing it.
int rebuild_receivers(lbm_ctx* c, const std::vector<int32_t>& rec_of) {
  const Grid& g = c->grid;
  const int64_t N = g.n;
  const size_t words = (size_t)((N + 15) / 16);
  const size_t ng = (size_t)std::max<int64_t>(c->ref_ngrp, 1), ne = (size_t)std::max<int64_t>(c->ref_nent, 1);
  dev_free(c, (void**)&c->ref_cls, sizeof(uint32_t) * words);
  const size_t kwords = (size_t)((N + 64 * 32 - 1) / (64 * 32));
  dev_free(c, (void**)&c->ref_keep, sizeof(uint32_t) * kwords);
  dev_free(c, (void**)&c->ref_inmask, sizeof(uint32_t) * (size_t)N);
  dev_free(c, (void**)&c->ref_rxbase, sizeof(int32_t) * (size_t)N);
  dev_free(c, (void**)&c->ref_gsrc, sizeof(uint32_t) * ng);
  dev_free(c, (void**)&c->ref_gbeg, sizeof(uint32_t) * ng);
  dev_free(c, (void**)&c->ref_gid, sizeof(uint32_t) * ng);
  dev_free(c, (void**)&c->ref_gsum, sizeof(float) * ng);
  dev_free(c, (void**)&c->ref_eidx, sizeof(uint32_t) * ne);
  dev_free(c, (void**)&c->ref_ew, sizeof(float) * ne);
  dev_free(c, (void**)&c->ref_es, sizeof(uint8_t) * ne);
  dev_free(c, (void**)&c->ref_partials, sizeof(float) * 3 * (size_t)std::max(c->ref_partials_cap, 1));
  c->ref_ngrp = c->ref_nent = 0;
  c->ref_partials_cap = 0;
  c->ref_geom = c->nspecial > 0;
  if (!c->ref_geom) return LBM_OK;

  struct Link {
    uint32_t t;   // receiving node
    uint8_t d;    // direction of the reflected population
    uint8_t s;    // slot of the reflecting node's record
    int32_t r;    // record
  };
  const int64_t plane = (int64_t)g.ny * g.nz;
  auto at = [&](int64_t x, int64_t y, int64_t z) { return z + (int64_t)g.nz * (y + (int64_t)g.ny * x); };
  std::vector<Link> links;
  std::vector<uint32_t> cls(words, 0u);
  for (int64_t n = 0; n < N; ++n) {
    const uint8_t fl = c->host_flags[n];
    if (fl & LBM_FLAG_INTERNAL) cls[n >> 4] |= 2u << (2 * (n & 15));
    const int32_t r = c->nrec > 0 ? rec_of[n] : -1;
    if (!(fl & LBM_FLAG_REFLECTED) || r < 0) continue;
    const int64_t x = n / plane, y = (n % plane) / g.nz, z = n % g.nz;
    const uint32_t* w = c->host_rec_dirs.data() + 14 * (size_t)r;
    for (int s = 1; s < 81; ++s) {
      const int d = (int)((w[s / 6] >> (5 * (s % 6))) & 0x1Fu);
      if (d == 0 || d > 26) continue;
      const int64_t tx = x + Lat<27>::E[d][0], ty = y + Lat<27>::E[d][1], tz = z + Lat<27>::E[d][2];
      if (tx < 0 || ty < 0 || tz < 0 || tx >= g.nx || ty >= g.ny || tz >= g.nz) continue;
      const int64_t t = at(tx, ty, tz);
      if (c->host_flags[t])
        return fail(LBM_ERR_INVALID, "reflection target " + std::to_string(t) +
                                         " is a flagged node (targets must be fluid)");
      links.push_back({(uint32_t)t, (uint8_t)d, (uint8_t)s, r});
    }
  }
  std::stable_sort(links.begin(), links.end(), [](const Link& a, const Link& b) {
    return a.t != b.t ? a.t < b.t : (a.d != b.d ? a.d < b.d : a.s < b.s);
  });
  std::vector<uint32_t> inmask((size_t)N, 0u), gsrc, gbeg, eidx;
  std::vector<int32_t> base((size_t)N, -1);
  std::vector<float> ew;
  std::vector<uint8_t> es;
  for (size_t i = 0; i < links.size();) {
    const uint32_t t = links[i].t;
    base[t] = (int32_t)gsrc.size();
    cls[t >> 4] |= 1u << (2 * (t & 15));
    const int64_t tx = t / plane, ty = (t % plane) / g.nz, tz = t % g.nz;
    while (i < links.size() && links[i].t == t) {
      const int d = links[i].d;
      inmask[t] |= 1u << d;
      // the reflecting node x = t - e_d (inside the box: it holds a record)
      const int64_t xx = tx - Lat<27>::E[d][0], xy = ty - Lat<27>::E[d][1], xz = tz - Lat<27>::E[d][2];
      gsrc.push_back(pidx<27>((uint32_t)at(xx, xy, xz), d));
      gbeg.push_back((uint32_t)eidx.size());
      for (; i < links.size() && links[i].t == t && links[i].d == d; ++i) {
        const int s = links[i].s, k = s / 3;
        // x's streamed value of direction k: A_k(x - e_k), 0 outside
        const int64_t sx = xx - Lat<27>::E[k][0], sy = xy - Lat<27>::E[k][1], sz = xz - Lat<27>::E[k][2];
        const bool in = sx >= 0 && sy >= 0 && sz >= 0 && sx < g.nx && sy < g.ny && sz < g.nz;
        eidx.push_back(in ? pidx<27>((uint32_t)at(sx, sy, sz), k) : ~0u);
        ew.push_back(c->host_rec_w[81 * (size_t)links[i].r + (size_t)s]);
        es.push_back((uint8_t)s);
      }
    }
  }
  if (eidx.size() >= (1ull << 31)) return fail(LBM_ERR_UNSUPPORTED, "too many reflected links");
  // The gather's threads take the groups by reflecting node x, then
  // direction (gsrc = pidx(x, d) gives both): the groups of one x read the
  // same neighbours of x (up to three slots share an incoming direction),
  // so a wavefront touches fewer cache lines than in receiver order (on
  // the 128^3 angled cube 193 k lines over all wavefronts against 295 k).
  // Each thread writes its sum to its group's slot; the tables follow the
  // thread order.
  const size_t ngs = gsrc.size();
  std::vector<uint32_t> gid(ngs);
  for (size_t q = 0; q < ngs; ++q) gid[q] = (uint32_t)q;
  auto xd = [&](uint32_t q) {
    const uint32_t e = gsrc[q];
    const uint64_t x = (uint64_t)(e / (27u * kTile)) * kTile + e % kTile;
    return x * 32u + (e / kTile) % 27u;
  };
  std::stable_sort(gid.begin(), gid.end(), [&](uint32_t a, uint32_t b) { return xd(a) < xd(b); });
  {
    std::vector<uint32_t> tsrc(ngs), tbeg(ngs), te(eidx.size());
    std::vector<float> tw(ew.size());
    std::vector<uint8_t> ts(es.size());
    size_t pos = 0;
    for (size_t i = 0; i < ngs; ++i) {
      const uint32_t q = gid[i];
      const size_t b = gbeg[q], e = q + 1 < ngs ? gbeg[q + 1] : eidx.size();
      tsrc[i] = gsrc[q];
      tbeg[i] = (uint32_t)pos;
      for (size_t k = b; k < e; ++k, ++pos) {
        te[pos] = eidx[k];
        tw[pos] = ew[k];
        ts[pos] = es[k];
      }
    }
    gsrc.swap(tsrc);
    gbeg.swap(tbeg);
    eidx.swap(te);
    ew.swap(tw);
    es.swap(ts);
  }
  c->ref_ngrp = (int64_t)gsrc.size();
  c->ref_nent = (int64_t)eidx.size();
  c->ref_partials_cap = ref_gather_partials(c->ref_ngrp);
  const size_t ng2 = std::max<size_t>(gsrc.size(), 1), ne2 = std::max<size_t>(eidx.size(), 1);
  int rc;
  // tiles within two nodes of a reflecting node: the gather reads those
  // within one, and the fused pass all of them (on the 128^3 angled cube a
  // reach of 2 to 8 nodes gave 85.0-85.1 us per step, 1 node 85.6, 16
  // nodes 85.6-85.7: the caches start to lose them)
  constexpr int kKeepReach = 2;
  std::vector<uint32_t> keep(kwords, 0u);
  for (int64_t n = 0; n < N; ++n) {
    if (!(c->host_flags[n] & LBM_FLAG_REFLECTED) || (c->nrec > 0 ? rec_of[n] : -1) < 0) continue;
    const int64_t x = n / plane, y = (n % plane) / g.nz, z = n % g.nz;
    for (int dx = -kKeepReach; dx <= kKeepReach; ++dx)
      for (int dy = -kKeepReach; dy <= kKeepReach; ++dy)
        for (int dz = -kKeepReach; dz <= kKeepReach; ++dz) {
          const int64_t a = x + dx, b = y + dy, cz = z + dz;
          if (a < 0 || b < 0 || cz < 0 || a >= g.nx || b >= g.ny || cz >= g.nz) continue;
          const uint64_t tile = (uint64_t)at(a, b, cz) / 64u;
          keep[tile >> 5] |= 1u << (tile & 31u);
        }
  }
  if ((rc = dev_alloc(c, (void**)&c->ref_cls, sizeof(uint32_t) * words, "hipMalloc(ref classes)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_keep, sizeof(uint32_t) * kwords, "hipMalloc(ref tiles)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_inmask, sizeof(uint32_t) * (size_t)N, "hipMalloc(ref in-masks)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_rxbase, sizeof(int32_t) * (size_t)N, "hipMalloc(ref receivers)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_gsrc, sizeof(uint32_t) * ng2, "hipMalloc(ref groups)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_gbeg, sizeof(uint32_t) * ng2, "hipMalloc(ref groups)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_gid, sizeof(uint32_t) * ng2, "hipMalloc(ref groups)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_gsum, sizeof(float) * ng2, "hipMalloc(ref group sums)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_eidx, sizeof(uint32_t) * ne2, "hipMalloc(ref entries)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_ew, sizeof(float) * ne2, "hipMalloc(ref weights)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_es, sizeof(uint8_t) * ne2, "hipMalloc(ref slots)")) ||
      (rc = dev_alloc(c, (void**)&c->ref_partials, sizeof(float) * 3 * (size_t)std::max(c->ref_partials_cap, 1),
                      "hipMalloc(ref force partials)")))
    return rc;
  if ((rc = h2d(c, c->ref_cls, cls.data(), sizeof(uint32_t) * words)) ||
      (rc = h2d(c, c->ref_keep, keep.data(), sizeof(uint32_t) * kwords)) ||
      (rc = h2d(c, c->ref_inmask, inmask.data(), sizeof(uint32_t) * (size_t)N)) ||
      (rc = h2d(c, c->ref_rxbase, base.data(), sizeof(int32_t) * (size_t)N)))
    return rc;
  if (!gsrc.empty() &&
      ((rc = h2d(c, c->ref_gsrc, gsrc.data(), sizeof(uint32_t) * gsrc.size())) ||
       (rc = h2d(c, c->ref_gbeg, gbeg.data(), sizeof(uint32_t) * gbeg.size())) ||
       (rc = h2d(c, c->ref_gid, gid.data(), sizeof(uint32_t) * gid.size())) ||
       (rc = h2d(c, c->ref_eidx, eidx.data(), sizeof(uint32_t) * eidx.size())) ||
       (rc = h2d(c, c->ref_ew, ew.data(), sizeof(float) * ew.size())) ||
       (rc = h2d(c, c->ref_es, es.data(), sizeof(uint8_t) * es.size()))))
    return rc;
  return LBM_OK;
}

// Nodes with a nonzero flag byte and their reflection record (or -1).
int rebuild_special(lbm_ctx* c) {
  if (!c->special_dirty) return LBM_OK;
  const int64_t N = c->grid.n;
  std::vector<int32_t> rec_of;
  if (c->nrec > 0) {
    rec_of.assign((size_t)N, -1);
    for (int32_t r = 0; r < c->nrec; ++r) {
      const int32_t node = c->host_rec_node[r];
      if (rec_of[node] >= 0)
        return fail(LBM_ERR_INVALID, "two reflection records for node " +
                                         std::to_string(node));
      rec_of[node] = r;
    }
  }
  std::vector<int32_t> sp, sr;
  for (int64_t n = 0; n < N; ++n) {
    const uint8_t fl = c->host_flags[n];
    if (!fl) continue;
    const int32_t r = c->nrec > 0 ? rec_of[n] : -1;
    if ((fl & LBM_FLAG_REFLECTED) && r >= 0 && !(fl & LBM_FLAG_INTERNAL))
      return fail(LBM_ERR_INVALID,
                  "reflecting node " + std::to_string(n) +
                      " is not internal (reflection runs in place and needs "
                      "reflecting nodes inside the obstacle)");
    sp.push_back((int32_t)n);
    sr.push_back(r);
  }
  dev_free(c, (void**)&c->special, sizeof(int32_t) * (size_t)std::max(c->nspecial, 1));
  dev_free(c, (void**)&c->special_rec, sizeof(int32_t) * (size_t)std::max(c->nspecial, 1));
  dev_free(c, (void**)&c->force_partials, sizeof(float) * 3 * (size_t)std::max(c->force_partials_cap, 1));
  c->nspecial = (int32_t)sp.size();
  c->force_partials_cap = std::max(1, ref_reflect_blocks(c->nspecial));
  int rc;
  if ((rc = dev_alloc(c, (void**)&c->special, sizeof(int32_t) * std::max<size_t>(sp.size(), 1), "hipMalloc(special)")))
    return rc;
  if ((rc = dev_alloc(c, (void**)&c->special_rec, sizeof(int32_t) * std::max<size_t>(sr.size(), 1), "hipMalloc(special_rec)")))
    return rc;
  if ((rc = dev_alloc(c, (void**)&c->force_partials, sizeof(float) * 3 * (size_t)c->force_partials_cap, "hipMalloc(force partials)")))
    return rc;
  if (!sp.empty()) {
    if ((rc = h2d(c, c->special, sp.data(), sizeof(int32_t) * sp.size()))) return rc;
    if ((rc = h2d(c, c->special_rec, sr.data(), sizeof(int32_t) * sr.size()))) return rc;
  }
  if ((rc = rebuild_receivers(c, rec_of))) return rc;
  c->special_dirty = false;
  return LBM_OK;
}

int step_standard(lbm_ctx* c, int n) {
  const StdParams p = make_std_params(c->cfg);
  for (int s = 0; s < n; ++s) {
    int rc = launch_std_step(c->q, device_storage(c->cfg.mode, c->cfg.storage), c->grid, c->f[c->cur],
                             c->f[c->cur ^ 1], c->mask, c->cls, p,
                             (int)(c->steps_done & 1), c->stream);
    if (rc) return rc;
    c->cur ^= 1;
    ++c->steps_done;
  }
  return LBM_OK;
}

int step_reference_fused(lbm_ctx* c, int n) {
  for (int s = 0; s < n; ++s) {
    int rc = launch_ref_stream_collide(c->grid, (const float*)c->f[c->cur],
                                       (float*)c->f[c->cur ^ 1], c->cfg.tau,
                                       c->stream);
    if (rc) return rc;
    c->cur ^= 1;
  }
  return LBM_OK;
}

int reflect(lbm_ctx* c, int with_force) {
  int rc = rebuild_special(c);
  if (rc) return rc;
  return launch_ref_reflect(c->grid, (float*)c->f[c->cur], c->nspecial,
                            c->special, c->special_rec, c->flags, c->rec_dirs,
                            c->rec_w, c->force_partials,
                            ref_reflect_blocks(c->nspecial),
                            with_force ? c->force : nullptr, c->stream);
}

// The reference's main loop S -> R -> C -> F (LatticeBoltzmannCuda.cpp:
// 85-93), one fused pass per step.
int step_reference_full(lbm_ctx* c, int n, int with_force) {
  int rc = rebuild_special(c);
  if (rc) return rc;
  const bool geom = c->ref_geom;
  for (int s = 0; s < n; ++s) {
    if ((rc = launch_ref_step_full(c->grid, (const float*)c->f[c->cur], (float*)c->f[c->cur ^ 1],
                                   c->cfg.tau, c->flow, geom ? c->ref_cls : nullptr,
                                   c->ref_inmask, c->ref_rxbase, c->ref_ngrp, c->ref_gsrc,
                                   c->ref_gbeg, c->ref_gid, c->ref_gsum, c->ref_nent, c->ref_eidx, c->ref_keep, c->ref_ew,
                                   c->ref_es, geom && with_force ? c->ref_partials : nullptr,
                                   geom && with_force ? c->force : nullptr, c->cur, c->stream)))
      return rc;
    c->cur ^= 1;
  }
  return LBM_OK;
}

// The same loop as separate stage passes (stream A -> B, reflect in place,
// collide + flow in place): the stage kernels the fused pass must match.
int step_reference_staged(lbm_ctx* c, int n, int with_force) {
  int rc = rebuild_special(c);
  if (rc) return rc;
  for (int s = 0; s < n; ++s) {
    if ((rc = launch_ref_stream(c->grid, (const float*)c->f[c->cur],
                                (float*)c->f[c->cur ^ 1], c->stream)))
      return rc;
    c->cur ^= 1;
    if ((rc = reflect(c, with_force))) return rc;
    if ((rc = launch_ref_collide(c->grid, (float*)c->f[c->cur], c->cfg.tau,
                                 &c->flow, c->stream)))
      return rc;
  }
  return LBM_OK;
}

}  // namespace

int check_config(const lbm_config* cfg) { return validate_config(cfg); }

int read_moments(lbm_ctx* c, void* stage, float* rho, float* u) {
  const size_t N = c->grid.n;
  float* dr = rho ? (float*)stage : nullptr;
  float* du = u ? (float*)stage + (rho ? N : 0) : nullptr;
  int rc;
  if (c->cfg.mode == LBM_MODE_REFERENCE)
    rc = launch_ref_moments(c->grid, (const float*)c->f[c->cur], dr, du, c->stream);
  else
    rc = launch_std_moments(c->q, device_storage(c->cfg.mode, c->cfg.storage), c->grid,
                            c->f[c->cur], c->mask, c->cfg.force, dr, du, c->stream);
  if (rc) return rc;
  if (rho && (rc = d2h(c, rho, dr, sizeof(float) * N))) return rc;
  if (u && (rc = d2h(c, u, du, sizeof(float) * 3 * N))) return rc;
  return sync(c);
}

}  // namespace lbm

using namespace lbm;

extern "C" const char* lbm_last_error(void) { return g_last_error.c_str(); }

extern "C" int lbm_abi_version(void) { return LBM_ABI_VERSION; }

// The entry points of include/lbm.h; lbm_exports.hip exports each one
// with C linkage behind the exception boundary.
namespace lbm {
namespace api {

int lbm_device_info(int* device_count, char* name, size_t name_len) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  if (device_count) *device_count = n;
  if (name && name_len) {
    name[0] = 0;
    if (n > 0) {
      hipDeviceProp_t prop;
      LBM_HIP(hipGetDeviceProperties(&prop, 0));
      std::snprintf(name, name_len, "%s (%s, %d CUs)", prop.name,
                    prop.gcnArchName, prop.multiProcessorCount);
    }
  }
  return LBM_OK;
}

int lbm_create(const lbm_config* cfg, void* stream, lbm_ctx** out) {
  if (!out) return fail(LBM_ERR_INVALID, "null output handle");
  *out = nullptr;
  int rc = validate_config(cfg);
  if (rc) return rc;
  lbm_ctx* c = new (std::nothrow) lbm_ctx();
  if (!c) return fail(LBM_ERR_NOMEM, "host allocation of context");
  c->cfg = *cfg;
  c->q = cfg->lattice;
  c->elem = storage_bytes(cfg->storage);
  Grid& g = c->grid;
  g.nx = (uint32_t)cfg->nx;
  g.ny = (uint32_t)cfg->ny;
  g.nz = (uint32_t)cfg->nz;
  g.n = g.nx * g.ny * g.nz;
  g.dz = make_magic(g.nz);
  g.dy = make_magic(g.ny);
  g.flags = wrap_bits(cfg->periodic);
  c->flow.kind = LBM_FLOW_NONE;
  auto bail = [&](int code) {
    lbm_destroy(c);
    return code;
  };
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) return bail(hip_fail(e, "hipGetDevice"));
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreate"));
    c->own_stream = true;
  }
  if ((e = hipEventCreate(&c->ev0)) != hipSuccess ||
      (e = hipEventCreate(&c->ev1)) != hipSuccess)
    return bail(hip_fail(e, "hipEventCreate"));
  const size_t fbytes = population_elems(c->q, g.n) * c->elem;
  for (int k = 0; k < 2; ++k) {
    if ((rc = dev_alloc(c, &c->f[k], fbytes, "hipMalloc(populations)")))
      return bail(rc);
    if ((e = hipMemsetAsync(c->f[k], 0, fbytes, c->stream)) != hipSuccess)
      return bail(hip_fail(e, "hipMemsetAsync(populations)"));
  }
  if ((rc = dev_alloc(c, (void**)&c->force, 3 * sizeof(float), "hipMalloc(force)")))
    return bail(rc);
  if ((e = hipMemsetAsync(c->force, 0, 3 * sizeof(float), c->stream)) != hipSuccess)
    return bail(hip_fail(e, "hipMemsetAsync(force)"));
  c->red_cap = (int)std::min<uint32_t>(1024u, (g.n + 255) / 256);
  if ((rc = dev_alloc(c, (void**)&c->red, sizeof(double) * 4 * c->red_cap, "hipMalloc(reduction)")))
    return bail(rc);
  if ((rc = dev_alloc(c, (void**)&c->red_out, sizeof(double) * 4, "hipMalloc(reduction)")))
    return bail(rc);
  if (cfg->mode == LBM_MODE_STANDARD) {
    if ((rc = dev_alloc(c, (void**)&c->mask, (has_mask16(*cfg) ? 6 : 4) * (size_t)g.n,
                        "hipMalloc(mask)")))
      return bail(rc);
    if ((rc = dev_alloc(c, (void**)&c->cls, sizeof(uint32_t) * ((g.n + 15) / 16), "hipMalloc(class map)")))
      return bail(rc);
    if ((rc = launch_build_mask(c->q, g, nullptr, cfg->periodic, cfg->lid_face,
                                c->mask, c->cls, c->stream)) ||
        (has_mask16(*cfg) && (rc = launch_mask16(g, c->mask, c->stream))))
      return bail(rc);
  } else {
    if ((rc = dev_alloc(c, (void**)&c->flags, g.n, "hipMalloc(flags)")))
      return bail(rc);
    if ((e = hipMemsetAsync(c->flags, 0, g.n, c->stream)) != hipSuccess)
      return bail(hip_fail(e, "hipMemsetAsync(flags)"));
    c->host_flags.assign(g.n, 0);
  }
  if ((rc = sync(c))) return bail(rc);
  *out = c;
  return LBM_OK;
}

int lbm_destroy(lbm_ctx* c) {
  if (!c) return LBM_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (void* p : {c->f[0], c->f[1], (void*)c->mask, (void*)c->cls, (void*)c->flags,
                  (void*)c->rec_node, (void*)c->rec_dirs, (void*)c->rec_w,
                  (void*)c->special, (void*)c->special_rec,
                  (void*)c->flow_mask, (void*)c->force,
                  (void*)c->force_partials, (void*)c->red, (void*)c->red_out,
                  (void*)c->ref_cls, (void*)c->ref_keep, (void*)c->ref_inmask, (void*)c->ref_rxbase,
                  (void*)c->ref_gsrc, (void*)c->ref_gbeg, (void*)c->ref_gid, (void*)c->ref_gsum,
                  (void*)c->ref_eidx, (void*)c->ref_ew, (void*)c->ref_es,
                  (void*)c->ref_partials, c->stage})
    if (p) (void)hipFree(p);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return LBM_OK;
}

int lbm_get_config(const lbm_ctx* c, lbm_config* out) {
  if (int rc = check_ctx(c)) return rc;
  if (!out) return fail(LBM_ERR_INVALID, "null output");
  *out = c->cfg;
  return LBM_OK;
}

int64_t lbm_node_count(const lbm_ctx* c) { return c ? (int64_t)c->grid.n : -1; }

int64_t lbm_device_bytes(const lbm_ctx* c) { return c ? c->device_bytes : -1; }

double lbm_cell_reynolds(const lbm_config* cfg) {
  if (int rc = validate_config(cfg)) return -(double)rc;
  if (cfg->mode != LBM_MODE_STANDARD)
    return -(double)fail(LBM_ERR_STATE, "lbm_cell_reynolds needs a standard-mode configuration");
  return cell_reynolds(*cfg);
}

int lbm_set_tau(lbm_ctx* c, float tau) {
  if (int rc = check_ctx(c)) return rc;
  if (!(tau > 0.0f) || (c->cfg.mode == LBM_MODE_STANDARD && !(tau > 0.5f)))
    return fail(LBM_ERR_INVALID, "tau out of range");
  c->cfg.tau = tau;
  ++c->params_version;
  return LBM_OK;
}

int lbm_set_body_force(lbm_ctx* c, const float force[3]) {
  if (int rc = check_ctx(c)) return rc;
  if (!force) return fail(LBM_ERR_INVALID, "null force");
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(force[a])) return fail(LBM_ERR_INVALID, "body force must be finite");
  if (c->cfg.lattice == 9 && force[2] != 0.0f)
    return fail(LBM_ERR_INVALID, "D2Q9 has no z axis: the body force's z component must be 0");
  for (int a = 0; a < 3; ++a) c->cfg.force[a] = force[a];
  ++c->params_version;
  return LBM_OK;
}

int lbm_synchronize(lbm_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  return sync(c);
}

int lbm_upload_points(lbm_ctx* c, const lbm_lattice_point* pts, int64_t count) {
  if (int rc = check_ctx(c)) return rc;
  if (!pts) return fail(LBM_ERR_INVALID, "null points");
  if (count != (int64_t)c->grid.n)
    return fail(LBM_ERR_INVALID, "point count does not match the lattice");
  DevBuf st;
  const size_t bytes = sizeof(lbm_lattice_point) * (size_t)count;
  int rc;
  if ((rc = st.alloc(bytes))) return rc;
  if ((rc = h2d(c, st.p, pts, bytes))) return rc;
  const bool ref = c->cfg.mode == LBM_MODE_REFERENCE;
  if ((rc = launch_points_in(c->q, device_storage(c->cfg.mode, c->cfg.storage), c->grid,
                             (const lbm_lattice_point*)st.p, c->f[c->cur],
                             ref ? c->flags : nullptr, c->stream)))
    return rc;
  if ((rc = mirror_buffers(c))) return rc;
  if (ref) {
    for (int64_t n = 0; n < count; ++n)
      c->host_flags[n] =
          (uint8_t)((pts[n].isInternal ? LBM_FLAG_INTERNAL : 0u) |
                    (pts[n].isReflected ? LBM_FLAG_REFLECTED : 0u));
    c->special_dirty = true;
  }
  return sync(c);
}

int lbm_download_points(lbm_ctx* c, lbm_lattice_point* pts, int64_t count) {
  if (int rc = check_ctx(c)) return rc;
  if (!pts) return fail(LBM_ERR_INVALID, "null points");
  if (count != (int64_t)c->grid.n)
    return fail(LBM_ERR_INVALID, "point count does not match the lattice");
  DevBuf st;
  const size_t bytes = sizeof(lbm_lattice_point) * (size_t)count;
  int rc;
  if ((rc = st.alloc(bytes))) return rc;
  if ((rc = launch_points_out(c->q, device_storage(c->cfg.mode, c->cfg.storage), c->grid, c->f[c->cur],
                              c->flags, (lbm_lattice_point*)st.p, c->stream)))
    return rc;
  return d2h(c, pts, st.p, bytes);
}

static int set_pops(lbm_ctx* c, const void* f, int is_f64) {
  if (int rc = check_ctx(c)) return rc;
  if (!f) return fail(LBM_ERR_INVALID, "null populations");
  const size_t bytes = (size_t)c->q * c->grid.n * (is_f64 ? 8 : 4);
  DevBuf st;
  int rc;
  if ((rc = st.alloc(bytes))) return rc;
  if ((rc = h2d(c, st.p, f, bytes))) return rc;
  if ((rc = launch_convert_in(c->q, device_storage(c->cfg.mode, c->cfg.storage), c->grid, st.p, is_f64,
                              c->f[c->cur], c->stream)))
    return rc;
  if ((rc = mirror_buffers(c))) return rc;
  return sync(c);
}

static int get_pops(lbm_ctx* c, void* f, int is_f64) {
  if (int rc = check_ctx(c)) return rc;
  if (!f) return fail(LBM_ERR_INVALID, "null populations");
  const size_t bytes = (size_t)c->q * c->grid.n * (is_f64 ? 8 : 4);
  DevBuf st;
  int rc;
  if ((rc = st.alloc(bytes))) return rc;
  if ((rc = launch_convert_out(c->q, device_storage(c->cfg.mode, c->cfg.storage), c->grid, c->f[c->cur],
                               st.p, is_f64, c->stream)))
    return rc;
  return d2h(c, f, st.p, bytes);
}

int lbm_set_populations_f32(lbm_ctx* c, const float* f) { return set_pops(c, f, 0); }
int lbm_get_populations_f32(lbm_ctx* c, float* f) { return get_pops(c, f, 0); }
int lbm_set_populations_f64(lbm_ctx* c, const double* f) { return set_pops(c, f, 1); }
int lbm_get_populations_f64(lbm_ctx* c, double* f) { return get_pops(c, f, 1); }

int lbm_init_equilibrium(lbm_ctx* c, const double* rho, const double* u) {
  if (int rc = check_ctx(c)) return rc;
  const size_t N = c->grid.n;
  DevBuf dr, du;
  int rc;
  if (rho) {
    if ((rc = dr.alloc(sizeof(double) * N))) return rc;
    if ((rc = h2d(c, dr.p, rho, sizeof(double) * N))) return rc;
  }
  if (u) {
    if ((rc = du.alloc(sizeof(double) * 3 * N))) return rc;
    if ((rc = h2d(c, du.p, u, sizeof(double) * 3 * N))) return rc;
  }
  if ((rc = launch_init_eq(c->q, device_storage(c->cfg.mode, c->cfg.storage), c->grid, (const double*)dr.p,
                           (const double*)du.p, c->f[c->cur], c->stream)))
    return rc;
  if ((rc = mirror_buffers(c))) return rc;
  return sync(c);
}

int lbm_upload_solid(lbm_ctx* c, const uint8_t* solid) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_STANDARD, "lbm_upload_solid")) return rc;
  DevBuf ds;
  int rc;
  bool any_solid = false;
  if (solid) {
    if ((rc = ds.alloc(c->grid.n))) return rc;
    if ((rc = h2d(c, ds.p, solid, c->grid.n))) return rc;
    for (uint32_t n = 0; n < c->grid.n && !any_solid; ++n) any_solid = solid[n] != 0;
  }
  c->grid.flags = (c->grid.flags & ~kReadMasks) | (any_solid ? kReadMasks : 0u);
  ++c->params_version;
  if ((rc = launch_build_mask(c->q, c->grid, (const uint8_t*)ds.p,
                              c->cfg.periodic, c->cfg.lid_face, c->mask,
                              c->cls, c->stream)) ||
      (has_mask16(c->cfg) && (rc = launch_mask16(c->grid, c->mask, c->stream))))
    return rc;
  return sync(c);
}

int lbm_get_mask(lbm_ctx* c, uint32_t* mask) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_STANDARD, "lbm_get_mask")) return rc;
  if (!mask) return fail(LBM_ERR_INVALID, "null mask");
  return d2h(c, mask, c->mask, sizeof(uint32_t) * c->grid.n);
}

int lbm_upload_flags(lbm_ctx* c, const uint8_t* flags) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_REFERENCE, "lbm_upload_flags")) return rc;
  if (!flags) return fail(LBM_ERR_INVALID, "null flags");
  std::memcpy(c->host_flags.data(), flags, c->grid.n);
  c->special_dirty = true;
  return h2d(c, c->flags, flags, c->grid.n);
}

int lbm_get_flags(lbm_ctx* c, uint8_t* flags) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_REFERENCE, "lbm_get_flags")) return rc;
  if (!flags) return fail(LBM_ERR_INVALID, "null flags");
  return d2h(c, flags, c->flags, c->grid.n);
}

int lbm_upload_reflections(lbm_ctx* c, int32_t count, const int32_t* node,
                           const uint32_t* dirs, const float* weights) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_REFERENCE, "lbm_upload_reflections"))
    return rc;
  if (count < 0 || (count > 0 && (!node || !dirs || !weights)))
    return fail(LBM_ERR_INVALID, "bad reflection records");
  for (int32_t r = 0; r < count; ++r)
    if (node[r] < 0 || (uint32_t)node[r] >= c->grid.n)
      return fail(LBM_ERR_INVALID, "reflection record node out of range");
  const size_t old = (size_t)std::max(c->nrec, 1);
  dev_free(c, (void**)&c->rec_node, sizeof(int32_t) * old);
  dev_free(c, (void**)&c->rec_dirs, sizeof(uint32_t) * 14 * old);
  dev_free(c, (void**)&c->rec_w, sizeof(float) * 81 * old);
  c->nrec = 0;
  c->host_rec_node.clear();
  c->host_rec_dirs.clear();
  c->host_rec_w.clear();
  c->special_dirty = true;
  const size_t m = (size_t)std::max(count, 1);
  int rc;
  if ((rc = dev_alloc(c, (void**)&c->rec_node, sizeof(int32_t) * m, "hipMalloc(records)")) ||
      (rc = dev_alloc(c, (void**)&c->rec_dirs, sizeof(uint32_t) * 14 * m, "hipMalloc(records)")) ||
      (rc = dev_alloc(c, (void**)&c->rec_w, sizeof(float) * 81 * m, "hipMalloc(records)")))
    return rc;
  if (count > 0) {
    if ((rc = h2d(c, c->rec_node, node, sizeof(int32_t) * count)) ||
        (rc = h2d(c, c->rec_dirs, dirs, sizeof(uint32_t) * 14 * count)) ||
        (rc = h2d(c, c->rec_w, weights, sizeof(float) * 81 * count)))
      return rc;
  }
  c->nrec = count;
  c->host_rec_node.assign(node, node + count);
  c->host_rec_dirs.assign(dirs, dirs + 14 * (size_t)count);
  c->host_rec_w.assign(weights, weights + 81 * (size_t)count);
  // the reflection records usually complete the geometry: build the fused
  // step's receiver tables now rather than in the first step (a later flag
  // change marks them for rebuilding again)
  return rebuild_special(c);
}

int lbm_set_flow(lbm_ctx* c, int kind, const int32_t* params,
                 const float* induced, const uint8_t* mask) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_REFERENCE, "lbm_set_flow")) return rc;
  if (kind < LBM_FLOW_NONE || kind > LBM_FLOW_MASK)
    return fail(LBM_ERR_INVALID, "unknown flow kind");
  if (kind != LBM_FLOW_NONE && !induced)
    return fail(LBM_ERR_INVALID, "null induced flow");
  if ((kind == LBM_FLOW_POINT || kind == LBM_FLOW_PLANE || kind == LBM_FLOW_BOX) && !params)
    return fail(LBM_ERR_INVALID, "null flow parameters");
  if (kind == LBM_FLOW_PLANE && (params[0] < 0 || params[0] > 2))
    return fail(LBM_ERR_INVALID, "plane axis must be 0..2");
  if (kind == LBM_FLOW_MASK && !mask)
    return fail(LBM_ERR_INVALID, "null flow mask");
  FlowSpec fl{};
  fl.kind = kind;
  const int np = kind == LBM_FLOW_POINT ? 3 : (kind == LBM_FLOW_PLANE ? 2 : (kind == LBM_FLOW_BOX ? 6 : 0));
  for (int k = 0; k < np; ++k) fl.params[k] = params[k];
  if (induced)
    for (int i = 0; i < 27; ++i) fl.induced[i] = induced[i];
  if (kind == LBM_FLOW_MASK) {
    if (!c->flow_mask) {
      int rc = dev_alloc(c, (void**)&c->flow_mask, c->grid.n, "hipMalloc(flow mask)");
      if (rc) return rc;
    }
    int rc = h2d(c, c->flow_mask, mask, c->grid.n);
    if (rc) return rc;
    fl.mask = c->flow_mask;
  }
  c->flow = fl;
  return LBM_OK;
}

int lbm_stream(lbm_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_REFERENCE, "lbm_stream")) return rc;
  int rc = launch_ref_stream(c->grid, (const float*)c->f[c->cur],
                             (float*)c->f[c->cur ^ 1], c->stream);
  if (rc) return rc;
  c->cur ^= 1;
  return LBM_OK;
}

int lbm_collide(lbm_ctx* c, float tau) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_REFERENCE, "lbm_collide")) return rc;
  return launch_ref_collide(c->grid, (float*)c->f[c->cur], tau, nullptr,
                            c->stream);
}

int lbm_reflect(lbm_ctx* c, int with_force) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_REFERENCE, "lbm_reflect")) return rc;
  return reflect(c, with_force);
}

int lbm_flow(lbm_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_REFERENCE, "lbm_flow")) return rc;
  return launch_ref_flow(c->grid, (float*)c->f[c->cur], c->flow, c->stream);
}

int lbm_prime_points(lbm_ctx* c) { return check_ctx(c); }

int lbm_set_force(lbm_ctx* c, const float in[3]) {
  if (int rc = check_ctx(c)) return rc;
  if (!in) return fail(LBM_ERR_INVALID, "null force");
  return h2d(c, c->force, in, 3 * sizeof(float));
}

int lbm_get_force(lbm_ctx* c, float out[3]) {
  if (int rc = check_ctx(c)) return rc;
  if (!out) return fail(LBM_ERR_INVALID, "null force");
  return d2h(c, out, c->force, 3 * sizeof(float));
}

int lbm_step(lbm_ctx* c, int n) {
  if (int rc = check_ctx(c)) return rc;
  if (n < 0) return fail(LBM_ERR_INVALID, "negative step count");
  TraceRange tr("lbm_step");
  return c->cfg.mode == LBM_MODE_STANDARD ? step_standard(c, n)
                                          : step_reference_fused(c, n);
}

int lbm_step_each(lbm_ctx* const* ctxs, int32_t count, int n) {
  if (!ctxs || count < 0 || n < 0) return fail(LBM_ERR_INVALID, "bad arguments");
  TraceRange tr("lbm_step_each");
  for (int32_t k = 0; k < count; ++k) {
    if (int rc = check_ctx(ctxs[k])) return rc;
    if (ctxs[k]->cfg.mode != LBM_MODE_STANDARD)
      return fail(LBM_ERR_STATE, "lbm_step_each needs standard-mode lattices");
  }
  for (int s = 0; s < n; ++s)
    for (int32_t k = 0; k < count; ++k)
      if (int rc = step_standard(ctxs[k], 1)) return rc;
  return LBM_OK;
}

int lbm_step_reference(lbm_ctx* c, int n, int with_force) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = need_mode(c, LBM_MODE_REFERENCE, "lbm_step_reference")) return rc;
  if (n < 0) return fail(LBM_ERR_INVALID, "negative step count");
  TraceRange tr("lbm_step_reference");
  return step_reference_full(c, n, with_force);
}

int lbm_time_steps(lbm_ctx* c, int n, float* ms) {
  if (int rc = check_ctx(c)) return rc;
  if (!ms) return fail(LBM_ERR_INVALID, "null output");
  LBM_HIP(hipEventRecord(c->ev0, c->stream));
  int rc = api::lbm_step(c, n);
  if (rc) return rc;
  LBM_HIP(hipEventRecord(c->ev1, c->stream));
  LBM_HIP(hipEventSynchronize(c->ev1));
  LBM_HIP(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return LBM_OK;
}

int lbm_time_steps_reference(lbm_ctx* c, int n, int with_force, float* ms) {
  if (int rc = check_ctx(c)) return rc;
  if (!ms) return fail(LBM_ERR_INVALID, "null output");
  int rc = rebuild_special(c);
  if (rc) return rc;
  LBM_HIP(hipEventRecord(c->ev0, c->stream));
  if ((rc = api::lbm_step_reference(c, n, with_force))) return rc;
  LBM_HIP(hipEventRecord(c->ev1, c->stream));
  LBM_HIP(hipEventSynchronize(c->ev1));
  LBM_HIP(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return LBM_OK;
}

int lbm_moments(lbm_ctx* c, float* rho, float* u) {
  if (int rc = check_ctx(c)) return rc;
  const size_t N = c->grid.n;
  int rc;
  // rho and u staged in the context's buffer (4 N + 12 N bytes), kept
  // from call to call
  if (c->stage_bytes < 16 * N) {
    dev_free(c, &c->stage, c->stage_bytes);
    c->stage_bytes = 0;
    if ((rc = dev_alloc(c, &c->stage, 16 * N, "hipMalloc(moments staging)"))) return rc;
    c->stage_bytes = 16 * N;
  }
  return read_moments(c, c->stage, rho, u);
}

int lbm_conservation(lbm_ctx* c, double out[4]) {
  if (int rc = check_ctx(c)) return rc;
  if (!out) return fail(LBM_ERR_INVALID, "null output");
  int rc = launch_conservation(c->q, device_storage(c->cfg.mode, c->cfg.storage), c->grid, c->f[c->cur],
                               c->mask, c->red, c->red_cap, c->red_out,
                               c->stream);
  if (rc) return rc;
  return d2h(c, out, c->red_out, 4 * sizeof(double));
}

int lbm_test_coordinates(lbm_ctx* c, int* ok) {
  if (int rc = check_ctx(c)) return rc;
  if (!ok) return fail(LBM_ERR_INVALID, "null output");
  DevBuf counts, dok;
  int rc;
  if ((rc = counts.alloc(sizeof(uint32_t) * c->grid.n)) || (rc = dok.alloc(sizeof(int))))
    return rc;
  const int one = 1;
  LBM_HIP(hipMemsetAsync(counts.p, 0, sizeof(uint32_t) * c->grid.n, c->stream));
  if ((rc = h2d(c, dok.p, &one, sizeof(int)))) return rc;
  if ((rc = launch_test_coordinates(c->grid, (uint32_t*)counts.p, (int*)dok.p,
                                    c->stream)))
    return rc;
  return d2h(c, ok, dok.p, sizeof(int));
}

int lbm_measure_copy_bandwidth(int64_t bytes, int iters, double* gbs) {
  if (bytes < 16 || iters < 1 || !gbs)
    return fail(LBM_ERR_INVALID, "bad copy-bandwidth arguments");
  const uint64_t nbytes = (uint64_t)bytes / 16 * 16;
  DevBuf a, b;
  int rc;
  if ((rc = a.alloc(nbytes)) || (rc = b.alloc(nbytes))) return rc;
  hipStream_t s;
  LBM_HIP(hipStreamCreate(&s));
  hipEvent_t e0, e1;
  hipError_t e;
  if ((e = hipEventCreate(&e0)) != hipSuccess) {
    (void)hipStreamDestroy(s);
    return hip_fail(e, "hipEventCreate");
  }
  if ((e = hipEventCreate(&e1)) != hipSuccess) {
    (void)hipEventDestroy(e0);
    (void)hipStreamDestroy(s);
    return hip_fail(e, "hipEventCreate");
  }
  float ms = 0.0f;
  rc = LBM_OK;
  if ((e = hipMemsetAsync(a.p, 0, nbytes, s)) != hipSuccess) rc = hip_fail(e, "hipMemsetAsync");
  for (int w = 0; w < 3 && rc == LBM_OK; ++w) rc = launch_copy_stream(a.p, b.p, nbytes, s);
  if (rc == LBM_OK && (e = hipEventRecord(e0, s)) != hipSuccess) rc = hip_fail(e, "hipEventRecord");
  for (int k = 0; k < iters && rc == LBM_OK; ++k) rc = launch_copy_stream(a.p, b.p, nbytes, s);
  if (rc == LBM_OK && (e = hipEventRecord(e1, s)) != hipSuccess) rc = hip_fail(e, "hipEventRecord");
  if (rc == LBM_OK && (e = hipEventSynchronize(e1)) != hipSuccess) rc = hip_fail(e, "hipEventSynchronize");
  if (rc == LBM_OK && (e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) rc = hip_fail(e, "hipEventElapsedTime");
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  if (rc == LBM_OK) *gbs = 2.0 * (double)nbytes * iters / (ms * 1e-3) / 1e9;
  return rc;
}

int64_t lbm_random_spheres(int32_t nx, int32_t ny, int32_t nz, uint64_t seed,
                           double frac, double rmin, double rmax,
                           int32_t periodic_x, uint8_t* solid,
                           int64_t max_spheres, double* spheres) {
#pragma clang fp contract(off)
  if (nx < 1 || ny < 1 || nz < 1 || !solid || !(frac >= 0.0 && frac <= 1.0) ||
      !(rmin > 0.0) || rmax < rmin) {
    set_error("bad random-sphere arguments");
    return -LBM_ERR_INVALID;
  }
  const int64_t N = (int64_t)nx * ny * nz;
  std::memset(solid, 0, (size_t)N);
  std::mt19937_64 gen(seed);
  auto u01 = [&gen]() { return (double)(gen() >> 11) * 0x1.0p-53; };
  int64_t count = 0, ns = 0;
  const double target = frac * (double)N;
  while ((double)count < target) {
    const double c[3] = {u01() * nx, u01() * ny, u01() * nz};
    const double r = rmin + (rmax - rmin) * u01();
    if (spheres && ns < max_spheres) {
      spheres[4 * ns + 0] = c[0];
      spheres[4 * ns + 1] = c[1];
      spheres[4 * ns + 2] = c[2];
      spheres[4 * ns + 3] = r;
    }
    ++ns;
    const int64_t lo[3] = {(int64_t)std::floor(c[0] - r), (int64_t)std::floor(c[1] - r),
                           (int64_t)std::floor(c[2] - r)};
    const int64_t hi[3] = {(int64_t)std::ceil(c[0] + r), (int64_t)std::ceil(c[1] + r),
                           (int64_t)std::ceil(c[2] + r)};
    const double r2 = r * r;
    for (int64_t x = lo[0]; x <= hi[0]; ++x) {
      int64_t xi = x;
      if (xi < 0 || xi >= nx) {
        if (!periodic_x) continue;
        xi = ((xi % nx) + nx) % nx;
      }
      const double dx = (double)x - c[0];
      const double dx2 = dx * dx;
      for (int64_t y = std::max<int64_t>(lo[1], 0); y <= hi[1] && y < ny; ++y) {
        const double dy = (double)y - c[1];
        const double dxy = dx2 + dy * dy;
        for (int64_t z = std::max<int64_t>(lo[2], 0); z <= hi[2] && z < nz; ++z) {
          const double dz = (double)z - c[2];
          if (dxy + dz * dz <= r2) {
            uint8_t& s = solid[z + nz * (y + ny * xi)];
            if (!s) {
              s = 1;
              ++count;
            }
          }
        }
      }
    }
  }
  return ns;
}

}  // namespace api
}  // namespace lbm
