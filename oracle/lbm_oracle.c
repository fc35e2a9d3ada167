/*
 * lbm_oracle.c -- CPU restatement of the reference lattice-Boltzmann path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product.  Function-by-function citations of the reference are in
 * lbm_oracle.h; the parity pins of each function are listed there too.
 *
 * Built with -ffp-contract=off so that every float operation rounds the way
 * it is written.
 */
#define _GNU_SOURCE
#include "lbm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================== */
/* tables: simulation/source/Constants.cpp:6-7, SimCalcFuncs.cu:12-14     */
/* ===================================================================== */
const int ORC_E27[27][3] = {
    {0, 0, 0},   {-1, 0, 0},  {0, -1, 0},   {0, 0, -1},  {0, 0, 1},
    {0, 1, 0},   {1, 0, 0},   {-1, -1, 0},  {-1, 0, -1}, {-1, 0, 1},
    {-1, 1, 0},  {0, -1, -1}, {0, -1, 1},   {0, 1, -1},  {0, 1, 1},
    {1, -1, 0},  {1, 0, -1},  {1, 0, 1},    {1, 1, 0},   {-1, -1, -1},
    {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},   {1, -1, -1}, {1, -1, 1},
    {1, 1, -1},  {1, 1, 1}};
const int ORC_INV27[27] = {0,  6,  5,  4,  3,  2,  1,  18, 17, 16, 15, 14, 13, 12,
                           11, 10, 9,  8,  7,  26, 25, 24, 23, 22, 21, 20, 19};
const float ORC_SPLIT27[27] = {
    0.0f,       1.0f,       1.0f,       1.0f,       1.0f,       1.0f,
    1.0f,       0.5f,       0.5f,       0.5f,       0.5f,       0.5f,
    0.5f,       0.5f,       0.5f,       0.5f,       0.5f,       0.5f,
    0.5f,       1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f,
    1.0f / 3.0f, 1.0f / 3.0f, 1.0f / 3.0f};

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static inline int64_t idx3(int x, int y, int z, int ny, int nz) {
  return (int64_t)z + (int64_t)nz * ((int64_t)y + (int64_t)ny * x);
}

static inline int in_box(int x, int y, int z, int nx, int ny, int nz) {
  return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz;
}

/* ===================================================================== */
/* reference mode                                                         */
/* ===================================================================== */

/* SimCalcFuncs.cu:79-96: for each i, j = INVERSE[i]:
 *   T[x].f[j] = L[x + e_i].f[j] if x + e_i is in the box, else 0. */
void orc_ref_stream(int nx, int ny, int nz, const float* L, float* T) {
  const int64_t N = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t n = idx3(x, y, z, ny, nz);
        for (int i = 0; i < 27; ++i) {
          const int j = ORC_INV27[i];
          const int sx = x + ORC_E27[i][0], sy = y + ORC_E27[i][1],
                    sz = z + ORC_E27[i][2];
          T[j * N + n] = in_box(sx, sy, sz, nx, ny, nz)
                             ? L[j * N + idx3(sx, sy, sz, ny, nz)]
                             : 0.0f;
        }
      }
}

/* SimCalcFuncs.cu:98-156.  rho and the SPLIT-weighted velocity in float;
 * the equilibrium expression mixes float with double literals, so it is
 * evaluated in double (only u.u is a float sum) and rounded to float; the
 * relaxation (eq - f)/tau is a float divide added to f. */
static inline void ref_collide_node(float* fn /* 27 values, stride s */,
                                    int64_t s, float tau) {
  float density = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
  for (int i = 0; i < 27; ++i) {
    const float fi = fn[i * s];
    density += fi;
    ux += fi * (float)ORC_E27[i][0] * ORC_SPLIT27[i];
    uy += fi * (float)ORC_E27[i][1] * ORC_SPLIT27[i];
    uz += fi * (float)ORC_E27[i][2] * ORC_SPLIT27[i];
  }
  if (density != 0.0f) {
    ux /= density;
    uy /= density;
    uz /= density;
  }
  const float usq = (ux * ux) + (uy * uy) + (uz * uz);
  for (int i = 0; i < 27; ++i) {
    const float ct = ((float)ORC_E27[i][0] * ux) + ((float)ORC_E27[i][1] * uy) +
                     ((float)ORC_E27[i][2] * uz);
    double wgt;
    if (i == 0)
      wgt = 8.0 / 27.0;
    else if (i < 7)
      wgt = 2.0 / 27.0;
    else if (i < 19)
      wgt = 1.0 / 54.0;
    else
      wgt = 1.0 / 216.0;
    const float eq = (float)(wgt * density *
                             (1 + (3.0 * ct) + (4.5 * ct * ct) - (1.5 * usq)));
    const float fi = fn[i * s];
    fn[i * s] = fi + (eq - fi) / tau;
  }
}

void orc_ref_collide(int nx, int ny, int nz, float* f, float tau) {
  const int64_t N = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) ref_collide_node(f + n, N, tau);
}

/* Lattice.cpp:450-477 / SimCalcFuncs.cu:18-30: slot s lives in word s/6,
 * bits 5*(s%6) .. 5*(s%6)+4. */
static inline int slot_dir(const uint32_t* words, int s) {
  return (int)((words[s / 6] >> (5 * (s % 6))) & 0x1Fu);
}

static inline void set_slot_dir(uint32_t* words, int s, unsigned v) {
  const unsigned sh = 5u * (unsigned)(s % 6);
  words[s / 6] = (words[s / 6] & ~(0x1Fu << sh)) | ((v & 0x1Fu) << sh);
}

/* SimCalcFuncs.cu:158-198 with Lattice.cpp:163-210 around it:
 *   T = copy(L)
 *   for every node with isReflected, slots s = 1..80 with direction d != 0
 *   and x + e_d inside the box:
 *       T[x+e_d].f[d] += L[x].f[s/3] * W[s]
 *       F += L[x].f[s/3] * W[s] * (e_d*SPLIT[d] - DIRECTIONS[s]*SPLIT[s])
 *   for every node with isInternal: T[x].f[*] = 0.
 * The reference indexes DIRECTIONS/SPLIT with the slot s (not s/3); for
 * s >= 27 that read is past the table and is defined here as 0. The
 * reference adds with atomics in an unspecified order; nodes are visited
 * here in index order and slots in ascending order.  force (nullable)
 * receives every term in float, one addition per term, as the reference's
 * atomics do; force64 (nullable) receives the exact (FP64) sum of this
 * pass. */
void orc_ref_reflect(int nx, int ny, int nz, const float* L, float* T,
                     const uint8_t* flags, int nrec, const int32_t* rec_node,
                     const uint32_t* rec_dirs, const float* rec_w,
                     float* force, double* force64) {
  double acc[3] = {0.0, 0.0, 0.0};
  const int64_t N = (int64_t)nx * ny * nz;
  memcpy(T, L, sizeof(float) * 27 * (size_t)N);
  int32_t* rec_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
  for (int64_t n = 0; n < N; ++n) rec_of[n] = -1;
  for (int r = 0; r < nrec; ++r)
    if (rec_node[r] >= 0 && rec_node[r] < N) rec_of[rec_node[r]] = r;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t n = idx3(x, y, z, ny, nz);
        if ((flags[n] & ORC_FLAG_REFLECTED) && rec_of[n] >= 0) {
          const int r = rec_of[n];
          const uint32_t* words = rec_dirs + 14 * (int64_t)r;
          const float* W = rec_w + 81 * (int64_t)r;
          for (int s = 1; s < 81; ++s) {
            const int d = slot_dir(words, s);
            if (d == 0 || d > 26) continue; /* 27..31 name no direction */
            const int tx = x + ORC_E27[d][0], ty = y + ORC_E27[d][1],
                      tz = z + ORC_E27[d][2];
            if (!in_box(tx, ty, tz, nx, ny, nz)) continue;
            const float val = L[(s / 3) * N + n] * W[s];
            T[d * N + idx3(tx, ty, tz, ny, nz)] += val;
            for (int a = 0; a < 3; ++a) {
              const float td = (float)ORC_E27[d][a] * ORC_SPLIT27[d];
              const float ts =
                  s < 27 ? (float)ORC_E27[s][a] * ORC_SPLIT27[s] : 0.0f;
              const float term = val * (td - ts);
              if (force) force[a] += term;
              acc[a] += (double)term;
            }
          }
        }
        if (flags[n] & ORC_FLAG_INTERNAL)
          for (int i = 0; i < 27; ++i) T[i * N + n] = 0.0f;
      }
  free(rec_of);
  if (force64)
    for (int a = 0; a < 3; ++a) force64[a] += acc[a];
}

static inline int flow_match(int kind, const int32_t* p, int x, int y, int z) {
  switch (kind) {
    case ORC_FLOW_POINT:
      return x == p[0] && y == p[1] && z == p[2];
    case ORC_FLOW_PLANE: {
      const int c = p[0] == 0 ? x : (p[0] == 1 ? y : z);
      return c == p[1];
    }
    case ORC_FLOW_BOX:
      return x >= p[0] && y >= p[1] && z >= p[2] && x <= p[3] && y <= p[4] &&
             z <= p[5];
    default:
      return 0;
  }
}

/* SimCalcFuncs.cu:200-215: where the criterion holds, f_i += induced_i.
 * The criteria of FlowCriterion.cu:5-16 are POINT(1,1,1) and PLANE(x=0). */
void orc_ref_flow(int nx, int ny, int nz, float* f, int kind,
                  const int32_t* params, const float* induced) {
  const int64_t N = (int64_t)nx * ny * nz;
  if (kind == ORC_FLOW_NONE) return;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        if (!flow_match(kind, params, x, y, z)) continue;
        const int64_t n = idx3(x, y, z, ny, nz);
        for (int i = 0; i < 27; ++i) f[i * N + n] += induced[i];
      }
}

void orc_ref_run_main(int nx, int ny, int nz, float* f, const uint8_t* flags,
                      int nrec, const int32_t* rec_node,
                      const uint32_t* rec_dirs, const float* rec_w, float tau,
                      int flow_kind, const int32_t* flow_params,
                      const float* induced, float* force, double* force64,
                      int steps) {
  const int64_t N = (int64_t)nx * ny * nz;
  float* T = (float*)malloc(sizeof(float) * 27 * (size_t)N);
  for (int s = 0; s < steps; ++s) {
    orc_ref_stream(nx, ny, nz, f, T);
    orc_ref_reflect(nx, ny, nz, T, f, flags, nrec, rec_node, rec_dirs, rec_w,
                    force, force64);
    orc_ref_collide(nx, ny, nz, f, tau);
    orc_ref_flow(nx, ny, nz, f, flow_kind, flow_params, induced);
  }
  free(T);
}

void orc_ref_run_lattice(int nx, int ny, int nz, float* f, float tau,
                         int steps) {
  const int64_t N = (int64_t)nx * ny * nz;
  float* T = (float*)malloc(sizeof(float) * 27 * (size_t)N);
  for (int s = 0; s < steps; ++s) {
    orc_ref_stream(nx, ny, nz, f, T);
    memcpy(f, T, sizeof(float) * 27 * (size_t)N);
    orc_ref_collide(nx, ny, nz, f, tau);
  }
  free(T);
}

void orc_ref_moments(int nx, int ny, int nz, const float* f, float* rho,
                     float* u) {
  const int64_t N = (int64_t)nx * ny * nz;
  for (int64_t n = 0; n < N; ++n) {
    float d = 0.0f, ux = 0.0f, uy = 0.0f, uz = 0.0f;
    for (int i = 0; i < 27; ++i) {
      const float fi = f[i * N + n];
      d += fi;
      ux += fi * (float)ORC_E27[i][0] * ORC_SPLIT27[i];
      uy += fi * (float)ORC_E27[i][1] * ORC_SPLIT27[i];
      uz += fi * (float)ORC_E27[i][2] * ORC_SPLIT27[i];
    }
    if (d != 0.0f) {
      ux /= d;
      uy /= d;
      uz /= d;
    }
    rho[n] = d;
    u[n] = ux;
    u[N + n] = uy;
    u[2 * N + n] = uz;
  }
}

/* ===================================================================== */
/* geometry                                                               */
/* ===================================================================== */
struct orc_mesh {
  int nv, nf;
  float* v;  /* nv*3 */
  int* f;    /* nf*3 (triangles only) */
  int triangles_only;
  int closed;
};

/* OBJ text: "v x y z" and "f a b c ..." (1-based, "a/b/c" forms accepted).
 * A face with other than 3 corners makes the mesh invalid, as CGAL's
 * is_triangle_mesh does in Model.cpp:31. */
orc_mesh* orc_mesh_load(const char* path) {
  FILE* fp = fopen(path, "r");
  if (!fp) return NULL;
  orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
  int capv = 64, capf = 64;
  m->v = (float*)malloc(sizeof(float) * 3 * capv);
  m->f = (int*)malloc(sizeof(int) * 3 * capf);
  m->triangles_only = 1;
  char line[1024];
  while (fgets(line, sizeof line, fp)) {
    char* p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (p[0] == 'v' && (p[1] == ' ' || p[1] == '\t')) {
      float x, y, z;
      if (sscanf(p + 1, "%f %f %f", &x, &y, &z) == 3) {
        if (m->nv == capv) {
          capv *= 2;
          m->v = (float*)realloc(m->v, sizeof(float) * 3 * capv);
        }
        m->v[3 * m->nv] = x;
        m->v[3 * m->nv + 1] = y;
        m->v[3 * m->nv + 2] = z;
        m->nv++;
      }
    } else if (p[0] == 'f' && (p[1] == ' ' || p[1] == '\t')) {
      int ids[64], k = 0;
      char* q = p + 1;
      while (*q && *q != '#' && *q != '\n' && k < 64) {
        while (*q == ' ' || *q == '\t') ++q;
        if (!*q || *q == '#' || *q == '\n' || *q == '\r') break;
        char* end;
        long id = strtol(q, &end, 10);
        if (end == q) break;
        ids[k++] = (int)(id > 0 ? id - 1 : m->nv + id);
        q = end;
        while (*q && *q != ' ' && *q != '\t' && *q != '\n') ++q;
      }
      if (k != 3) {
        m->triangles_only = 0;
        continue;
      }
      if (m->nf == capf) {
        capf *= 2;
        m->f = (int*)realloc(m->f, sizeof(int) * 3 * capf);
      }
      for (int c = 0; c < 3; ++c) m->f[3 * m->nf + c] = ids[c];
      m->nf++;
    }
  }
  fclose(fp);
  /* closed: every directed edge a->b has its opposite b->a
   * (CGAL::is_closed == no border halfedge, Model.cpp:46-49). */
  m->closed = m->triangles_only && m->nf > 0;
  for (int t = 0; t < m->nf && m->closed; ++t)
    for (int c = 0; c < 3 && m->closed; ++c) {
      const int a = m->f[3 * t + c], b = m->f[3 * t + (c + 1) % 3];
      int found = 0;
      for (int u = 0; u < m->nf && !found; ++u)
        for (int e = 0; e < 3; ++e)
          if (m->f[3 * u + e] == b && m->f[3 * u + (e + 1) % 3] == a) {
            found = 1;
            break;
          }
      if (!found) m->closed = 0;
    }
  return m;
}

void orc_mesh_free(orc_mesh* m) {
  if (!m) return;
  free(m->v);
  free(m->f);
  free(m);
}

int orc_mesh_nverts(const orc_mesh* m) { return m ? m->nv : 0; }
int orc_mesh_nfaces(const orc_mesh* m) { return m ? m->nf : 0; }
int orc_mesh_closed(const orc_mesh* m) { return m ? m->closed : 0; }
int orc_mesh_valid(const orc_mesh* m) {
  return m ? (m->triangles_only && m->closed) : 0;
}

void orc_mesh_bbox(const orc_mesh* m, float out[6]) {
  for (int a = 0; a < 3; ++a) {
    out[a] = m->nv ? m->v[a] : 0.0f;
    out[3 + a] = m->nv ? m->v[a] : 0.0f;
  }
  for (int i = 1; i < m->nv; ++i)
    for (int a = 0; a < 3; ++a) {
      const float c = m->v[3 * i + a];
      if (c < out[a]) out[a] = c;
      if (c > out[3 + a]) out[3 + a] = c;
    }
}

static void tri_d(const orc_mesh* m, int t, double a[3], double b[3],
                  double c[3]) {
  for (int k = 0; k < 3; ++k) {
    a[k] = m->v[3 * m->f[3 * t] + k];
    b[k] = m->v[3 * m->f[3 * t + 1] + k];
    c[k] = m->v[3 * m->f[3 * t + 2] + k];
  }
}

static void sub3(const double* a, const double* b, double* o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
static void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static double norm3(const double* a) { return sqrt(dot3(a, a)); }

/* The point lies on the (closed) triangle, up to double rounding. */
static int on_triangle(const double* p, const double* a, const double* b,
                       const double* c) {
  double ab[3], ac[3], ap[3], n[3];
  sub3(b, a, ab);
  sub3(c, a, ac);
  sub3(p, a, ap);
  cross3(ab, ac, n);
  const double nn = norm3(n);
  if (nn == 0.0) return 0;
  const double dist = dot3(n, ap);
  if (fabs(dist) > 1e-12 * nn * (norm3(ap) + 1e-30)) return 0;
  /* barycentric coordinates by sub-areas */
  double bp[3], cp[3], t1[3], t2[3], t3[3], bc[3];
  sub3(p, b, bp);
  sub3(p, c, cp);
  sub3(c, b, bc);
  cross3(ab, ap, t1);             /* ~ w_c */
  cross3(ap, ac, t2);             /* ~ w_b */
  cross3(bc, bp, t3);             /* ~ w_a */
  const double eps = -1e-12 * nn;
  return dot3(t1, n) >= eps && dot3(t2, n) >= eps && dot3(t3, n) >= eps;
}

/* Moeller-Trumbore in double.  Returns 1 on a hit with t >= tmin; *edge is
 * set when the hit is within rounding of an edge or vertex. */
static int ray_tri(const double* o, const double* d, const double* a,
                   const double* b, const double* c, double tmin, double* t,
                   int* edge) {
  double e1[3], e2[3], pv[3], tv[3], qv[3];
  sub3(b, a, e1);
  sub3(c, a, e2);
  cross3(d, e2, pv);
  const double det = dot3(e1, pv);
  const double scale = norm3(e1) * norm3(e2) * norm3(d);
  if (fabs(det) <= 1e-14 * scale) return 0; /* parallel */
  const double inv = 1.0 / det;
  sub3(o, a, tv);
  const double u = dot3(tv, pv) * inv;
  cross3(tv, e1, qv);
  const double v = dot3(d, qv) * inv;
  const double tol = 1e-12;
  if (u < -tol || v < -tol || u + v > 1.0 + tol) return 0;
  const double tt = dot3(e2, qv) * inv;
  if (tt < tmin) return 0;
  *t = tt;
  *edge = (u < tol || v < tol || u + v > 1.0 - tol);
  return 1;
}

/* Model.cpp:51-63: CGAL::Side_of_triangle_mesh; ON_BOUNDARY counts as
 * inside.  Restated as: on any face -> inside; else parity of crossings of
 * a ray in a generic direction (re-cast when a crossing hits an edge). */
int orc_mesh_inside(const orc_mesh* m, float x, float y, float z) {
  if (!orc_mesh_valid(m)) return 0;
  const double p[3] = {x, y, z};
  double a[3], b[3], c[3];
  for (int t = 0; t < m->nf; ++t) {
    tri_d(m, t, a, b, c);
    if (on_triangle(p, a, b, c)) return 1;
  }
  static const double dirs[6][3] = {{0.8212, 0.4361, 0.3679},
                                    {-0.3119, 0.9013, 0.3009},
                                    {0.2411, -0.3307, 0.9124},
                                    {-0.7071, -0.5019, 0.4981},
                                    {0.5513, 0.5711, -0.6083},
                                    {-0.1237, -0.8871, -0.4447}};
  for (int r = 0; r < 6; ++r) {
    int count = 0, degenerate = 0;
    for (int t = 0; t < m->nf && !degenerate; ++t) {
      tri_d(m, t, a, b, c);
      double tt;
      int edge;
      if (ray_tri(p, dirs[r], a, b, c, 1e-300, &tt, &edge)) {
        if (edge) degenerate = 1;
        ++count;
      }
    }
    if (!degenerate) return count & 1;
  }
  return 0;
}

static float fnorm3(float x, float y, float z) {
  return sqrtf((x * x) + (y * y) + (z * z));
}

static void quat_mult(const float a[4], const float b[4], float r[4]) {
  r[0] = (a[0] * b[0]) - (a[1] * b[1]) - (a[2] * b[2]) - (a[3] * b[3]);
  r[1] = (a[0] * b[1]) + (a[1] * b[0]) + (a[2] * b[3]) - (a[3] * b[2]);
  r[2] = (a[0] * b[2]) - (a[1] * b[3]) + (a[2] * b[0]) + (a[3] * b[1]);
  r[3] = (a[0] * b[3]) + (a[1] * b[2]) - (a[2] * b[1]) + (a[3] * b[0]);
}

/* Geometry.cpp:22-42: q v q^-1 with q = (cos(a/2), sin(a/2) axis), in float
 * (the reference calls the float overloads of sqrt/cos/sin). */
void orc_rotate_vector(const float axis_in[3], const float v[3], float angle,
                       float out[3]) {
  float axis[3] = {axis_in[0], axis_in[1], axis_in[2]};
  const float mag = fnorm3(axis[0], axis[1], axis[2]);
  axis[0] /= mag;
  axis[1] /= mag;
  axis[2] /= mag;
  angle = (float)((double)angle / 2.0);
  const float c = cosf(angle), s = sinf(angle), cn = cosf(-angle),
              sn = sinf(-angle);
  const float q[4] = {c, s * axis[0], s * axis[1], s * axis[2]};
  const float qi[4] = {cn, sn * axis[0], sn * axis[1], sn * axis[2]};
  const float pt[4] = {0.0f, v[0], v[1], v[2]};
  float t[4], r[4];
  quat_mult(q, pt, t);
  quat_mult(t, qi, r);
  out[0] = r[1];
  out[1] = r[2];
  out[2] = r[3];
}

/* Model.cpp:75-91: first hit of the ray from p1 through p2; the unit vector
 * p1 - p2 rotated by pi about the hit face's normal; (0,0,0) when nothing
 * is hit.  The face normal is (v1-v0) x (v2-v0) in float. */
void orc_mesh_reflection_vector(const orc_mesh* m, const float p1[3],
                                const float p2[3], float out[3]) {
  float v[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  const float nv = fnorm3(v[0], v[1], v[2]);
  v[0] /= nv;
  v[1] /= nv;
  v[2] /= nv;
  const double o[3] = {p1[0], p1[1], p1[2]};
  const double d[3] = {(double)p2[0] - p1[0], (double)p2[1] - p1[1],
                       (double)p2[2] - p1[2]};
  /* nearest hit; hits within 1e-9 of the nearest (a ray through an edge
   * or vertex shared by several faces) go to the lowest face index */
  int best = -1;
  double bt = 0.0;
  double a[3], b[3], c[3];
  for (int t = 0; t < m->nf; ++t) {
    tri_d(m, t, a, b, c);
    double tt;
    int edge;
    if (ray_tri(o, d, a, b, c, 0.0, &tt, &edge) && (best < 0 || tt < bt)) bt = tt, best = t;
  }
  if (best >= 0)
    for (int t = 0; t < best; ++t) {
      tri_d(m, t, a, b, c);
      double tt;
      int edge;
      if (ray_tri(o, d, a, b, c, 0.0, &tt, &edge) && tt <= bt + 1e-9) {
        best = t;
        break;
      }
    }
  if (best < 0) {
    out[0] = out[1] = out[2] = 0.0f;
    return;
  }
  const float* A = m->v + 3 * m->f[3 * best];
  const float* B = m->v + 3 * m->f[3 * best + 1];
  const float* C = m->v + 3 * m->f[3 * best + 2];
  const float e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  const float e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
  const float n[3] = {(e1[1] * e2[2]) - (e1[2] * e2[1]),
                      (e1[2] * e2[0]) - (e1[0] * e2[2]),
                      (e1[0] * e2[1]) - (e1[1] * e2[0])};
  orc_rotate_vector(n, v, (float)M_PI, out);
}

/* Lattice.cpp:262-321: score_j = 2 - |v - e_j/|e_j|| for all 27 directions
 * (rest included), ordered by descending score, ties to the higher index
 * (std::priority_queue over pair<float,int>). */
void orc_distribute_vector(const float v[3], float score[27], int32_t idx[27]) {
  float unit[27][3];
  for (int j = 0; j < 27; ++j) {
    const int k = abs(ORC_E27[j][0]) + abs(ORC_E27[j][1]) + abs(ORC_E27[j][2]);
    const double inv = k == 0 ? 0.0 : (k == 1 ? 1.0 : (k == 2 ? 1.0 / 1.4142135623730951 : 1.0 / 1.7320508075688772));
    for (int a = 0; a < 3; ++a)
      unit[j][a] = (float)(ORC_E27[j][a] == 0 ? 0.0 : ORC_E27[j][a] * inv);
  }
  float sc[27];
  for (int j = 0; j < 27; ++j) {
    const float dx = v[0] - unit[j][0], dy = v[1] - unit[j][1],
                dz = v[2] - unit[j][2];
    sc[j] = (float)(2.0 - (double)sqrtf((dx * dx) + (dy * dy) + (dz * dz)));
  }
  int used[27] = {0};
  for (int r = 0; r < 27; ++r) {
    int best = -1;
    for (int j = 26; j >= 0; --j) {
      if (used[j]) continue;
      if (best < 0 || sc[j] > sc[best]) best = j;
    }
    used[best] = 1;
    score[r] = sc[best];
    idx[r] = best;
  }
}

/* Lattice.cpp:228-260: isInternal(x) = mesh contains (x*h, y*h, z*h). */
int orc_insert_model(const orc_mesh* m, int nx, int ny, int nz, float spacing,
                     uint8_t* internal) {
  float bb[6];
  orc_mesh_bbox(m, bb);
  if (bb[0] < 0.0f || bb[1] < 0.0f || bb[2] < 0.0f ||
      bb[3] > nx * spacing || bb[4] > ny * spacing || bb[5] > nz * spacing)
    return 0;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z)
        internal[idx3(x, y, z, ny, nz)] = (uint8_t)orc_mesh_inside(
            m, (float)x * spacing, (float)y * spacing, (float)z * spacing);
  return 1;
}

/* Lattice.cpp:323-426.  For each internal node x and each direction k with
 * x + e_k inside the box and not internal: reflect along the ray from
 * (x+e_k)h to xh, rank the 27 directions, keep the first 3 whose target
 * x + e_d is inside the box and not internal (the reference does not check
 * the box; an outside target is defined here as not eligible), write them
 * to slots INVERSE[k]*3 + rank and normalise their weights. */
int orc_preprocess_model(const orc_mesh* m, int nx, int ny, int nz,
                         float spacing, uint8_t* flags, int max_rec,
                         int32_t* rec_node, uint32_t* rec_dirs, float* rec_w) {
  int nrec = 0;
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t n = idx3(x, y, z, ny, nz);
        if (!(flags[n] & ORC_FLAG_INTERNAL)) continue;
        int r = -1;
        for (int k = 0; k < 27; ++k) {
          const int sx = x + ORC_E27[k][0], sy = y + ORC_E27[k][1],
                    sz = z + ORC_E27[k][2];
          if (!in_box(sx, sy, sz, nx, ny, nz)) continue;
          if (flags[idx3(sx, sy, sz, ny, nz)] & ORC_FLAG_INTERNAL) continue;
          const float p1[3] = {spacing * (float)sx, spacing * (float)sy,
                               spacing * (float)sz};
          const float p2[3] = {spacing * (float)x, spacing * (float)y,
                               spacing * (float)z};
          float rv[3], sc[27];
          int32_t id[27];
          orc_mesh_reflection_vector(m, p1, p2, rv);
          orc_distribute_vector(rv, sc, id);
          int counter = 0;
          float norm = 0.0f;
          const int base = ORC_INV27[k] * 3;
          for (int i = 0; i < 27; ++i) {
            const int d = id[i];
            const int tx = x + ORC_E27[d][0], ty = y + ORC_E27[d][1],
                      tz = z + ORC_E27[d][2];
            if (in_box(tx, ty, tz, nx, ny, nz) &&
                !(flags[idx3(tx, ty, tz, ny, nz)] & ORC_FLAG_INTERNAL)) {
              if (r < 0) {
                if (nrec >= max_rec) return -1;
                r = nrec++;
                rec_node[r] = (int32_t)n;
                memset(rec_dirs + 14 * (int64_t)r, 0, 14 * sizeof(uint32_t));
                memset(rec_w + 81 * (int64_t)r, 0, 81 * sizeof(float));
              }
              set_slot_dir(rec_dirs + 14 * (int64_t)r, base + counter,
                           (unsigned)d);
              rec_w[81 * (int64_t)r + base + counter] = sc[i];
              norm += sc[i];
              ++counter;
            }
            if (counter == 3) break;
          }
          if (norm != 0.0f && r >= 0)
            for (int i = 0; i < 3; ++i)
              if (slot_dir(rec_dirs + 14 * (int64_t)r, base + i) != 0)
                rec_w[81 * (int64_t)r + base + i] /= norm;
        }
        if (r >= 0) flags[n] |= ORC_FLAG_REFLECTED;
      }
  return nrec;
}

/* Utils.cpp:3-28 */
unsigned orc_compress_ints(const int* a, int num, int bits) {
  unsigned out = 0, max = 1;
  for (int b = 0; b < bits; ++b) max *= 2;
  for (int i = 0; i < num - 1; ++i)
    if ((unsigned)a[i] < max) {
      out |= (unsigned)a[i];
      out <<= 5;
    }
  if ((unsigned)a[num - 1] < max) out |= (unsigned)a[num - 1];
  return out;
}

/* ===================================================================== */
/* standard mode                                                          */
/* ===================================================================== */
static const int E9[9][3] = {{0, 0, 0},  {-1, 0, 0}, {0, -1, 0},
                             {0, 1, 0},  {1, 0, 0},  {-1, -1, 0},
                             {-1, 1, 0}, {1, -1, 0}, {1, 1, 0}};
static const int INV9[9] = {0, 4, 3, 2, 1, 8, 7, 6, 5};

int orc_std_q_valid(int q) { return q == 9 || q == 19 || q == 27; }

static void std_table(int q, int e[27][3], int inv[27], double w[27]) {
  for (int i = 0; i < q; ++i) {
    if (q == 9) {
      for (int a = 0; a < 3; ++a) e[i][a] = E9[i][a];
      inv[i] = INV9[i];
      w[i] = i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
    } else {
      for (int a = 0; a < 3; ++a) e[i][a] = ORC_E27[i][a];
      inv[i] = ORC_INV27[i];
      if (q == 19)
        w[i] = i == 0 ? 1.0 / 3.0 : (i < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
      else
        w[i] = i == 0 ? 8.0 / 27.0
                      : (i < 7 ? 2.0 / 27.0 : (i < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
    }
  }
}

void orc_std_tables(int q, int32_t* e, int32_t* inv, double* w) {
  int E[27][3], I[27];
  double W[27];
  std_table(q, E, I, W);
  for (int i = 0; i < q; ++i) {
    for (int a = 0; a < 3; ++a) e[3 * i + a] = E[i][a];
    inv[i] = I[i];
    w[i] = W[i];
  }
}

void orc_std_build_mask(int q, int nx, int ny, int nz, const uint8_t* solid,
                        const int32_t periodic[3], int lid_face,
                        uint32_t* mask) {
  int E[27][3], I[27];
  double W[27];
  std_table(q, E, I, W);
  const int dims[3] = {nx, ny, nz};
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t n = idx3(x, y, z, ny, nz);
        if (solid && solid[n]) {
          mask[n] = ORC_MASK_SOLID;
          continue;
        }
        uint32_t m = 0;
        const int c[3] = {x, y, z};
        for (int i = 1; i < q; ++i) {
          int s[3], out = 0;
          for (int a = 0; a < 3; ++a) {
            s[a] = c[a] - E[i][a];
            if (s[a] < 0 || s[a] >= dims[a]) {
              if (periodic[a])
                s[a] = (s[a] + dims[a]) % dims[a];
              else
                out = 1;
            }
          }
          if (out || (solid && solid[idx3(s[0], s[1], s[2], ny, nz)]))
            m |= 1u << i;
        }
        if (lid_face >= 0) {
          const int ax = lid_face / 2, side = lid_face % 2;
          if ((side == 1 && c[ax] == dims[ax] - 1) || (side == 0 && c[ax] == 0))
            m |= ORC_MASK_LID;
        }
        mask[n] = m;
      }
}

void orc_std_init_eq_f64(int q, int nx, int ny, int nz, const double* rho,
                         const double* u, double* f) {
  int E[27][3], I[27];
  double W[27];
  std_table(q, E, I, W);
  const int64_t N = (int64_t)nx * ny * nz;
  for (int64_t n = 0; n < N; ++n) {
    const double r = rho ? rho[n] : 1.0;
    const double ux = u ? u[n] : 0.0, uy = u ? u[N + n] : 0.0,
                 uz = u ? u[2 * N + n] : 0.0;
    const double usq = ux * ux + uy * uy + uz * uz;
    for (int i = 0; i < q; ++i) {
      const double eu = E[i][0] * ux + E[i][1] * uy + E[i][2] * uz;
      f[i * N + n] = W[i] * r * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq);
    }
  }
}

float orc_round_fp16(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  const uint32_t sign = b & 0x80000000u;
  uint32_t a = b & 0x7fffffffu;
  if (a >= 0x7f800000u) return x;
  float r;
  if (a >= 0x477ff000u) { /* rounds to infinity in binary16 */
    a = 0x7f800000u;
    memcpy(&r, &a, 4);
  } else if (a < 0x38800000u) { /* below 2^-14: binary16 subnormal grid 2^-24 */
    float ax;
    memcpy(&ax, &a, 4);
    const float qs = 16777216.0f; /* 2^24 */
    const float k = nearbyintf(ax * qs); /* exact scaling, round-to-even */
    r = k / qs;
  } else {
    a = a + 0x0fffu + ((a >> 13) & 1u);
    a &= ~0x1fffu;
    memcpy(&r, &a, 4);
  }
  uint32_t rb;
  memcpy(&rb, &r, 4);
  rb |= sign;
  memcpy(&r, &rb, 4);
  return r;
}

float orc_round_bf16(float x) {
  uint32_t b;
  memcpy(&b, &x, 4);
  if ((b & 0x7fffffffu) >= 0x7f800000u) return x;
  b = b + 0x7fffu + ((b >> 16) & 1u);
  b &= 0xffff0000u;
  float r;
  memcpy(&r, &b, 4);
  return r;
}

/* 16-bit storage of the shifted deviation v = f_i - w_i, as the GPU library
 * stores it: FP16 holds v * 2^12 (so deviations down to 1.5e-8 are normal
 * binary16 numbers), BF16 holds v.  Lattices with a body force round
 * stochastically: the bits the conversion drops get a random addend before
 * they are cut -- the GPU's rule, so the emulation rounds with the same
 * distribution (its floats can differ from the GPU's by the order of
 * operations, so individual roundings can differ).  The addend is the top
 * bits of (low 24 bits of the value's float pattern) x (an odd 24-bit
 * multiplier of the direction) + (a node hash: the low 24 bits of the
 * node's element index in the GPU's blocked layout, (n / 64) * Q * 64 +
 * n % 64, times an odd constant).  Returns the stored value, unscaled. */
uint32_t orc_sr_mul(int i) { return ((0x9E3779B9u * (uint32_t)(i + 1)) >> 8) | 1u; }

static uint32_t orc_sr_hash(int64_t n, int q) {
  const uint32_t own = (uint32_t)((n >> 6) * (int64_t)q * 64 + (n & 63));
  return (own & 0xffffffu) * 0x5BD1E9u;
}

static uint32_t orc_sat_add(uint32_t a, uint32_t b) {
  const uint32_t s = a + b;
  return s < a ? 0xffffffffu : s;
}

static float orc_store16(int storage, float v, int sr, uint32_t mul, uint32_t hash) {
  if (storage == 2) {
    if (sr) {
      /* gfx950's stochastically rounding FP32 -> FP16 conversion, whose
       * random word is the low 32 bits of (24-bit value bits) x multiplier
       * + hash (scaling by 2^12 keeps the mantissa and the exponent's
       * parity, so the value bits are those of the scaled value) and r its
       * top 13 bits: a scaled value in FP16's normal range (or not finite)
       * gets r added to the 13 bits the conversion drops, which are then
       * cut; below 2^-14 it becomes floor(|x| 2^24 + r / 2^13) units of
       * 2^-24 with its sign */
      uint32_t b;
      memcpy(&b, &v, 4);
      const uint32_t r = ((b & 0xffffffu) * mul + hash) >> 19;
      const float x = v * 4096.0f;
      if (fabsf(x) < 6.103515625e-05f) {
        const double units = floor((double)fabsf(x) * 16777216.0 + (double)r / 8192.0);
        const float y = (float)(units / 16777216.0);
        return (x < 0.0f || (x == 0.0f && signbit(x)) ? -y : y) / 4096.0f;
      }
      b = orc_sat_add(b, r) & 0xffffe000u;
      memcpy(&v, &b, 4);
    }
    return orc_round_fp16(v * 4096.0f) / 4096.0f;
  }
  if (storage == 3) {
    if (!sr) return orc_round_bf16(v);
    uint32_t b;
    memcpy(&b, &v, 4);
    const uint32_t r = ((b & 0xffffffu) * mul + hash) >> 16;
    b = orc_sat_add(b, r) & 0xffff0000u;
    float o;
    memcpy(&o, &b, 4);
    return o;
  }
  return v;
}

float orc_store16_sr(int storage, float v, int direction, uint32_t hash) {
  return orc_store16(storage, v, 1, orc_sr_mul(direction), hash);
}

/* One pull step of the standard model for one node, float compute.
 *   f_i = A_i(x - e_i)                  open link (periodic wrap)
 *   f_i = A_inv(i)(x) + 6 w_i rho_w e_i.u_w   bounced link (wall term only
 *                                          when the source lies beyond the
 *                                          moving face)
 *   rho = sum f, u = (sum f e + F/2)/rho
 *   f*_i = f_i + (feq_i - f_i)/tau + (1 - 1/(2 tau)) w_i [3(e_i-u) + 9(e_i.u)e_i].F
 */
#define STD_STEP_BODY(REAL)                                                    \
  {                                                                            \
    const int dims[3] = {nx, ny, nz};                                          \
    const REAL omega = (REAL)1 / (REAL)p->tau;                                 \
    const REAL guo = (REAL)1 - (REAL)0.5 * omega;                              \
    const REAL Fx = (REAL)p->force[0], Fy = (REAL)p->force[1],                 \
               Fz = (REAL)p->force[2];                                         \
    const int lax = p->lid_face >= 0 ? p->lid_face / 2 : 0;                    \
    const int lsg = p->lid_face >= 0 ? (p->lid_face % 2 ? 1 : -1) : 0;         \
    REAL wt[27], lidt[27];                                                     \
    for (int i = 0; i < q; ++i) {                                              \
      wt[i] = (REAL)W[i];                                                      \
      lidt[i] = (REAL)6 * wt[i] * (REAL)p->wall_rho *                          \
                ((REAL)E[i][0] * (REAL)p->lid_u[0] +                           \
                 (REAL)E[i][1] * (REAL)p->lid_u[1] +                           \
                 (REAL)E[i][2] * (REAL)p->lid_u[2]);                           \
    }                                                                          \
    _Pragma("omp parallel for schedule(static)") for (int x = 0; x < nx; ++x)  \
      for (int y = 0; y < ny; ++y)                                             \
        for (int z = 0; z < nz; ++z) {                                         \
          const int64_t n = idx3(x, y, z, ny, nz);                             \
          const uint32_t m = mask[n];                                          \
          if (m & ORC_MASK_SOLID) {                                            \
            for (int i = 0; i < q; ++i) STORE(B, i * N + n, i, wt[i]);         \
            continue;                                                          \
          }                                                                    \
          const int c[3] = {x, y, z};                                          \
          REAL f[27];                                                          \
          for (int i = 0; i < q; ++i) {                                        \
            if ((m >> i) & 1u) {                                               \
              f[i] = LOAD(A, I[i] * N + n, I[i]);                              \
              if ((m & ORC_MASK_LID) && E[i][lax] == -lsg) f[i] += lidt[i];    \
            } else {                                                           \
              int s[3];                                                        \
              for (int a = 0; a < 3; ++a) {                                    \
                s[a] = c[a] - E[i][a];                                         \
                if (s[a] < 0) s[a] += dims[a];                                 \
                if (s[a] >= dims[a]) s[a] -= dims[a];                          \
              }                                                                \
              f[i] = LOAD(A, i * N + idx3(s[0], s[1], s[2], ny, nz), i);       \
            }                                                                  \
          }                                                                    \
          REAL rho = 0, jx = 0, jy = 0, jz = 0;                                \
          for (int i = 0; i < q; ++i) {                                        \
            rho += f[i];                                                       \
            jx += f[i] * (REAL)E[i][0];                                        \
            jy += f[i] * (REAL)E[i][1];                                        \
            jz += f[i] * (REAL)E[i][2];                                        \
          }                                                                    \
          const REAL ir = (REAL)1 / rho;                                       \
          const REAL ux = (jx + (REAL)0.5 * Fx) * ir,                          \
                     uy = (jy + (REAL)0.5 * Fy) * ir,                          \
                     uz = (jz + (REAL)0.5 * Fz) * ir;                          \
          const REAL usq = ux * ux + uy * uy + uz * uz;                        \
          const REAL uF = ux * Fx + uy * Fy + uz * Fz;                         \
          for (int i = 0; i < q; ++i) {                                        \
            const REAL eu = (REAL)E[i][0] * ux + (REAL)E[i][1] * uy +          \
                            (REAL)E[i][2] * uz;                                \
            const REAL eF = (REAL)E[i][0] * Fx + (REAL)E[i][1] * Fy +          \
                            (REAL)E[i][2] * Fz;                                \
            const REAL feq = wt[i] * rho *                                     \
                             ((REAL)1 + (REAL)3 * eu + (REAL)4.5 * eu * eu -   \
                              (REAL)1.5 * usq);                                \
            const REAL src = guo * wt[i] *                                     \
                             ((REAL)3 * (eF - uF) + (REAL)9 * eu * eF);        \
            STORE(B, i * N + n, i, f[i] + omega * (feq - f[i]) + src);         \
          }                                                                    \
        }                                                                      \
  }

static void std_step_f32(int q, int nx, int ny, int nz, const float* A,
                         float* B, const uint32_t* mask,
                         const orc_std_params* p, int storage) {
  int E[27][3], I[27];
  double W[27];
  std_table(q, E, I, W);
  const int64_t N = (int64_t)nx * ny * nz;
#define LOAD(ARR, IDX, DIR) ((ARR)[IDX])
#define STORE(ARR, IDX, DIR, VAL) ((ARR)[IDX] = (VAL))
  STD_STEP_BODY(float)
#undef LOAD
#undef STORE
}

/* The step of a lattice with shifted 16-bit storage, in the GPU library's
 * shifted form: the stored g_i = f_i - w_i are collided as
 *   drho = sum g, u = (sum e g + F/2) / (1 + drho),
 *   tm1 = (4.5 e.u + 3) e.u - 1.5 u.u                          (t - 1),
 *   g*_i = g_i + omega (w_i (drho tm1 + (tm1 + drho)) - g_i)
 *                         (= w rho t - w - g for rho = 1 + drho;
 *                          16-bit storage: w_i (tm1 rho + drho))
 *          [+ gf3_i + gw_i (-3 u.F) + gf9_i e.u   with a body force]
 * (gw = (1 - omega/2) w, gf3 = 3 gw e.F, gf9 = 9 gw e.F), each operation
 * rounded to float; solid nodes store the rest state g = 0.  Working on
 * g and drho -- never on f or on numbers near 1 -- keeps the rounding error
 * of the collision proportional to the deviations, so increments far below
 * f survive. */
static void std_step_shifted(int q, int nx, int ny, int nz, const float* A,
                             float* B, const uint32_t* mask,
                             const orc_std_params* p, int storage) {
  int E[27][3], I[27];
  double W[27];
  std_table(q, E, I, W);
  const int64_t N = (int64_t)nx * ny * nz;
  const int dims[3] = {nx, ny, nz};
  const float omega = 1.0f / p->tau;
  const float guo = 1.0f - 0.5f * omega;
  const float F[3] = {p->force[0], p->force[1], p->force[2]};
  const int forced = F[0] != 0.0f || F[1] != 0.0f || F[2] != 0.0f;
  const int lax = p->lid_face >= 0 ? p->lid_face / 2 : 0;
  const int lsg = p->lid_face >= 0 ? (p->lid_face % 2 ? 1 : -1) : 0;
  float w[27], lidt[27], gw[27], gf3[27], gf9[27];
  for (int i = 0; i < q; ++i) {
    w[i] = (float)W[i];
    lidt[i] = 6.0f * w[i] * p->wall_rho *
              ((float)E[i][0] * p->lid_u[0] + (float)E[i][1] * p->lid_u[1] +
               (float)E[i][2] * p->lid_u[2]);
    const float eF = (float)E[i][0] * F[0] + (float)E[i][1] * F[1] + (float)E[i][2] * F[2];
    gw[i] = guo * w[i];
    gf3[i] = gw[i] * (3.0f * eF);
    gf9[i] = gw[i] * (9.0f * eF);
  }
#pragma omp parallel for schedule(static)
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t n = idx3(x, y, z, ny, nz);
        const uint32_t m = mask[n];
        if (m & ORC_MASK_SOLID) {
          for (int i = 0; i < q; ++i) B[i * N + n] = 0.0f;
          continue;
        }
        const int c[3] = {x, y, z};
        float g[27];
        for (int i = 0; i < q; ++i) {
          if ((m >> i) & 1u) {
            g[i] = A[I[i] * N + n];
            if ((m & ORC_MASK_LID) && E[i][lax] == -lsg) g[i] += lidt[i];
          } else {
            int s3[3];
            for (int a = 0; a < 3; ++a) {
              s3[a] = c[a] - E[i][a];
              if (s3[a] < 0) s3[a] += dims[a];
              if (s3[a] >= dims[a]) s3[a] -= dims[a];
            }
            g[i] = A[i * N + idx3(s3[0], s3[1], s3[2], ny, nz)];
          }
        }
        float drho = 0.0f, jx = 0.0f, jy = 0.0f, jz = 0.0f;
        for (int i = 0; i < q; ++i) {
          drho += g[i];
          jx += (float)E[i][0] * g[i];
          jy += (float)E[i][1] * g[i];
          jz += (float)E[i][2] * g[i];
        }
        const float rho = 1.0f + drho;
        const float ir = 1.0f / rho;
        const float ux = (jx + 0.5f * F[0]) * ir, uy = (jy + 0.5f * F[1]) * ir,
                    uz = (jz + 0.5f * F[2]) * ir;
        const float m15usq = -1.5f * (ux * ux + uy * uy + uz * uz);
        const float m3uF = -3.0f * (ux * F[0] + uy * F[1] + uz * F[2]);
        for (int i = 0; i < q; ++i) {
          const float eu = (float)E[i][0] * ux + (float)E[i][1] * uy + (float)E[i][2] * uz;
          /* t - 1, and w rho t - w = w ((t - 1) + drho + drho (t - 1)), all
           * formed from small terms -- for 16-bit storage as
           * w ((t - 1) rho + drho) with rho = 1 + drho rounded once, whose
           * error (6e-8 of the term) lies far below the 16-bit value's
           * resolution */
          const float tm1 = (4.5f * eu + 3.0f) * eu + m15usq;
          const float d = (storage == 2 || storage == 3) ? w[i] * (tm1 * rho + drho) - g[i]
                                  : w[i] * (drho * tm1 + (tm1 + drho)) - g[i];
          float out = omega * d + g[i];
          if (forced) {
            out = out + gf3[i];
            out = gw[i] * m3uF + out;
            out = gf9[i] * eu + out;
          }
          B[i * N + n] = orc_store16(storage, out, forced, orc_sr_mul(i), orc_sr_hash(n, q));
        }
      }
}

static void std_step_f64(int q, int nx, int ny, int nz, const double* A,
                         double* B, const uint32_t* mask,
                         const orc_std_params* p) {
  int E[27][3], I[27];
  double W[27];
  std_table(q, E, I, W);
  const int64_t N = (int64_t)nx * ny * nz;
#define LOAD(ARR, IDX, DIR) ((ARR)[IDX])
#define STORE(ARR, IDX, DIR, VAL) ((ARR)[IDX] = (VAL))
  STD_STEP_BODY(double)
#undef LOAD
#undef STORE
}

void orc_std_run_f32(int q, int nx, int ny, int nz, float* fA,
                     const uint32_t* mask, const orc_std_params* p, int steps,
                     int storage) {
  int E[27][3], I[27];
  double W[27];
  std_table(q, E, I, W);
  const int64_t N = (int64_t)nx * ny * nz;
  const size_t bytes = sizeof(float) * (size_t)q * (size_t)N;
  float* A = (float*)malloc(bytes);
  float* B = (float*)malloc(bytes);
  /* storage values: shifted and rounded for the FP16/BF16 emulation */
  for (int i = 0; i < q; ++i)
    for (int64_t n = 0; n < N; ++n) {
      const float v = fA[i * N + n];
      A[i * N + n] = storage ? orc_store16(storage, v - (float)W[i], 0, 0u, 0u) : v;
    }
  memcpy(B, A, bytes);
  for (int s = 0; s < steps; ++s) {
    if (storage)
      std_step_shifted(q, nx, ny, nz, A, B, mask, p, storage);
    else
      std_step_f32(q, nx, ny, nz, A, B, mask, p, 0);
    float* t = A;
    A = B;
    B = t;
  }
  for (int i = 0; i < q; ++i)
    for (int64_t n = 0; n < N; ++n)
      fA[i * N + n] = storage ? A[i * N + n] + (float)W[i] : A[i * N + n];
  free(A);
  free(B);
}

void orc_std_run_f64(int q, int nx, int ny, int nz, double* fA,
                     const uint32_t* mask, const orc_std_params* p, int steps) {
  const int64_t N = (int64_t)nx * ny * nz;
  const size_t bytes = sizeof(double) * (size_t)q * (size_t)N;
  double* B = (double*)malloc(bytes);
  double* A = fA;
  memcpy(B, A, bytes);
  for (int s = 0; s < steps; ++s) {
    std_step_f64(q, nx, ny, nz, A, B, mask, p);
    double* t = A;
    A = B;
    B = t;
  }
  if (A != fA) {
    memcpy(fA, A, bytes);
    free(A);
  } else {
    free(B);
  }
}

void orc_std_moments_f64(int q, int nx, int ny, int nz, const double* f,
                         const uint32_t* mask, const float force[3],
                         double* rho, double* u) {
  int E[27][3], I[27];
  double W[27];
  std_table(q, E, I, W);
  const int64_t N = (int64_t)nx * ny * nz;
  for (int64_t n = 0; n < N; ++n) {
    if (mask && (mask[n] & ORC_MASK_SOLID)) {
      rho[n] = 0.0;
      u[n] = u[N + n] = u[2 * N + n] = 0.0;
      continue;
    }
    double r = 0, jx = 0, jy = 0, jz = 0;
    for (int i = 0; i < q; ++i) {
      const double fi = f[i * N + n];
      r += fi;
      jx += fi * E[i][0];
      jy += fi * E[i][1];
      jz += fi * E[i][2];
    }
    rho[n] = r;
    u[n] = (jx - 0.5 * force[0]) / r;
    u[N + n] = (jy - 0.5 * force[1]) / r;
    u[2 * N + n] = (jz - 0.5 * force[2]) / r;
  }
}

/* ===================================================================== */
/* MT19937-64 (Matsumoto & Nishimura 2004) and random spheres             */
/* ===================================================================== */
typedef struct {
  uint64_t mt[312];
  int mti;
} mt64;

static void mt64_seed(mt64* s, uint64_t seed) {
  s->mt[0] = seed;
  for (int i = 1; i < 312; ++i)
    s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) +
               (uint64_t)i;
  s->mti = 312;
}

static uint64_t mt64_next(mt64* s) {
  static const uint64_t MAG[2] = {0ULL, 0xB5026F5AA96619E9ULL};
  const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
  if (s->mti >= 312) {
    int i;
    uint64_t x;
    for (i = 0; i < 312 - 156; ++i) {
      x = (s->mt[i] & UM) | (s->mt[i + 1] & LM);
      s->mt[i] = s->mt[i + 156] ^ (x >> 1) ^ MAG[x & 1ULL];
    }
    for (; i < 311; ++i) {
      x = (s->mt[i] & UM) | (s->mt[i + 1] & LM);
      s->mt[i] = s->mt[i + (156 - 312)] ^ (x >> 1) ^ MAG[x & 1ULL];
    }
    x = (s->mt[311] & UM) | (s->mt[0] & LM);
    s->mt[311] = s->mt[155] ^ (x >> 1) ^ MAG[x & 1ULL];
    s->mti = 0;
  }
  uint64_t x = s->mt[s->mti++];
  x ^= (x >> 29) & 0x5555555555555555ULL;
  x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
  x ^= (x << 37) & 0xFFF7EEE000000000ULL;
  x ^= (x >> 43);
  return x;
}

uint64_t orc_mt64_nth(uint64_t seed, int64_t n) {
  mt64 s;
  mt64_seed(&s, seed);
  uint64_t v = 0;
  for (int64_t i = 0; i < n; ++i) v = mt64_next(&s);
  return v;
}

static double mt64_u01(mt64* s) {
  return (double)(mt64_next(s) >> 11) * (1.0 / 9007199254740992.0);
}

int64_t orc_random_spheres(int nx, int ny, int nz, uint64_t seed, double frac,
                           double rmin, double rmax, int periodic_x,
                           uint8_t* solid, int64_t max_spheres,
                           double* spheres) {
  const int64_t N = (int64_t)nx * ny * nz;
  memset(solid, 0, (size_t)N);
  mt64 s;
  mt64_seed(&s, seed);
  int64_t count = 0, ns = 0;
  while ((double)count < frac * (double)N) {
    const double cx = mt64_u01(&s) * nx;
    const double cy = mt64_u01(&s) * ny;
    const double cz = mt64_u01(&s) * nz;
    const double r = rmin + (rmax - rmin) * mt64_u01(&s);
    if (ns < max_spheres && spheres) {
      spheres[4 * ns] = cx;
      spheres[4 * ns + 1] = cy;
      spheres[4 * ns + 2] = cz;
      spheres[4 * ns + 3] = r;
    }
    ++ns;
    const int x0 = (int)floor(cx - r), x1 = (int)ceil(cx + r);
    const int y0 = (int)floor(cy - r), y1 = (int)ceil(cy + r);
    const int z0 = (int)floor(cz - r), z1 = (int)ceil(cz + r);
    for (int x = x0; x <= x1; ++x) {
      int xi = x;
      if (xi < 0 || xi >= nx) {
        if (!periodic_x) continue;
        xi = ((xi % nx) + nx) % nx;
      }
      const double dx = (double)x - cx;
      for (int y = y0 < 0 ? 0 : y0; y <= y1 && y < ny; ++y) {
        const double dy = (double)y - cy;
        for (int z = z0 < 0 ? 0 : z0; z <= z1 && z < nz; ++z) {
          const double dz = (double)z - cz;
          if (dx * dx + dy * dy + dz * dz <= r * r) {
            const int64_t n = idx3(xi, y, z, ny, nz);
            if (!solid[n]) {
              solid[n] = 1;
              ++count;
            }
          }
        }
      }
    }
  }
  return ns;
}
