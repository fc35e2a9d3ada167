// Triangle-mesh model without CGAL (reference: simulation/source/Model.cpp:22-91).
//
//   importModel        Wavefront OBJ text -> vertices + triangles
//   isModelClosed      every directed edge a->b has its opposite b->a
//   isPointInsideModel on the surface -> inside; else generalized winding
//                      number |W| > 1/2 (solid angles, Van Oosterom &
//                      Strackee 1983)
//   reflectionVector   first triangle hit by the ray p1 -> p2 (hits within
//                      1e-9 of the nearest, i.e. through a shared edge or
//                      vertex, go to the lowest face index), and the unit
//                      vector p1 - p2 rotated by pi about its normal
//
// Geometry tests run in double on the float vertex coordinates; a point is
// "on" a triangle only when it lies in its plane up to double rounding, so
// points a float ulp away from a face are decided like CGAL's float kernel
// decides them (outside when they are outside).
#include "Model.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <utility>

namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

V3 vert(const TriangleMesh& m, int v) {
  return {m.vertices[3 * v], m.vertices[3 * v + 1], m.vertices[3 * v + 2]};
}

// p on triangle (a, b, c): coplanar up to rounding and inside all three
// edges (edge functions of the triangle's own normal).
bool on_triangle(V3 p, V3 a, V3 b, V3 c) {
  const V3 n = cross(b - a, c - a);
  const double nn = norm(n);
  if (nn == 0.0) return false;
  const double off = dot(n, p - a);
  const double scale = nn * (norm(p - a) + norm(b - a));
  if (std::fabs(off) > 1e-12 * scale) return false;
  const double tol = -1e-12 * nn * nn;
  return dot(cross(b - a, p - a), n) >= tol && dot(cross(c - b, p - b), n) >= tol &&
         dot(cross(a - c, p - c), n) >= tol;
}

// Signed solid angle of triangle (a, b, c) seen from p.
double solid_angle(V3 p, V3 a, V3 b, V3 c) {
  const V3 ra = a - p, rb = b - p, rc = c - p;
  const double la = norm(ra), lb = norm(rb), lc = norm(rc);
  const double num = dot(ra, cross(rb, rc));
  const double den = la * lb * lc + dot(ra, rb) * lc + dot(rb, rc) * la + dot(rc, ra) * lb;
  return 2.0 * std::atan2(num, den);
}

// Ray o + t d, t >= 0, against triangle (a, b, c): plane intersection then
// edge functions.  Returns t or a negative value for no hit.
double ray_hit(V3 o, V3 d, V3 a, V3 b, V3 c) {
  const V3 n = cross(b - a, c - a);
  const double denom = dot(n, d);
  const double scale = norm(n) * norm(d);
  if (scale == 0.0 || std::fabs(denom) <= 1e-14 * scale) return -1.0;
  const double t = dot(n, a - o) / denom;
  if (t < 0.0) return -1.0;
  const V3 q{o.x + t * d.x, o.y + t * d.y, o.z + t * d.z};
  const double tol = -1e-12 * dot(n, n);
  if (dot(cross(b - a, q - a), n) < tol || dot(cross(c - b, q - b), n) < tol ||
      dot(cross(a - c, q - c), n) < tol)
    return -1.0;
  return t;
}

bool parse_index(const std::string& tok, int nverts, int& out) {
  const char* s = tok.c_str();
  char* end = nullptr;
  const long v = std::strtol(s, &end, 10);
  if (end == s) return false;
  out = v > 0 ? (int)(v - 1) : (int)(nverts + v);
  return true;
}

}  // namespace

void Model::importModel(std::string filename) {
  modelMesh = TriangleMesh{};
  meshValid = true;
  closed = false;
  loaded = false;
  std::ifstream in(filename);
  if (!in) {
    // The reference returns early and keeps a mesh-less model that later
    // dereferences a null tree; here the model is simply invalid.
    meshValid = false;
    return;
  }
  bool triangles_only = true;
  std::string line;
  while (std::getline(in, line)) {
    const std::size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream ls(line);
    std::string tag;
    if (!(ls >> tag)) continue;
    if (tag == "v") {
      float x, y, z;
      if (ls >> x >> y >> z) {
        modelMesh.vertices.push_back(x);
        modelMesh.vertices.push_back(y);
        modelMesh.vertices.push_back(z);
      }
    } else if (tag == "f") {
      std::vector<int> ids;
      std::string tok;
      const int nv = (int)modelMesh.number_of_vertices();
      while (ls >> tok) {
        int id;
        if (!parse_index(tok, nv, id)) break;
        ids.push_back(id);
      }
      if (ids.size() != 3) {
        triangles_only = false;
        continue;
      }
      for (int id : ids) modelMesh.faces.push_back(id);
    }
  }
  loaded = true;
  const int nv = (int)modelMesh.number_of_vertices();
  for (int id : modelMesh.faces)
    if (id < 0 || id >= nv) triangles_only = false;
  // closed: each directed edge exactly once, and its opposite present
  std::map<std::pair<int, int>, int> edges;
  for (std::size_t t = 0; t < modelMesh.number_of_faces(); ++t)
    for (int k = 0; k < 3; ++k)
      ++edges[{modelMesh.faces[3 * t + k], modelMesh.faces[3 * t + (k + 1) % 3]}];
  closed = triangles_only && !edges.empty();
  for (const auto& e : edges) {
    if (!closed) break;
    auto it = edges.find({e.first.second, e.first.first});
    if (e.second != 1 || it == edges.end() || it->second != 1) closed = false;
  }
  meshValid = triangles_only && closed;
}

bool Model::isModelClosed() { return closed; }

bool Model::isPointInsideModel(Point_3 point) {
  if (!meshValid) return false;
  const V3 p{point.x, point.y, point.z};
  const std::size_t nf = modelMesh.number_of_faces();
  for (std::size_t t = 0; t < nf; ++t) {
    const V3 a = vert(modelMesh, modelMesh.faces[3 * t]);
    const V3 b = vert(modelMesh, modelMesh.faces[3 * t + 1]);
    const V3 c = vert(modelMesh, modelMesh.faces[3 * t + 2]);
    if (on_triangle(p, a, b, c)) return true;
  }
  double omega = 0.0;
  for (std::size_t t = 0; t < nf; ++t) {
    omega += solid_angle(p, vert(modelMesh, modelMesh.faces[3 * t]),
                         vert(modelMesh, modelMesh.faces[3 * t + 1]),
                         vert(modelMesh, modelMesh.faces[3 * t + 2]));
  }
  const double winding = omega / (4.0 * M_PI);
  return std::fabs(winding) > 0.5;
}

MeshBounds Model::bounding_box() {
  MeshBounds b{};
  const std::size_t nv = modelMesh.number_of_vertices();
  for (int a = 0; a < 3; ++a) {
    b.lo[a] = nv ? modelMesh.vertices[a] : 0.0f;
    b.hi[a] = b.lo[a];
  }
  for (std::size_t v = 1; v < nv; ++v)
    for (int a = 0; a < 3; ++a) {
      const float c = modelMesh.vertices[3 * v + a];
      if (c < b.lo[a]) b.lo[a] = c;
      if (c > b.hi[a]) b.hi[a] = c;
    }
  return b;
}

Point_3 Model::reflectionVector(Point_3 point_1, Point_3 point_2) {
  Point_3 v{point_1.x - point_2.x, point_1.y - point_2.y, point_1.z - point_2.z};
  v.normaliseVector();
  const V3 o{point_1.x, point_1.y, point_1.z};
  const V3 d{(double)point_2.x - point_1.x, (double)point_2.y - point_1.y,
             (double)point_2.z - point_1.z};
  long best = -1;
  double best_t = std::numeric_limits<double>::infinity();
  const std::size_t nf = modelMesh.number_of_faces();
  for (std::size_t t = 0; t < nf; ++t) {
    const double h = ray_hit(o, d, vert(modelMesh, modelMesh.faces[3 * t]),
                             vert(modelMesh, modelMesh.faces[3 * t + 1]),
                             vert(modelMesh, modelMesh.faces[3 * t + 2]));
    if (h >= 0.0 && h < best_t) {
      best_t = h;
      best = (long)t;
    }
  }
  if (best < 0) return Point_3(0.0f, 0.0f, 0.0f);
  // a ray through an edge or vertex shared by several faces: the lowest
  // face index among the hits within 1e-9 of the nearest one
  for (long t = 0; t < best; ++t) {
    const double h = ray_hit(o, d, vert(modelMesh, modelMesh.faces[3 * t]),
                             vert(modelMesh, modelMesh.faces[3 * t + 1]),
                             vert(modelMesh, modelMesh.faces[3 * t + 2]));
    if (h >= 0.0 && h <= best_t + 1e-9) {
      best = t;
      break;
    }
  }
  // face normal (v1 - v0) x (v2 - v0) in float, as a float-kernel
  // compute_face_normal gives it (its sign does not matter for a pi turn)
  const float* A = &modelMesh.vertices[3 * modelMesh.faces[3 * best]];
  const float* B = &modelMesh.vertices[3 * modelMesh.faces[3 * best + 1]];
  const float* C = &modelMesh.vertices[3 * modelMesh.faces[3 * best + 2]];
  const float e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  const float e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
  const Point_3 n((e1[1] * e2[2]) - (e1[2] * e2[1]), (e1[2] * e2[0]) - (e1[0] * e2[2]),
                  (e1[0] * e2[1]) - (e1[1] * e2[0]));
  return rotateVectorAroundAxis(n, v, (float)M_PI);
}
