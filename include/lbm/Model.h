// Triangle-mesh obstacle model (reference: simulation/include/Model.h:23-47,
// simulation/source/Model.cpp:22-91).  The reference builds on CGAL
// (Surface_mesh, Side_of_triangle_mesh, AABB_tree); this is a self-contained
// re-design with the same semantics:
//   - importModel reads Wavefront OBJ text; the mesh is valid when every
//     face is a triangle and the surface is closed (every edge a->b has its
//     opposite b->a);
//   - isPointInsideModel: points on the surface count as inside
//     (Side_of_triangle_mesh ON_BOUNDARY); otherwise the generalized
//     winding number decides;
//   - reflectionVector: the first triangle hit by the ray from point_1
//     through point_2, and the unit vector point_1 - point_2 rotated by pi
//     about that triangle's normal; (0,0,0) when nothing is hit.
#pragma once

#include <string>
#include <vector>

#include "Geometry.h"

struct MeshBounds {
  float lo[3], hi[3];
  float xmin() const { return lo[0]; }
  float ymin() const { return lo[1]; }
  float zmin() const { return lo[2]; }
  float xmax() const { return hi[0]; }
  float ymax() const { return hi[1]; }
  float zmax() const { return hi[2]; }
};

struct TriangleMesh {
  std::vector<float> vertices;  // 3 per vertex
  std::vector<int> faces;       // 3 per triangle
  std::size_t number_of_vertices() const { return vertices.size() / 3; }
  std::size_t number_of_faces() const { return faces.size() / 3; }
};

class Model {
 public:
  void importModel(std::string filename);
  bool isModelClosed();
  bool isPointInsideModel(Point_3 point);
  const TriangleMesh& getMesh() const { return modelMesh; }
  bool checkError() { return !meshValid; }
  MeshBounds bounding_box();
  Point_3 reflectionVector(Point_3 point_1, Point_3 point_2);
  bool isLoaded() const { return loaded; }

 private:
  TriangleMesh modelMesh;
  bool meshValid = true;
  bool closed = false;
  bool loaded = false;
};
