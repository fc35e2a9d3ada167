// Vector type and rotation of the reference's Geometry API
// (simulation/include/Geometry.h:8-31).  Host-only; no GPU headers.
#pragma once

#include <cmath>

struct Point_3 {
  float x, y, z;

  Point_3(float x_i, float y_i, float z_i) : x(x_i), y(y_i), z(z_i) {}

  bool operator==(const Point_3& a) const {
    return a.x == x && a.y == y && a.z == z;
  }

  void normaliseVector() {
    const float norm = std::sqrt((x * x) + (y * y) + (z * z));
    x /= norm;
    y /= norm;
    z /= norm;
  }
};

// Rotates `vector` by `angle` (radians) about `axis` with the unit
// quaternion q = (cos(a/2), sin(a/2) axis): v' = q v q^-1, in float.
Point_3 rotateVectorAroundAxis(Point_3 axis, Point_3 vector, float angle);
