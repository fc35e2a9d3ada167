// Quaternion rotation of simulation/source/Geometry.cpp:7-42, in float as
// there (the float overloads of sqrt/cos/sin).
#include "Geometry.h"

#include <array>
#include <cmath>

namespace {

using Quat = std::array<float, 4>;  // (w, x, y, z)

// Hamilton product a * b.
Quat hamilton(const Quat& a, const Quat& b) {
  return Quat{(a[0] * b[0]) - (a[1] * b[1]) - (a[2] * b[2]) - (a[3] * b[3]),
              (a[0] * b[1]) + (a[1] * b[0]) + (a[2] * b[3]) - (a[3] * b[2]),
              (a[0] * b[2]) - (a[1] * b[3]) + (a[2] * b[0]) + (a[3] * b[1]),
              (a[0] * b[3]) + (a[1] * b[2]) - (a[2] * b[1]) + (a[3] * b[0])};
}

}  // namespace

Point_3 rotateVectorAroundAxis(Point_3 axis, Point_3 vector, float angle) {
  const float len = std::sqrt((axis.x * axis.x) + (axis.y * axis.y) + (axis.z * axis.z));
  axis.x /= len;
  axis.y /= len;
  axis.z /= len;
  const float half = (float)((double)angle / 2.0);
  const float c = std::cos(half), s = std::sin(half);
  const float ci = std::cos(-half), si = std::sin(-half);
  const Quat q{c, s * axis.x, s * axis.y, s * axis.z};
  const Quat q_inv{ci, si * axis.x, si * axis.y, si * axis.z};
  const Quat v{0.0f, vector.x, vector.y, vector.z};
  const Quat r = hamilton(hamilton(q, v), q_inv);
  return Point_3(r[1], r[2], r[3]);
}
