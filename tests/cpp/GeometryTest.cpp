// Re-expression of the reference's test/GeometryTest.cpp:11-32 (quaternion
// rotation) against the facade.
#include <cmath>

#include "Constants.h"
#include "Geometry.h"
#include "minitest.h"

TEST(GeometryTest, TestQuaternionRotation) {
  const Point_3 axis(std::cos(M_PI / 4), std::sin(M_PI / 4), 0.0);
  const Point_3 x_axis(1.0, 0.0, 0.0);
  // half turn about the diagonal of the xy plane maps x onto y
  Point_3 r = rotateVectorAroundAxis(axis, x_axis, (float)M_PI);
  EXPECT_NEAR(r.x, 0.0, CONSTANTS::GEOMETRIC_TOLERANCE);
  EXPECT_NEAR(r.y, 1.0, CONSTANTS::GEOMETRIC_TOLERANCE);
  EXPECT_NEAR(r.z, 0.0, CONSTANTS::GEOMETRIC_TOLERANCE);
  // quarter turn
  r = rotateVectorAroundAxis(axis, x_axis, (float)(M_PI / 2));
  EXPECT_NEAR(r.x, 0.5, CONSTANTS::GEOMETRIC_TOLERANCE);
  EXPECT_NEAR(r.y, 0.5, CONSTANTS::GEOMETRIC_TOLERANCE);
  EXPECT_NEAR(r.z, -std::cos(M_PI / 4), CONSTANTS::GEOMETRIC_TOLERANCE);
}

TEST(GeometryTest, TestNormalise) {
  Point_3 v(3.0f, 0.0f, 4.0f);
  v.normaliseVector();
  EXPECT_NEAR(v.x, 0.6, 1e-7);
  EXPECT_NEAR(v.z, 0.8, 1e-7);
  EXPECT_TRUE(Point_3(1, 2, 3) == Point_3(1, 2, 3));
}

int main(int argc, char** argv) { return RUN_ALL_TESTS(); }
