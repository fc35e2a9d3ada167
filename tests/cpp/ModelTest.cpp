// Re-expression of the reference's test/ModelTest.cpp:5-74 against the
// CGAL-free Model.
#include "Model.h"
#include "minitest.h"
#include "test_data.h"

TEST(ModelTest, importModel) {
  Model m;
  m.importModel(data_file("triangle.obj"));
  EXPECT_EQ(m.getMesh().number_of_vertices(), 3u);
}

TEST(ModelTest, isModelClosed) {
  Model m;
  m.importModel(data_file("cube.obj"));
  EXPECT_TRUE(m.isModelClosed());
  m.importModel(data_file("triangle.obj"));
  EXPECT_FALSE(m.isModelClosed());
}

TEST(ModelTest, isPointInsideModel) {
  Model m;
  m.importModel(data_file("cube.obj"));
  EXPECT_TRUE(m.isPointInsideModel(Point_3(0.0f, 0.0f, 0.0f)));
  EXPECT_FALSE(m.isPointInsideModel(Point_3(2.0f, 0.0f, 0.0f)));
  // the surface itself is inside (Side_of_triangle_mesh ON_BOUNDARY)
  EXPECT_TRUE(m.isPointInsideModel(Point_3(1.0f, 0.25f, -0.5f)));
  EXPECT_TRUE(m.isPointInsideModel(Point_3(1.0f, 1.0f, 1.0f)));
}

TEST(ModelTest, checkErrorModel) {
  Model m;
  m.importModel(data_file("quad.obj"));
  EXPECT_TRUE(m.checkError());
  Model c;
  c.importModel(data_file("cube.obj"));
  EXPECT_FALSE(c.checkError());
}

TEST(ModelTest, intersectionNormal) {
  Model m;
  m.importModel(data_file("compat_cube.obj"));
  Point_3 r = m.reflectionVector(Point_3(0.5f, 0.5f, -1.0f), Point_3(0.5f, 0.5f, 1.0f));
  EXPECT_NEAR(r.x, 0.0, 1e-6);
  EXPECT_NEAR(r.y, 0.0, 1e-6);
  EXPECT_NEAR(r.z, -1.0, 1e-6);
  // pointing away from the mesh: no hit
  r = m.reflectionVector(Point_3(0.5f, 0.5f, -1.0f), Point_3(0.5f, 0.5f, -2.0f));
  EXPECT_TRUE(r == Point_3(0.0f, 0.0f, 0.0f));
}

int main(int argc, char** argv) { return RUN_ALL_TESTS(); }
