// Re-expression of the reference's test/CoordinatesTest.cpp:14-229 through
// the Lattice facade (GPU).
#include <cstdio>
#include <memory>

#include "Constants.h"
#include "Lattice.h"
#include "minitest.h"
#include "test_data.h"

TEST(CoordinatesTest, TestLoadAndRetrieve) {
  dim3 threads(1, 2, 1), blocks(1, 1, 1);
  Lattice lat(1, 2, 1, blocks, threads, FluidData(1.0, 1.0), 0.1f);
  std::unique_ptr<LatticePoint[]> pts(new LatticePoint[2]());
  pts[0].particle_distribution[5] = 2.0f;
  lat.load_data(pts.get());
  std::unique_ptr<LatticePoint[]> back(lat.retrieve_data());
  EXPECT_NEAR(back[0].particle_distribution[5], pts[0].particle_distribution[5], 1e-6);
}

static std::unique_ptr<LatticePoint[]> grid_points(int n) {
  std::unique_ptr<LatticePoint[]> pts(new LatticePoint[(size_t)n * n * n]());
  int k = 0;
  for (int x = 0; x < n; ++x)
    for (int y = 0; y < n; ++y)
      for (int z = 0; z < n; ++z, ++k) {
        pts[k].x = x;
        pts[k].y = y;
        pts[k].z = z;
      }
  return pts;
}

TEST(CoordinatesTest, TestCoordsMatch) {
  dim3 threads(10, 10, 10), blocks(10, 10, 10);
  Lattice lat(100, 100, 100, blocks, threads, FluidData(1.0, 1.0), 0.1f);
  auto pts = grid_points(100);
  lat.load_data(pts.get());
  std::unique_ptr<LatticePoint[]> back(lat.retrieve_data());
  int bad = 0;
  for (int x = 0; x < 100; ++x)
    for (int y = 0; y < 100; ++y)
      for (int z = 0; z < 100; ++z) {
        const LatticePoint& p = back[z + (y * 100) + (x * 10000)];
        bad += p.x != x || p.y != y || p.z != z;
      }
  EXPECT_EQ(bad, 0);
}

TEST(CoordinatesTest, TestCoordinatesOnCopy) {
  dim3 threads(10, 10, 10), blocks(10, 10, 10);
  Lattice lat(100, 100, 100, blocks, threads, FluidData(1.0, 1.0), 0.1f);
  auto pts = grid_points(100);
  lat.load_data(pts.get());
  bool result = true;
  // the reference's spelling of the call (CoordinatesTest.cpp:130)
  RunCudaTestFunctions::run_test_coordinates(threads, blocks, lat.getCudaDataPointer(), 100,
                                             &result);
  EXPECT_TRUE(result);
}

TEST(CoordinatesTest, TestCoordinatesWithModel) {
  dim3 threads(1, 1, 1), blocks(10, 10, 10);
  Lattice lat(10, 10, 10, blocks, threads, FluidData(1.0, 1.0), 0.1f);
  std::unique_ptr<LatticePoint[]> pts(new LatticePoint[1000]());
  lat.load_data(pts.get());
  lat.insertModel(data_file("compat_cube.obj"));
  std::unique_ptr<LatticePoint[]> back(lat.retrieve_data());
  for (int y = 0; y < 10; ++y)
    for (int x = 0; x < 10; ++x) EXPECT_FALSE(back[(10 * y) + x].isInternal);
  for (int y = 1; y < 9; ++y)
    for (int x = 1; x < 9; ++x) EXPECT_TRUE(back[500 + (10 * y) + x].isInternal);
}

TEST(CoordinatesTest, TestDistributeVector) {
  dim3 one(1, 1, 1);
  Lattice lat(1, 1, 1, one, one, FluidData(1.0, 1.0), 0.1f);
  auto r = lat.distributeVector(Point_3(1.0f, 0.1f, 0.1f));
  const float norm = r[0].first + r[1].first + r[2].first;
  EXPECT_NEAR(r[0].first / norm, 0.4134132615337746, CONSTANTS::GEOMETRIC_TOLERANCE);
  EXPECT_NEAR(r[1].first / norm, 0.2932933692331128, CONSTANTS::GEOMETRIC_TOLERANCE);
  EXPECT_NEAR(r[2].first / norm, 0.2932933692331128, CONSTANTS::GEOMETRIC_TOLERANCE);
  EXPECT_EQ(r[0].second, 6);
  EXPECT_EQ(r[1].second, 18);
  EXPECT_EQ(r[2].second, 17);
}

TEST(CoordinatesTest, TestReflection) {
  dim3 threads(3, 3, 3), blocks(1, 1, 1);
  Lattice lat(3, 3, 3, blocks, threads, FluidData(1.0, 1.0), 0.5f);
  std::unique_ptr<LatticePoint[]> pts(new LatticePoint[27]());
  lat.load_data(pts.get());
  lat.insertModel(data_file("cube_angled.obj"));
  lat.simulateCollision();
  lat.preProcessModel();
  std::unique_ptr<LatticePoint[]> back(lat.retrieve_data());
  LatticePoint& centre = back[13];
  EXPECT_TRUE(centre.reflections != nullptr);
  if (!centre.reflections) return;
  EXPECT_NEAR(centre.reflections->reflection_weight[1 * 3 + 0], 0.367402, 1e-6);
  EXPECT_NEAR(centre.reflections->reflection_weight[1 * 3 + 1], 0.341121, 1e-6);
  EXPECT_NEAR(centre.reflections->reflection_weight[1 * 3 + 2], 0.291477, 1e-6);
  EXPECT_EQ(lat.getReflectionDirection(&centre, 3), 17);
  EXPECT_EQ(lat.getReflectionDirection(&centre, 4), 4);
  EXPECT_EQ(lat.getReflectionDirection(&centre, 5), 26);
  EXPECT_TRUE(centre.isReflected && centre.isInternal);
}

// Not in the reference: the record-stream file and the snapshot file
// round-trip through a second lattice.
TEST(CoordinatesTest, TestRecordAndSnapshotFiles) {
  dim3 threads(4, 3, 5), blocks(2, 2, 1);
  const int n = 8 * 6 * 5;
  Lattice a(8, 6, 5, blocks, threads, FluidData(1.0, 0.9), 0.1f);
  std::unique_ptr<LatticePoint[]> pts(new LatticePoint[n]());
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < 27; ++i) pts[k].particle_distribution[i] = 0.001f * (1 + (k * 27 + i) % 97);
  a.load_data(pts.get());
  a.writeRecords("/tmp/lbm_test_records.bin");
  a.simulateLattice();
  a.saveSnapshot("/tmp/lbm_test_snapshot.bin");
  a.simulateLattice();
  std::unique_ptr<LatticePoint[]> want(a.retrieve_data());

  Lattice b(8, 6, 5, blocks, threads, FluidData(1.0, 0.9), 0.1f);
  b.readRecords("/tmp/lbm_test_records.bin");
  std::unique_ptr<LatticePoint[]> first(b.retrieve_data());
  int bad = 0;
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < 27; ++i)
      bad += first[k].particle_distribution[i] != pts[k].particle_distribution[i];
  EXPECT_EQ(bad, 0);
  b.loadSnapshot("/tmp/lbm_test_snapshot.bin");
  b.simulateLattice();
  std::unique_ptr<LatticePoint[]> got(b.retrieve_data());
  bad = 0;
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < 27; ++i)
      bad += got[k].particle_distribution[i] != want[k].particle_distribution[i];
  EXPECT_EQ(bad, 0);
  std::remove("/tmp/lbm_test_records.bin");
  std::remove("/tmp/lbm_test_snapshot.bin");
}

int main(int argc, char** argv) { return RUN_ALL_TESTS(); }
