// Re-expression of the reference's test/CalculationsTest.cpp:19-284 through
// the Lattice facade (GPU: every stage runs in liblbm_hip.so).
#include <memory>

#include "FlowCriterion.h"
#include "FlowSurface.h"
#include "Lattice.h"
#include "Utils.h"
#include "minitest.h"
#include "test_data.h"

namespace {
int at(Lattice& l, int x, int y, int z) {
  const dim3 d = l.getDimensions();
  return z + (y * d.z) + (x * d.z * d.y);
}
}  // namespace

TEST(CalculationsTest, TestStreaming) {
  dim3 threads(1, 2, 1), blocks(1, 1, 1);
  Lattice lat(1, 2, 1, blocks, threads, FluidData(1.0, 1.0), 0.1f);
  LatticeData data(lat.getDeviceHandle(), lat.getDimensions());
  std::unique_ptr<LatticePoint[]> pts(new LatticePoint[2]());
  pts[0].particle_distribution[5] = 2.0f;
  lat.load_data(pts.get());
  RunHipTestFunctions::run_prime_points(blocks, threads, data);
  lat.simulateStreaming();
  pts.reset(lat.retrieve_data());
  EXPECT_NEAR(pts[1].particle_distribution[5], 2.0, 1e-6);
}

TEST(CalculationsTest, TestCollision) {
  dim3 threads(1, 1, 1), blocks(1, 1, 1);
  Lattice lat(1, 1, 1, blocks, threads, FluidData(1.0, 1.0), 0.1f);
  std::unique_ptr<LatticePoint[]> pts(new LatticePoint[1]());
  pts[0].particle_distribution[0] = 2.0f;
  lat.load_data(pts.get());
  lat.simulateCollision();
  pts.reset(lat.retrieve_data());
  EXPECT_NEAR(pts[0].particle_distribution[0], 0.592592593, 1e-6);
  EXPECT_NEAR(pts[0].particle_distribution[1], 0.148148148, 1e-6);
  EXPECT_NEAR(pts[0].particle_distribution[2], 0.148148148, 1e-6);
}

TEST(CalculationsTest, TestFull) {
  dim3 threads(1, 2, 1), blocks(1, 1, 1);
  Lattice lat(1, 2, 1, blocks, threads, FluidData(1.0, 1.0), 0.1f);
  std::unique_ptr<LatticePoint[]> pts(new LatticePoint[2]());
  pts[0].particle_distribution[5] = 2.0f;
  lat.load_data(pts.get());
  lat.simulateLattice();
  pts.reset(lat.retrieve_data());
  EXPECT_NEAR(pts[1].particle_distribution[0], -0.296296296, 1e-6);
  EXPECT_NEAR(pts[1].particle_distribution[1], -0.074074074, 1e-6);
  EXPECT_NEAR(pts[1].particle_distribution[2], 0.148148148, 1e-6);
}

static void reflection_case(bool with_force) {
  dim3 threads(3, 3, 3), blocks(1, 1, 1);
  Lattice lat(3, 3, 3, blocks, threads, FluidData(1.0, 1.0), 0.5f);
  lat.insertModel(data_file("cube_angled.obj"));
  lat.preProcessModel();
  std::unique_ptr<LatticePoint[]> pts(lat.retrieve_data());
  pts[at(lat, 1, 1, 1)].particle_distribution[1] = 1.0f;
  lat.load_data(pts.get());
  ReflectionData force;
  if (with_force) lat.setReflectionData(&force);
  lat.simulateReflections();
  pts.reset(lat.retrieve_data());
  EXPECT_NEAR(pts[at(lat, 2, 1, 2)].particle_distribution[17], 0.367402, 1e-6);
  EXPECT_NEAR(pts[at(lat, 1, 1, 2)].particle_distribution[4], 0.341121, 1e-6);
  EXPECT_NEAR(pts[at(lat, 2, 2, 2)].particle_distribution[26], 0.291477, 1e-6);
  EXPECT_NEAR(pts[at(lat, 1, 1, 1)].particle_distribution[1], 0.0, 1e-6);
  if (with_force) {
    ReflectionData* r = lat.retrieve_reflection_data();
    EXPECT_NEAR(r->x, 0.28086, 1e-6);
    EXPECT_NEAR(r->y, -0.194318, 1e-6);
    EXPECT_NEAR(r->z, 0.648261, 1e-6);
  }
}

TEST(CalculationsTest, TestReflections) { reflection_case(false); }
TEST(CalculationsTest, TestReflectionsData) { reflection_case(true); }

TEST(CalculationsTest, TestFlowGeneration) {
  dim3 threads(3, 3, 3), blocks(1, 1, 1);
  Lattice lat(3, 3, 3, blocks, threads, FluidData(1.0, 1.0), 0.5f);
  LatticePoint flow{};
  flow.particle_distribution[0] = 1.0f;
  FlowData induced(&FlowCriterion::ptestCriteria, flow);
  lat.setFlowData(&induced);
  lat.simulateFlow();
  std::unique_ptr<LatticePoint[]> pts(lat.retrieve_data());
  EXPECT_NEAR(pts[at(lat, 1, 1, 1)].particle_distribution[0], 1.0, 1e-6);
  EXPECT_NEAR(pts[at(lat, 0, 0, 0)].particle_distribution[0], 0.0, 1e-6);
}

TEST(CalculationsTest, TestPlaneFlowGeneration) {
  dim3 threads(3, 3, 3), blocks(1, 1, 1);
  Lattice lat(3, 3, 3, blocks, threads, FluidData(1.0, 1.0), 0.5f);
  LatticePoint flow{};
  flow.particle_distribution[0] = 1.0f;
  FlowData induced(&FlowCriterion::ptestPlaneCriteria, flow);
  lat.setFlowData(&induced);
  lat.simulateFlow();
  std::unique_ptr<LatticePoint[]> pts(lat.retrieve_data());
  for (int y = 0; y < 3; ++y)
    for (int z = 0; z < 3; ++z) {
      EXPECT_NEAR(pts[at(lat, 0, y, z)].particle_distribution[0], 1.0, 1e-6);
      EXPECT_NEAR(pts[at(lat, 1, y, z)].particle_distribution[0], 0.0, 1e-6);
    }
}

TEST(CalculationsTest, CompressionTest) {
  int values[6] = {1, 2, 3, 4, 5, 6};
  EXPECT_EQ(Utils::compressInts(values, 6, 5), 35754150u);
}

TEST(CalculationsTest, TestLaunchShapeChecked) {
  bool threw = false;
  try {
    Lattice bad(4, 4, 4, dim3(1, 1, 1), dim3(2, 2, 2), FluidData(), 0.1f);
  } catch (const std::invalid_argument&) {
    threw = true;
  }
  EXPECT_TRUE(threw);
}

int main(int argc, char** argv) { return RUN_ALL_TESTS(); }
