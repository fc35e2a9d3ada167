// Command-line driver with the reference's protocol
// (LatticeBoltzmannCuda.cpp:11-95):
//
//   LatticeBoltzmannHip <model.obj> tx ty tz bx by bz density tau spacing steps
//                       [--staged] [--quiet] [--time]
//
// The lattice is (tx*bx) x (ty*by) x (tz*bz) D3Q27 nodes; every population
// starts at 0.001 (direction 6 at 0.002); the model is voxelised and
// preprocessed into reflection tables; the plane x = 0 receives the same
// vector each step; each step prints the force register, then runs
// streaming, reflections, collision and flow -- here as one fused pass
// (Lattice::simulateStep), or as the four separate calls with --staged.
// --quiet prints only the final force; --time adds one JSON line with the
// main loop's wall time (the loop is synchronised before and after) and
// the time of insertModel + preProcessModel (voxelisation and reflection
// tables, on the GPU).  --read-records FILE replaces the starting
// populations and flags with a record file (Lattice::readRecords) after
// the model is preprocessed; --write-records FILE writes the final state
// as one (Lattice::writeRecords: the 144-byte LatticePoint records of
// retrieve_data after a 20-byte header).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "FlowCriterion.h"
#include "FlowSurface.h"
#include "Lattice.h"
#include "LatticePoint.h"
#include "lbm.h"

int main(int argc, char* argv[]) {
  if (argc < 12) {
    std::fprintf(stderr,
                 "usage: %s model.obj tx ty tz bx by bz density tau spacing steps "
                 "[--staged] [--quiet] [--time] [--read-records FILE] "
                 "[--write-records FILE]\n",
                 argc > 0 ? argv[0] : "LatticeBoltzmannHip");
    return 1;
  }
  bool staged = false, quiet = false, timed = false;
  const char* read_records = nullptr;
  const char* write_records = nullptr;
  for (int a = 12; a < argc; ++a) {
    if (!std::strcmp(argv[a], "--staged")) staged = true;
    else if (!std::strcmp(argv[a], "--quiet")) quiet = true;
    else if (!std::strcmp(argv[a], "--time")) timed = true;
    else if (!std::strcmp(argv[a], "--read-records") && a + 1 < argc) read_records = argv[++a];
    else if (!std::strcmp(argv[a], "--write-records") && a + 1 < argc) write_records = argv[++a];
    else {
      std::fprintf(stderr, "unknown option %s\n", argv[a]);
      return 1;
    }
  }
  try {
    dim3 threads(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    dim3 blocks(std::atoi(argv[5]), std::atoi(argv[6]), std::atoi(argv[7]));
    FluidData fluid((float)std::atof(argv[8]), (float)std::atof(argv[9]));
    const int nx = (int)(threads.x * blocks.x), ny = (int)(threads.y * blocks.y),
              nz = (int)(threads.z * blocks.z);
    const long long n = (long long)nx * ny * nz;
    Lattice lattice(nx, ny, nz, blocks, threads, fluid, (float)std::atof(argv[10]));

    LatticePoint flowRef;
    for (int i = 0; i < 27; ++i) flowRef.particle_distribution[i] = 0.001f;
    flowRef.particle_distribution[6] = 0.002f;

    LatticePoint* data = lattice.retrieve_data();
    for (long long k = 0; k < n; ++k)
      for (int i = 0; i < 27; ++i) data[k].particle_distribution[i] = flowRef.particle_distribution[i];
    lattice.load_data(data);
    delete[] data;

    const auto s0 = std::chrono::steady_clock::now();
    lattice.insertModel(argv[1]);
    lattice.preProcessModel();
    const double setup = std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
    if (read_records) lattice.readRecords(read_records);

    FlowData inducedFlow(&FlowCriterion::ptestPlaneCriteria, flowRef);
    lattice.setFlowData(&inducedFlow);
    ReflectionData zero;
    lattice.setReflectionData(&zero);
    LatticeData latticeData(lattice.getDeviceHandle(), lattice.getDimensions());
    RunHipFunctions::run_prime_points(blocks, threads, latticeData);

    auto print_force = [&]() {
      ReflectionData* r = lattice.retrieve_reflection_data();
      std::cout << r->x << " " << r->y << " " << r->z << std::endl;
    };
    const int steps = std::atoi(argv[11]);
    lbm::check(lbm_synchronize(lattice.getDeviceHandle()), "lbm_synchronize");
    const auto t0 = std::chrono::steady_clock::now();
    for (int step = 0; step < steps; ++step) {
      if (!quiet) print_force();
      if (staged) {
        lattice.simulateStreaming();
        lattice.simulateReflections();
        lattice.simulateCollision();
        lattice.simulateFlow();
      } else {
        lattice.simulateStep();
      }
    }
    lbm::check(lbm_synchronize(lattice.getDeviceHandle()), "lbm_synchronize");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (quiet) print_force();
    if (write_records) lattice.writeRecords(write_records);
    if (timed)
      std::printf("{\"nodes\": %lld, \"steps\": %d, \"seconds\": %.6f, \"mlups\": %.1f, "
                  "\"reflection_records\": %zu, \"setup_seconds\": %.4f, \"staged\": %s, "
                  "\"printing\": %s}\n",
                  n, steps, secs, secs > 0 ? (double)n * steps / secs / 1e6 : 0.0,
                  lattice.reflectionRecordCount(), setup, staged ? "true" : "false",
                  quiet ? "false" : "true");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 2;
  }
  return 0;
}
