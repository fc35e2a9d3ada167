// Host-side simulation object of the reference (simulation/include/Lattice.h:15-106,
// simulation/source/Lattice.cpp), re-designed over the C-ABI of lbm.h.
// No GPU header is included: the lattice lives in an opaque lbm_ctx
// (D3Q27, reference mode) and every call goes through liblbm_hip.so.
//
// Differences from the reference that a caller can see:
//   - the device handle (getDeviceHandle) is an lbm_ctx*, not a pitched
//     pointer; the device layout is blocked structure-of-arrays;
//   - blocks x threads must equal the lattice extents (the reference's
//     kernels assume it, the constructor here checks it and throws
//     std::invalid_argument); the GPU launch shape is chosen internally;
//   - populations start at zero (the reference leaves them uninitialised);
//   - errors of the GPU layer throw lbm::Error instead of printing;
//   - retrieve_data()[n].reflections points into a table owned by the
//     Lattice (stable for its lifetime); d_reflections is NULL.
#pragma once

#include <array>
#include <cstdint>
#include <deque>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "FlowSurface.h"
#include "Geometry.h"
#include "LatticePoint.h"

struct lbm_ctx;
class Model;

namespace lbm {
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};
// Throws lbm::Error when rc != 0, with lbm_last_error() as message.
void check(int rc, const char* where);

// Host geometry preprocessing behind Lattice::insertModel/preProcessModel,
// usable without a GPU.  flags holds one byte per node (bit 0 internal,
// bit 1 reflected), node index z + nz*(y + ny*x).
//
// Lattice.cpp:228-260: sets bit 0 where the model contains (x h, y h, z h).
// Returns false (flags untouched) when the model is not loaded or does not
// fit in [0, N h] on every axis.
bool mark_internal(Model& model, int nx, int ny, int nz, float spacing,
                   std::vector<uint8_t>& flags);
// Lattice.cpp:323-426: reflection records of every internal node with a
// fluid neighbour; sets bit 1 on those nodes.  records/nodes are replaced.
void build_reflections(Model& model, int nx, int ny, int nz, float spacing,
                       std::vector<uint8_t>& flags,
                       std::deque<ReflectionValues>& records,
                       std::vector<int32_t>& nodes);
// Lattice.cpp:262-321
std::array<std::pair<float, int>, 27> distribute_vector(Point_3 vector);
}  // namespace lbm

// Launch-shape triple of the reference's API (CUDA's dim3 in the
// reference).  The facade only checks that blocks x threads matches the
// lattice; it never launches with it.
struct dim3 {
  unsigned x, y, z;
  dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

using LatticeHandle = lbm_ctx*;

struct LatticeData {
  LatticeHandle latticePtr;
  dim3 latticeDimensions;
  LatticeData(LatticeHandle ptr, dim3 dims) : latticePtr(ptr), latticeDimensions(dims) {}
  LatticeData() : latticePtr(nullptr), latticeDimensions(dim3()) {}
};

struct FluidData {
  float m_fluidDensity;
  float m_characteristicTimescale;
  FluidData(float fluidDensity, float characteristicTimescale)
      : m_fluidDensity(fluidDensity), m_characteristicTimescale(characteristicTimescale) {}
  FluidData() : m_fluidDensity(1.0f), m_characteristicTimescale(1.0f) {}
};

struct ReflectionData {
  float x = 0.0f;
  float y = 0.0f;
  float z = 0.0f;
};

class Lattice {
 public:
  Lattice(int x, int y, int z, dim3 blocks, dim3 threads, FluidData fluid,
          float spacing);
  ~Lattice();
  Lattice(const Lattice&) = delete;
  Lattice& operator=(const Lattice&) = delete;

  void load_data(LatticePoint* lattice_array);
  LatticePoint* retrieve_data();  // new[]; the caller deletes[]
  ReflectionData* retrieve_reflection_data();

  dim3 getDimensions() { return dim3(m_xResolution, m_yResolution, m_zResolution); }
  LatticeHandle getDeviceHandle() { return m_ctx; }
  // The reference's name for the same accessor (Lattice.h:51), so its
  // callers compile unchanged; it returns the opaque handle, not a
  // pitched pointer.
  LatticeHandle getCudaDataPointer() { return m_ctx; }

  void simulateStreaming();
  void simulateCollision();
  void simulateReflections();
  void simulateFlow();
  void simulateLattice();
  // One step of the reference main loop's simulateStreaming,
  // simulateReflections, simulateCollision, simulateFlow
  // (LatticeBoltzmannCuda.cpp:85-93) as one fused pass on the GPU; the
  // populations are bit-identical to the four calls.
  void simulateStep();

  // Files (no counterpart in the reference).  writeRecords / readRecords:
  // the retrieve_data / load_data record stream (Lattice.cpp:73-113) as a
  // file -- "LBMPTS1\0", the three extents as int32, then one 144-byte
  // LatticePoint per node in node order (reflection pointers not stored).
  // saveSnapshot / loadSnapshot: the whole resumable state
  // (lbm_snapshot_save / lbm_snapshot_load).
  void writeRecords(const std::string& path);
  void readRecords(const std::string& path);
  void saveSnapshot(const std::string& path);
  void loadSnapshot(const std::string& path);

  void insertModel(std::string filename);
  std::array<std::pair<float, int>, 27> distributeVector(Point_3 vector);
  void preProcessModel();

  void setFlowData(FlowData* flowData);
  void setReflectionData(ReflectionData* reflection);

  void setReflectionDirection(LatticePoint* point, const int index,
                              const unsigned int value);
  int getReflectionDirection(LatticePoint* point, const int index);

  // Number of reflection records built by preProcessModel.
  std::size_t reflectionRecordCount() const { return m_reflections.size(); }

 private:
  std::vector<uint8_t> downloadFlags();

  LatticeHandle m_ctx = nullptr;
  FluidData m_fluid;
  FlowData* m_flowData = nullptr;
  int m_xResolution, m_yResolution, m_zResolution;
  float m_latticeSpacing;
  std::unique_ptr<Model> m_simModel;
  dim3 m_threads, m_blocks;
  ReflectionData* m_reflectionData = nullptr;
  // stable storage: pointers handed out through LatticePoint::reflections
  std::deque<ReflectionValues> m_reflections;
  std::vector<int32_t> m_reflectionNodes;
};

// The reference's launch wrappers (simulation/include/SimCalcFuncs.cuh:13-26,
// namespace RunCudaFunctions).  Each runs the corresponding stage of the
// lattice behind `lattice`; the temporary lattice and the device-side
// FlowData / ReflectionData pointers of the reference are owned by the
// context, so those arguments are accepted and not used.
namespace RunHipFunctions {
void run_calculate_streaming(dim3 blocks, dim3 threads, LatticeData lattice,
                             LatticeData templattice);
void run_calculate_collision(dim3 blocks, dim3 threads, LatticeData lattice,
                             float timescale);
void run_calculate_reflections(dim3 blocks, dim3 threads, LatticeData lattice,
                               LatticeData templattice);
void run_calculate_reflections_data(dim3 blocks, dim3 threads,
                                    LatticeData lattice,
                                    LatticeData templattice,
                                    ReflectionData* reflection);
void run_prime_points(dim3 blocks, dim3 threads, LatticeData lattice);
void run_generate_flow(dim3 blocks, dim3 threads, LatticeData lattice,
                       FlowData* flowData);
}  // namespace RunHipFunctions

// The reference's CUDA test helper (test/CudaTestHelper.cuh:6-13), here on
// the GPU through lbm_test_coordinates.  result is a host pointer.
namespace RunHipTestFunctions {
void run_test_coordinates(dim3 threads, dim3 blocks, LatticeHandle lattice,
                          int y_resolution, bool* result);
void run_test_streaming(dim3 threads, dim3 blocks, LatticeData lattice,
                        LatticeData templattice);
void run_prime_points(dim3 threads, dim3 blocks, LatticeData lattice);
}  // namespace RunHipTestFunctions

// The reference's namespace names (SimCalcFuncs.cuh:13, CudaTestHelper.cuh:6)
// for the same functions, so the reference's callers and tests compile
// unchanged.  Nothing CUDA is behind them: they are the HIP functions above.
namespace RunCudaFunctions = RunHipFunctions;
namespace RunCudaTestFunctions = RunHipTestFunctions;
