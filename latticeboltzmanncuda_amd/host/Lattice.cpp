// Host-side Lattice facade over the C-ABI (reference:
// simulation/source/Lattice.cpp:22-477).  The lattice is a reference-mode
// D3Q27 context of liblbm_hip.so; geometry preprocessing (insertModel,
// preProcessModel) runs on the host as in the reference and uploads flag
// bytes and reflection records.
#include "Lattice.h"

#include <cstdio>
#include <cstring>
#include <memory>

#include <algorithm>
#include <cstring>
#include <string>

#include "Constants.h"
#include "Model.h"
#include "lbm.h"

namespace lbm {
void check(int rc, const char* where) {
  if (rc != LBM_OK) throw Error(rc, std::string(where) + ": " + lbm_last_error());
}
}  // namespace lbm

namespace {

constexpr unsigned kFieldMask = 0x1Fu;

inline int64_t node_index(int x, int y, int z, int ny, int nz) {
  return (int64_t)z + (int64_t)nz * ((int64_t)y + (int64_t)ny * x);
}

inline bool in_box(int x, int y, int z, int nx, int ny, int nz) {
  return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz;
}

// 5-bit slot fields, slot s in word s/6 at bit 5*(s%6)
// (Lattice.cpp:450-477, SimCalcFuncs.cu:18-30).
inline void set_slot(int* words, int s, unsigned value) {
  const unsigned shift = 5u * (unsigned)(s % 6);
  unsigned w = (unsigned)words[s / 6];
  w = (w & ~(kFieldMask << shift)) | ((value & kFieldMask) << shift);
  words[s / 6] = (int)w;
}

inline int get_slot(const int* words, int s) {
  return (int)(((unsigned)words[s / 6] >> (5u * (unsigned)(s % 6))) & kFieldMask);
}

}  // namespace

Lattice::Lattice(int x, int y, int z, dim3 blocks, dim3 threads,
                 FluidData fluid, float spacing)
    : m_fluid(fluid),
      m_xResolution(x),
      m_yResolution(y),
      m_zResolution(z),
      m_latticeSpacing(spacing),
      m_threads(threads),
      m_blocks(blocks) {
  if ((int64_t)blocks.x * threads.x != x || (int64_t)blocks.y * threads.y != y ||
      (int64_t)blocks.z * threads.z != z)
    throw std::invalid_argument(
        "Lattice: blocks x threads must equal the lattice extents");
  lbm_config cfg{};
  cfg.nx = x;
  cfg.ny = y;
  cfg.nz = z;
  cfg.lattice = LBM_D3Q27;
  cfg.storage = LBM_STORE_F32;
  cfg.mode = LBM_MODE_REFERENCE;
  cfg.tau = fluid.m_characteristicTimescale;
  cfg.lid_face = -1;
  cfg.wall_rho = 1.0f;
  lbm::check(lbm_create(&cfg, nullptr, &m_ctx), "lbm_create");
}

Lattice::~Lattice() { lbm_destroy(m_ctx); }

void Lattice::load_data(LatticePoint* lattice_array) {
  const int64_t n = (int64_t)m_xResolution * m_yResolution * m_zResolution;
  lbm::check(lbm_upload_points(m_ctx, reinterpret_cast<const lbm_lattice_point*>(lattice_array), n),
             "lbm_upload_points");
}

LatticePoint* Lattice::retrieve_data() {
  const int64_t n = (int64_t)m_xResolution * m_yResolution * m_zResolution;
  LatticePoint* out = new LatticePoint[n];
  const int rc = lbm_download_points(m_ctx, reinterpret_cast<lbm_lattice_point*>(out), n);
  if (rc != LBM_OK) {
    delete[] out;
    lbm::check(rc, "lbm_download_points");
  }
  for (std::size_t r = 0; r < m_reflectionNodes.size(); ++r)
    out[m_reflectionNodes[r]].reflections = &m_reflections[r];
  return out;
}

ReflectionData* Lattice::retrieve_reflection_data() {
  if (!m_reflectionData) return nullptr;
  float f[3];
  lbm::check(lbm_get_force(m_ctx, f), "lbm_get_force");
  m_reflectionData->x = f[0];
  m_reflectionData->y = f[1];
  m_reflectionData->z = f[2];
  return m_reflectionData;
}

void Lattice::simulateStreaming() { lbm::check(lbm_stream(m_ctx), "lbm_stream"); }

void Lattice::simulateCollision() {
  lbm::check(lbm_collide(m_ctx, m_fluid.m_characteristicTimescale), "lbm_collide");
}

void Lattice::simulateReflections() {
  lbm::check(lbm_reflect(m_ctx, m_reflectionData != nullptr ? 1 : 0), "lbm_reflect");
}

void Lattice::simulateFlow() {
  if (m_flowData) lbm::check(lbm_flow(m_ctx), "lbm_flow");
}

void Lattice::simulateLattice() { lbm::check(lbm_step(m_ctx, 1), "lbm_step"); }

void Lattice::writeRecords(const std::string& path) {
  const int64_t n = (int64_t)m_xResolution * m_yResolution * m_zResolution;
  std::unique_ptr<LatticePoint[]> pts(retrieve_data());
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw lbm::Error(LBM_ERR_INVALID, "cannot open " + path);
  const char magic[8] = {'L', 'B', 'M', 'P', 'T', 'S', '1', '\0'};
  const int32_t dims[3] = {m_xResolution, m_yResolution, m_zResolution};
  bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(dims, 4, 3, f) == 3;
  for (int64_t k = 0; ok && k < n; ++k) {
    LatticePoint p = pts[k];
    p.reflections = nullptr;  // a host address: meaningless in a file
    ok = std::fwrite(&p, sizeof(LatticePoint), 1, f) == 1;
  }
  if (std::fclose(f) != 0 || !ok) throw lbm::Error(LBM_ERR_INVALID, "write failed: " + path);
}

void Lattice::readRecords(const std::string& path) {
  const int64_t n = (int64_t)m_xResolution * m_yResolution * m_zResolution;
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw lbm::Error(LBM_ERR_INVALID, "cannot open " + path);
  char magic[8];
  int32_t dims[3];
  std::unique_ptr<LatticePoint[]> pts(new LatticePoint[(std::size_t)n]());
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "LBMPTS1", 8) == 0 &&
            std::fread(dims, 4, 3, f) == 3 && dims[0] == m_xResolution &&
            dims[1] == m_yResolution && dims[2] == m_zResolution &&
            std::fread(pts.get(), sizeof(LatticePoint), (std::size_t)n, f) == (std::size_t)n;
  std::fclose(f);
  if (!ok) throw lbm::Error(LBM_ERR_INVALID, "not a record file of this lattice: " + path);
  load_data(pts.get());
}

void Lattice::saveSnapshot(const std::string& path) {
  lbm::check(lbm_snapshot_save(m_ctx, path.c_str()), "lbm_snapshot_save");
}

void Lattice::loadSnapshot(const std::string& path) {
  lbm::check(lbm_snapshot_load(m_ctx, path.c_str()), "lbm_snapshot_load");
  lbm_config cfg{};
  lbm::check(lbm_get_config(m_ctx, &cfg), "lbm_get_config");
  m_fluid.m_characteristicTimescale = cfg.tau;
}

void Lattice::simulateStep() {
  lbm::check(lbm_step_reference(m_ctx, 1, m_reflectionData != nullptr ? 1 : 0),
             "lbm_step_reference");
}

std::vector<uint8_t> Lattice::downloadFlags() {
  std::vector<uint8_t> flags((std::size_t)m_xResolution * m_yResolution * m_zResolution);
  lbm::check(lbm_get_flags(m_ctx, flags.data()), "lbm_get_flags");
  return flags;
}

namespace lbm {

bool mark_internal(Model& model, int nx, int ny, int nz, float spacing,
                   std::vector<uint8_t>& flags) {
  if (!model.isLoaded()) return false;
  const MeshBounds b = model.bounding_box();
  if (b.xmin() < 0.0f || b.ymin() < 0.0f || b.zmin() < 0.0f ||
      b.xmax() > nx * spacing || b.ymax() > ny * spacing ||
      b.zmax() > nz * spacing)
    return false;
  flags.resize((std::size_t)nx * ny * nz, 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        const int64_t n = node_index(x, y, z, ny, nz);
        const bool inside = model.isPointInsideModel(
            Point_3((float)x * spacing, (float)y * spacing, (float)z * spacing));
        flags[n] = (uint8_t)((flags[n] & ~LBM_FLAG_INTERNAL) |
                             (inside ? LBM_FLAG_INTERNAL : 0u));
      }
  return true;
}

// score_j = 2 - |v - e_j/|e_j||, all 27 directions (rest included),
// highest score first, equal scores by higher index first (the order of
// std::priority_queue<std::pair<float, int>>).
std::array<std::pair<float, int>, 27> distribute_vector(Point_3 vector) {
  std::array<std::pair<float, int>, 27> out;
  for (int j = 0; j < 27; ++j) {
    const int* e = CONSTANTS::DIRECTIONS[j];
    const int len2 = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
    const double inv = len2 == 0 ? 0.0 : (len2 == 1 ? 1.0 : (len2 == 2 ? 1.0 / 1.4142135623730951 : 1.0 / 1.7320508075688772));
    float unit[3];
    for (int a = 0; a < 3; ++a) unit[a] = e[a] == 0 ? 0.0f : (float)(e[a] * inv);
    const float dx = vector.x - unit[0], dy = vector.y - unit[1], dz = vector.z - unit[2];
    out[j] = {(float)(2.0 - (double)std::sqrt((dx * dx) + (dy * dy) + (dz * dz))), j};
  }
  std::sort(out.begin(), out.end(),
            [](const std::pair<float, int>& a, const std::pair<float, int>& b) {
              return a > b;
            });
  return out;
}

// For every internal node x and direction k with x + e_k inside the box and
// not internal: reflect along the ray from (x + e_k) h to x h, rank the
// directions, keep the first three whose target x + e_d is inside the box
// and not internal (the reference reads outside its host array for targets
// beyond the box; they are not eligible here), store them in slots
// INVERSE[k]*3 + rank and normalise the three weights.
void build_reflections(Model& model, int nx, int ny, int nz, float h,
                       std::vector<uint8_t>& flags,
                       std::deque<ReflectionValues>& records,
                       std::vector<int32_t>& nodes) {
  auto internal = [&](int x, int y, int z) {
    return (flags[node_index(x, y, z, ny, nz)] & LBM_FLAG_INTERNAL) != 0;
  };
  records.clear();
  nodes.clear();
  for (int x = 0; x < nx; ++x)
    for (int y = 0; y < ny; ++y)
      for (int z = 0; z < nz; ++z) {
        if (!internal(x, y, z)) continue;
        ReflectionValues* rec = nullptr;
        for (int k = 0; k < 27; ++k) {
          const int* ek = CONSTANTS::DIRECTIONS[k];
          const int sx = x + ek[0], sy = y + ek[1], sz = z + ek[2];
          if (!in_box(sx, sy, sz, nx, ny, nz) || internal(sx, sy, sz)) continue;
          const Point_3 v = model.reflectionVector(
              Point_3(h * (float)sx, h * (float)sy, h * (float)sz),
              Point_3(h * (float)x, h * (float)y, h * (float)z));
          const std::array<std::pair<float, int>, 27> ranked = distribute_vector(v);
          const int base = CONSTANTS::INVERSE_DIRECTIONS[k] * 3;
          int counter = 0;
          float total = 0.0f;
          for (int i = 0; i < 27 && counter < 3; ++i) {
            const int d = ranked[i].second;
            const int* ed = CONSTANTS::DIRECTIONS[d];
            const int tx = x + ed[0], ty = y + ed[1], tz = z + ed[2];
            if (!in_box(tx, ty, tz, nx, ny, nz) || internal(tx, ty, tz)) continue;
            if (!rec) {
              records.emplace_back();
              rec = &records.back();
              std::memset(rec, 0, sizeof(ReflectionValues));
              nodes.push_back((int32_t)node_index(x, y, z, ny, nz));
            }
            set_slot(rec->reflection_directions, base + counter, (unsigned)d);
            rec->reflection_weight[base + counter] = ranked[i].first;
            total += ranked[i].first;
            ++counter;
          }
          if (total != 0.0f && rec)
            for (int i = 0; i < 3; ++i)
              if (get_slot(rec->reflection_directions, base + i) != 0)
                rec->reflection_weight[base + i] /= total;
        }
        if (rec) flags[node_index(x, y, z, ny, nz)] |= LBM_FLAG_REFLECTED;
      }
}

}  // namespace lbm

// Lattice.cpp:228-260.  The inside test of every node runs on the GPU
// (lbm_voxelize_mesh) with the decisions of Model::isPointInsideModel; the
// host loop lbm::mark_internal remains for checks.
void Lattice::insertModel(std::string filename) {
  m_simModel.reset(new Model());
  m_simModel->importModel(filename);
  if (!m_simModel->isLoaded()) return;
  const MeshBounds b = m_simModel->bounding_box();
  if (b.xmin() < 0.0f || b.ymin() < 0.0f || b.zmin() < 0.0f ||
      b.xmax() > m_xResolution * m_latticeSpacing ||
      b.ymax() > m_yResolution * m_latticeSpacing ||
      b.zmax() > m_zResolution * m_latticeSpacing)
    return;
  std::vector<uint8_t> flags = downloadFlags();
  std::vector<uint8_t> inside(flags.size(), 0);
  if (!m_simModel->checkError()) {  // an invalid mesh has no inside
    const TriangleMesh& mesh = m_simModel->getMesh();
    lbm::check(lbm_voxelize_mesh(m_ctx, mesh.vertices.data(), (int32_t)mesh.number_of_vertices(),
                                 mesh.faces.data(), (int32_t)mesh.number_of_faces(),
                                 m_latticeSpacing, inside.data(), nullptr),
               "lbm_voxelize_mesh");
  }
  for (std::size_t n = 0; n < flags.size(); ++n)
    flags[n] = (uint8_t)((flags[n] & ~LBM_FLAG_INTERNAL) | (inside[n] ? LBM_FLAG_INTERNAL : 0u));
  lbm::check(lbm_upload_flags(m_ctx, flags.data()), "lbm_upload_flags");
}

std::array<std::pair<float, int>, 27> Lattice::distributeVector(Point_3 vector) {
  return lbm::distribute_vector(vector);
}

// Lattice.cpp:323-426.  Records are rebuilt from scratch on every call;
// the tables are built on the GPU (lbm_build_reflections) with the values
// of the host builder lbm::build_reflections, which remains for checks.
void Lattice::preProcessModel() {
  if (!m_simModel) return;
  std::vector<uint8_t> flags = downloadFlags();
  int32_t ninternal = 0;
  for (uint8_t f : flags) ninternal += (f & LBM_FLAG_INTERNAL) ? 1 : 0;
  const TriangleMesh& mesh = m_simModel->getMesh();
  std::vector<int32_t> nodes((std::size_t)ninternal);
  std::vector<uint32_t> dirs(14 * (std::size_t)ninternal);
  std::vector<float> weights(81 * (std::size_t)ninternal);
  int32_t count = 0;
  lbm::check(lbm_build_reflections(m_ctx, mesh.vertices.data(), (int32_t)mesh.number_of_vertices(),
                                   mesh.faces.data(), (int32_t)mesh.number_of_faces(),
                                   m_latticeSpacing, ninternal, nodes.data(), dirs.data(),
                                   weights.data(), &count),
             "lbm_build_reflections");
  const std::size_t nrec = (std::size_t)count;
  nodes.resize(nrec);
  dirs.resize(14 * nrec);
  weights.resize(81 * nrec);
  m_reflections.clear();
  m_reflectionNodes = nodes;
  for (std::size_t r = 0; r < nrec; ++r) {
    m_reflections.emplace_back();
    ReflectionValues& rec = m_reflections.back();
    std::memcpy(rec.reflection_directions, &dirs[14 * r], 14 * sizeof(uint32_t));
    std::memcpy(rec.reflection_weight, &weights[81 * r], 81 * sizeof(float));
    flags[(std::size_t)nodes[r]] |= LBM_FLAG_REFLECTED;
  }
  lbm::check(lbm_upload_flags(m_ctx, flags.data()), "lbm_upload_flags");
  lbm::check(lbm_upload_reflections(m_ctx, (int32_t)nrec, m_reflectionNodes.data(),
                                    dirs.data(), weights.data()),
             "lbm_upload_reflections");
}

void Lattice::setFlowData(FlowData* flowData) {
  m_flowData = flowData;
  if (!flowData) {
    lbm::check(lbm_set_flow(m_ctx, LBM_FLOW_NONE, nullptr, nullptr, nullptr), "lbm_set_flow");
    return;
  }
  lbm::check(lbm_set_flow(m_ctx, (int)flowData->pointCriterion.kind,
                          flowData->pointCriterion.params,
                          flowData->inducedFlow.particle_distribution, nullptr),
             "lbm_set_flow");
}

void Lattice::setReflectionData(ReflectionData* reflection) {
  m_reflectionData = reflection;
  if (reflection) {
    const float f[3] = {reflection->x, reflection->y, reflection->z};
    lbm::check(lbm_set_force(m_ctx, f), "lbm_set_force");
  }
}

void Lattice::setReflectionDirection(LatticePoint* point, const int index,
                                     const unsigned int value) {
  set_slot(point->reflections->reflection_directions, index, value);
}

int Lattice::getReflectionDirection(LatticePoint* point, const int index) {
  return get_slot(point->reflections->reflection_directions, index);
}

// ---------------------------------------------------------------- wrappers
namespace RunHipFunctions {
void run_calculate_streaming(dim3, dim3, LatticeData lattice, LatticeData) {
  lbm::check(lbm_stream(lattice.latticePtr), "lbm_stream");
}
void run_calculate_collision(dim3, dim3, LatticeData lattice, float timescale) {
  lbm::check(lbm_collide(lattice.latticePtr, timescale), "lbm_collide");
}
void run_calculate_reflections(dim3, dim3, LatticeData lattice, LatticeData) {
  lbm::check(lbm_reflect(lattice.latticePtr, 0), "lbm_reflect");
}
void run_calculate_reflections_data(dim3, dim3, LatticeData lattice, LatticeData,
                                    ReflectionData*) {
  lbm::check(lbm_reflect(lattice.latticePtr, 1), "lbm_reflect");
}
void run_prime_points(dim3, dim3, LatticeData lattice) {
  lbm::check(lbm_prime_points(lattice.latticePtr), "lbm_prime_points");
}
void run_generate_flow(dim3, dim3, LatticeData lattice, FlowData*) {
  lbm::check(lbm_flow(lattice.latticePtr), "lbm_flow");
}
}  // namespace RunHipFunctions

namespace RunHipTestFunctions {
void run_test_coordinates(dim3, dim3, LatticeHandle lattice, int, bool* result) {
  int ok = 0;
  lbm::check(lbm_test_coordinates(lattice, &ok), "lbm_test_coordinates");
  if (!ok) *result = false;
}
void run_test_streaming(dim3, dim3, LatticeData lattice, LatticeData) {
  lbm::check(lbm_stream(lattice.latticePtr), "lbm_stream");
}
void run_prime_points(dim3, dim3, LatticeData lattice) {
  lbm::check(lbm_prime_points(lattice.latticePtr), "lbm_prime_points");
}
}  // namespace RunHipTestFunctions
