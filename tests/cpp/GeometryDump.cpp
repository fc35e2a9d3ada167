// Writes the facade's geometry preprocessing (flags + reflection records)
// for one mesh and lattice to a binary file, for comparison with the
// oracle.  Host builders by default (no GPU); with --gpu the GPU
// voxeliser and reflection-table builder of the C-ABI.
//   GeometryDump <mesh.obj> nx ny nz spacing <out.bin> [--gpu]
// out.bin: N flag bytes, int32 nrec, nrec int32 nodes, nrec*14 uint32
// direction words, nrec*81 float weights; int32 -1 alone if the model was
// ignored.
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <vector>

#include <cstring>

#include "Lattice.h"
#include "Model.h"
#include "lbm.h"

int main(int argc, char** argv) {
  if (argc != 7 && argc != 8) {
    std::fprintf(stderr, "usage: %s mesh nx ny nz spacing out [--gpu]\n", argv[0]);
    return 2;
  }
  const bool gpu = argc == 8 && !std::strcmp(argv[7], "--gpu");
  const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
  const float h = std::strtof(argv[5], nullptr);
  Model m;
  m.importModel(argv[1]);
  std::vector<uint8_t> flags((std::size_t)nx * ny * nz, 0);
  FILE* f = std::fopen(argv[6], "wb");
  if (!f) return 1;
  if (!lbm::mark_internal(m, nx, ny, nz, h, flags)) {
    const int32_t none = -1;
    std::fwrite(&none, 4, 1, f);
    std::fclose(f);
    return 0;
  }
  std::deque<ReflectionValues> recs;
  std::vector<int32_t> nodes;
  if (!gpu) {
    lbm::build_reflections(m, nx, ny, nz, h, flags, recs, nodes);
  } else {
    lbm_config cfg{};
    cfg.nx = nx;
    cfg.ny = ny;
    cfg.nz = nz;
    cfg.lattice = LBM_D3Q27;
    cfg.storage = LBM_STORE_F32;
    cfg.mode = LBM_MODE_REFERENCE;
    cfg.tau = 1.0f;
    cfg.lid_face = -1;
    cfg.wall_rho = 1.0f;
    lbm_ctx* ctx = nullptr;
    lbm::check(lbm_create(&cfg, nullptr, &ctx), "lbm_create");
    const TriangleMesh& mesh = m.getMesh();
    std::vector<uint8_t> inside(flags.size(), 0);
    if (!m.checkError())
      lbm::check(lbm_voxelize_mesh(ctx, mesh.vertices.data(), (int32_t)mesh.number_of_vertices(),
                                   mesh.faces.data(), (int32_t)mesh.number_of_faces(), h,
                                   inside.data(), nullptr),
                 "lbm_voxelize_mesh");
    int32_t ni = 0;
    for (std::size_t n = 0; n < flags.size(); ++n) {
      flags[n] = inside[n] ? LBM_FLAG_INTERNAL : 0;
      ni += inside[n] ? 1 : 0;
    }
    lbm::check(lbm_upload_flags(ctx, flags.data()), "lbm_upload_flags");
    std::vector<int32_t> nd((std::size_t)ni);
    std::vector<uint32_t> dw(14 * (std::size_t)ni);
    std::vector<float> ww(81 * (std::size_t)ni);
    int32_t count = 0;
    lbm::check(lbm_build_reflections(ctx, mesh.vertices.data(), (int32_t)mesh.number_of_vertices(),
                                     mesh.faces.data(), (int32_t)mesh.number_of_faces(), h, ni,
                                     nd.data(), dw.data(), ww.data(), &count),
               "lbm_build_reflections");
    lbm_destroy(ctx);
    for (int32_t r = 0; r < count; ++r) {
      recs.emplace_back();
      std::memcpy(recs.back().reflection_directions, &dw[14 * (std::size_t)r], 14 * 4);
      std::memcpy(recs.back().reflection_weight, &ww[81 * (std::size_t)r], 81 * 4);
      nodes.push_back(nd[(std::size_t)r]);
      flags[(std::size_t)nd[(std::size_t)r]] |= LBM_FLAG_REFLECTED;
    }
  }
  std::fwrite(flags.data(), 1, flags.size(), f);
  const int32_t nrec = (int32_t)recs.size();
  std::fwrite(&nrec, 4, 1, f);
  std::fwrite(nodes.data(), 4, nodes.size(), f);
  for (const auto& r : recs) std::fwrite(r.reflection_directions, 4, 14, f);
  for (const auto& r : recs) std::fwrite(r.reflection_weight, 4, 81, f);
  std::fclose(f);
  return 0;
}
