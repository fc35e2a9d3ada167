// Writes the facade's host geometry preprocessing (flags + reflection
// records) for one mesh and lattice to a binary file, for comparison with
// the oracle.  No GPU is used.
//   GeometryDump <mesh.obj> nx ny nz spacing <out.bin>
// out.bin: N flag bytes, int32 nrec, nrec int32 nodes, nrec*14 uint32
// direction words, nrec*81 float weights; int32 -1 alone if the model was
// ignored.
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <vector>

#include "Lattice.h"
#include "Model.h"

int main(int argc, char** argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s mesh nx ny nz spacing out\n", argv[0]);
    return 2;
  }
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
  lbm::build_reflections(m, nx, ny, nz, h, flags, recs, nodes);
  std::fwrite(flags.data(), 1, flags.size(), f);
  const int32_t nrec = (int32_t)recs.size();
  std::fwrite(&nrec, 4, 1, f);
  std::fwrite(nodes.data(), 4, nodes.size(), f);
  for (const auto& r : recs) std::fwrite(r.reflection_directions, 4, 14, f);
  for (const auto& r : recs) std::fwrite(r.reflection_weight, 4, 81, f);
  std::fclose(f);
  return 0;
}
