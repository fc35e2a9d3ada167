// Batched stepping of many resident lattices (internal).
//
// A batch owns standard-mode member contexts (each an lbm_ctx with its own
// population buffers, link mask and class map) that share one HIP stream.
// One persistent kernel per step updates every member: a fixed grid of
// workgroups walks a device-resident tile table (tile = `block` consecutive
// nodes of one lattice) and looks the lattice up in a descriptor table.
// Because the grid size does not depend on the resident set and both
// tables live at fixed device addresses, a captured hipGraph of a chunk of
// steps stays valid while lattices are admitted and evicted between
// chunks; only a table reallocation or a new launch shape re-captures it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <vector>

#include "lbm_internal.h"

namespace lbm {

// Device descriptor of one member lattice.
struct BatchLattice {
  Grid g;
  void* f[2];  // population buffers; the current one is f[(off + parity) & 1]
  const uint32_t* mask;
  const uint32_t* cls;
  StdParams p;
  int32_t q;
  int32_t off;
  float* u_prev;           // velocity snapshot for residuals (3 x N)
  uint32_t res_first;      // first residual chunk of this lattice
};

struct BatchTile {
  uint32_t lat;    // index into the descriptor table
  uint32_t first;  // first node of the tile
};

// Residuals are summed over fixed 256-node chunks, so a lattice's residual
// does not depend on the launch shape or on the other members.
constexpr uint32_t kResidualChunk = 256;
// Per residual chunk: sum |u - u_prev|^2, sum |u|^2, max |u|^2.
constexpr uint32_t kResidualPartials = 3;

// Tiles are kept in one table per kernel class (lattice type), and each
// step launches one persistent kernel per non-empty class: each kernel
// then holds only its own lattice type's registers and code.  Lattices
// with and without a moving wall share a class (the kernel branches once
// per tile): one launch fewer per step is worth more than the wall-free
// code path (C4 FP16: 144.7 vs 149.1 us per step).
constexpr int kBatchClasses = 3;
inline int batch_class(int q) { return q == 9 ? 0 : (q == 19 ? 1 : 2); }
inline int class_q(int c) { return c == 0 ? 9 : (c == 1 ? 19 : 27); }

int launch_batch_step(int storage, int cls, bool nt, int block, int grid,
                      const BatchLattice* lats, const BatchTile* tiles,
                      const uint32_t* ntiles, int parity, hipStream_t s);
int launch_batch_residual(int storage, const BatchLattice* lats,
                          const BatchTile* chunks, uint32_t nchunks,
                          double* partials, hipStream_t s);

}  // namespace lbm

struct lbm_batch {
  int storage = LBM_STORE_F32;
  int64_t budget = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int device_cus = 0;

  double max_cell_re = 0.0;        // admission stability limit, 0 = none
  std::vector<lbm_ctx*> slots;     // nullptr = free slot
  std::vector<float*> snapshots;   // per slot, 3 x N floats (or nullptr)
  std::vector<int64_t> slot_bytes;
  int64_t bytes = 0;               // device bytes of members + tables

  // device tables (capacities in entries)
  lbm::BatchLattice* d_lats = nullptr;
  uint32_t lats_cap = 0;
  lbm::BatchTile* d_tiles[lbm::kBatchClasses] = {};
  uint32_t tiles_cap[lbm::kBatchClasses] = {};
  uint32_t* d_ntiles = nullptr;  // tiles per class (device copy)
  lbm::BatchTile* d_chunks = nullptr;  // residual chunks
  uint32_t chunks_cap = 0;
  double* d_partials = nullptr;        // 2 per residual chunk
  uint32_t partials_cap = 0;           // chunks the partials buffer holds
  // host mirror of the current tables
  std::vector<lbm::BatchLattice> h_lats;
  std::vector<int32_t> h_lat_slot;     // descriptor index -> slot
  std::vector<uint32_t> h_res_count;   // residual chunks per descriptor
  std::vector<uint64_t> h_param_version;  // member parameter versions
  uint32_t ntiles_c[lbm::kBatchClasses] = {};
  uint32_t ntiles = 0, nchunks = 0;
  bool dirty = true;

  // launch shape
  int block = 256;
  int grid = 0;  // 0: one workgroup per tile (not persistent)
  int tables_block = 0;  // block size the tile table was built for
  bool nt = false;       // non-temporal stores: members exceed the Infinity Cache

  // captured chunk
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_chunk = 0, graph_block = 0;
  bool graph_nt = false;
  const void* graph_lats = nullptr;    // descriptor table the graph reads
  const void* graph_ntiles = nullptr;  // tile counts the graph reads
  // per class: tile table and grid the captured launches use (null / 0 for
  // a class that was empty)
  void* graph_tables[lbm::kBatchClasses] = {};
  int graph_grids[lbm::kBatchClasses] = {};
};
