// C-ABI of lattice batches: admission under an HBM budget, one persistent
// kernel per step over every resident lattice, hipGraph-captured chunks of
// steps, an autotuner for the launch shape, and per-lattice residuals.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "lbm_batch.h"
#include "lbm_layout.h"
#include "lbm_store.h"

using namespace lbm;

namespace {

int check_batch(const lbm_batch* b) {
  if (!b) return fail(LBM_ERR_INVALID, "null batch");
  return LBM_OK;
}

int64_t member_bytes(const lbm_config& cfg, int track) {
  const int64_t n = (int64_t)cfg.nx * cfg.ny * cfg.nz;
  return estimate_ctx_bytes(cfg) + (track ? 12 * n : 0);
}

void drop_graph(lbm_batch* b) {
  // replays of the executable graph may still be running on the stream
  if (b->graph_exec && b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->graph_exec = nullptr;
  b->graph = nullptr;
  b->graph_chunk = 0;
}

template <class T>
int grow(lbm_batch* b, T** p, uint32_t* cap, uint32_t need, const char* what) {
  if (need <= *cap && *p) return LBM_OK;
  uint32_t c = std::max<uint32_t>(need, 64);
  c = c + c / 2;
  // captured launches point at the old table: the graph goes first (after
  // its replays finish), so no instantiated graph ever references freed
  // memory, whether or not the new allocation succeeds
  drop_graph(b);
  if (*p) {
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(*p);
    b->bytes -= (int64_t)(*cap) * (int64_t)sizeof(T);
  }
  *p = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)c);
  if (e != hipSuccess) {
    *p = nullptr;
    return hip_fail(e, what);
  }
  b->bytes += (int64_t)c * (int64_t)sizeof(T);
  *cap = c;
  return LBM_OK;
}

bool tables_stale(const lbm_batch* b) {
  if (b->dirty || b->tables_block != b->block) return true;
  for (size_t k = 0; k < b->h_lats.size(); ++k) {
    const lbm_ctx* c = b->slots[b->h_lat_slot[k]];
    if (b->h_lats[k].off != c->cur || b->h_param_version[k] != c->params_version) return true;
  }
  return false;
}

// Rebuilds the descriptor, tile and residual-chunk tables of the resident
// set (in slot order) and copies them to the device on the batch stream.
int sync_tables(lbm_batch* b) {
  if (!tables_stale(b)) return LBM_OK;
  std::vector<BatchTile> tiles[kBatchClasses], chunks;
  uint64_t pop_bytes = 0;
  b->h_lats.clear();
  b->h_lat_slot.clear();
  b->h_res_count.clear();
  b->h_param_version.clear();
  for (size_t s = 0; s < b->slots.size(); ++s) {
    lbm_ctx* c = b->slots[s];
    if (!c) continue;
    BatchLattice d{};
    d.g = c->grid;
    d.f[0] = c->f[0];
    d.f[1] = c->f[1];
    d.mask = c->mask;
    d.cls = c->cls;
    d.p = make_std_params(c->cfg);
    d.q = c->q;
    d.off = c->cur;
    d.u_prev = b->snapshots[s];
    d.res_first = (uint32_t)chunks.size();
    pop_bytes += 2 * population_elems(c->q, c->grid.n) * storage_bytes(b->storage);
    const uint32_t idx = (uint32_t)b->h_lats.size();
    const int cls = batch_class(c->q);
    for (uint32_t first = 0; first < c->grid.n; first += (uint32_t)b->block)
      tiles[cls].push_back({idx, first});
    uint32_t nres = 0;
    for (uint32_t first = 0; first < c->grid.n; first += kResidualChunk, ++nres)
      chunks.push_back({idx, first});
    b->h_lats.push_back(d);
    b->h_lat_slot.push_back((int32_t)s);
    b->h_res_count.push_back(nres);
    b->h_param_version.push_back(c->params_version);
  }
  int rc;
  if ((rc = grow(b, &b->d_lats, &b->lats_cap, (uint32_t)b->h_lats.size(), "hipMalloc(batch lattices)")) ||
      (rc = grow(b, &b->d_chunks, &b->chunks_cap, (uint32_t)chunks.size(), "hipMalloc(batch chunks)")))
    return rc;
  for (int k = 0; k < kBatchClasses; ++k)
    if (!tiles[k].empty() &&
        (rc = grow(b, &b->d_tiles[k], &b->tiles_cap[k], (uint32_t)tiles[k].size(), "hipMalloc(batch tiles)")))
      return rc;
  if (!b->d_partials || b->partials_cap < b->chunks_cap) {
    if (b->d_partials) {
      (void)hipStreamSynchronize(b->stream);
      (void)hipFree(b->d_partials);
      b->bytes -= (int64_t)sizeof(double) * kResidualPartials * b->partials_cap;
    }
    b->partials_cap = 0;
    hipError_t e =
        hipMalloc((void**)&b->d_partials, sizeof(double) * kResidualPartials * b->chunks_cap);
    if (e != hipSuccess) {
      b->d_partials = nullptr;
      return hip_fail(e, "hipMalloc(batch partials)");
    }
    b->partials_cap = b->chunks_cap;
    b->bytes += (int64_t)sizeof(double) * kResidualPartials * b->partials_cap;
  }
  if (!b->d_ntiles) {
    hipError_t e = hipMalloc((void**)&b->d_ntiles, sizeof(uint32_t) * kBatchClasses);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(batch tile counts)");
    b->bytes += sizeof(uint32_t) * kBatchClasses;
  }
  b->nt = streaming_stores(pop_bytes);
  b->ntiles = 0;
  for (int k = 0; k < kBatchClasses; ++k) {
    b->ntiles_c[k] = (uint32_t)tiles[k].size();
    b->ntiles += b->ntiles_c[k];
  }
  b->nchunks = (uint32_t)chunks.size();
  if (!b->h_lats.empty())
    LBM_HIP(hipMemcpyAsync(b->d_lats, b->h_lats.data(), sizeof(BatchLattice) * b->h_lats.size(),
                           hipMemcpyHostToDevice, b->stream));
  for (int k = 0; k < kBatchClasses; ++k)
    if (!tiles[k].empty())
      LBM_HIP(hipMemcpyAsync(b->d_tiles[k], tiles[k].data(), sizeof(BatchTile) * tiles[k].size(),
                             hipMemcpyHostToDevice, b->stream));
  if (!chunks.empty())
    LBM_HIP(hipMemcpyAsync(b->d_chunks, chunks.data(), sizeof(BatchTile) * chunks.size(),
                           hipMemcpyHostToDevice, b->stream));
  LBM_HIP(hipMemcpyAsync(b->d_ntiles, b->ntiles_c, sizeof(uint32_t) * kBatchClasses,
                         hipMemcpyHostToDevice, b->stream));
  // the host vectors die here: wait for the copies
  LBM_HIP(hipStreamSynchronize(b->stream));
  b->dirty = false;
  b->tables_block = b->block;
  return LBM_OK;
}

// Workgroups launched for class k: the configured persistent grid (fixed,
// so a captured graph survives admissions and evictions; surplus
// workgroups find no tile and exit), or one workgroup per tile when grid is
// 0; 0 for an empty class, which is not launched.
int class_grid(const lbm_batch* b, int k) {
  const uint32_t t = b->ntiles_c[k];
  if (t == 0) return 0;
  return b->grid > 0 ? b->grid : (int)t;
}

int launch_steps(lbm_batch* b, int n) {
  for (int s = 0; s < n; ++s)
    for (int k = 0; k < kBatchClasses; ++k) {
      const int grid = class_grid(b, k);
      if (grid == 0) continue;
      int rc = launch_batch_step(device_storage(LBM_MODE_STANDARD, b->storage), k, b->nt,
                                 b->block, grid, b->d_lats, b->d_tiles[k],
                                 b->d_ntiles + k, s & 1, b->stream);
      if (rc) return rc;
    }
  return LBM_OK;
}

void advance_members(lbm_batch* b, int n) {
  for (lbm_ctx* c : b->slots)
    if (c) {
      c->cur ^= (n & 1);
      c->steps_done += (uint64_t)n;
    }
}

int capture_chunk(lbm_batch* b, int chunk) {
  drop_graph(b);
  hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
  int rc = launch_steps(b, chunk);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  b->graph = g;
  e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(b);
    return hip_fail(e, "hipGraphInstantiate");
  }
  b->graph_chunk = chunk;
  b->graph_block = b->block;
  b->graph_nt = b->nt;
  b->graph_lats = b->d_lats;
  b->graph_ntiles = b->d_ntiles;
  for (int k = 0; k < kBatchClasses; ++k) {
    b->graph_grids[k] = class_grid(b, k);
    b->graph_tables[k] = b->graph_grids[k] ? (void*)b->d_tiles[k] : nullptr;
  }
  return LBM_OK;
}

// The captured launches stay valid while every class launches the same
// grid from the same tables (the tables' contents may change).  A table
// is only ever freed after the graph is dropped (grow), so an address
// that matches here is the live table the graph was captured with, never
// a freed one that happens to be reused.
bool graph_current(const lbm_batch* b, int chunk) {
  if (!b->graph_exec || b->graph_chunk != chunk || b->graph_block != b->block ||
      b->graph_nt != b->nt || b->graph_lats != b->d_lats || b->graph_ntiles != b->d_ntiles)
    return false;
  for (int k = 0; k < kBatchClasses; ++k) {
    const int grid = class_grid(b, k);
    if (b->graph_grids[k] != grid) return false;
    if (grid && b->graph_tables[k] != (void*)b->d_tiles[k]) return false;
  }
  return true;
}

int ensure_graph(lbm_batch* b, int chunk) {
  if (graph_current(b, chunk)) return LBM_OK;
  return capture_chunk(b, chunk);
}

// Launch-shape cache of the autotuner: (storage, lattice-mix bits, log2
// nodes) -> (block, grid).
std::mutex g_tune_mutex;
std::map<std::tuple<int, int, int>, std::pair<int, int>> g_tune_cache;

std::tuple<int, int, int> tune_key(const lbm_batch* b) {
  int mix = 0;
  int64_t nodes = 0;
  for (const lbm_ctx* c : b->slots)
    if (c) {
      mix |= c->q == 9 ? 1 : (c->q == 19 ? 2 : 4);
      nodes += c->grid.n;
    }
  int lg = 0;
  while ((1ll << (lg + 1)) <= nodes) ++lg;
  return {b->storage, mix, lg};
}

}  // namespace

// Entry points of include/lbm.h (exported by lbm_exports.hip).
namespace lbm {
namespace api {

int lbm_batch_create(int storage, int64_t budget_bytes, void* stream,
                     lbm_batch** out) {
  if (!out) return fail(LBM_ERR_INVALID, "null output handle");
  *out = nullptr;
  if (storage_bytes(storage) == 0) return fail(LBM_ERR_INVALID, "unknown storage type");
  if (budget_bytes <= 0) return fail(LBM_ERR_INVALID, "budget must be positive");
  lbm_batch* b = new (std::nothrow) lbm_batch();
  if (!b) return fail(LBM_ERR_NOMEM, "host allocation of batch");
  b->storage = storage;
  b->budget = budget_bytes;
  hipError_t e;
  int dev = 0;
  hipDeviceProp_t prop;
  if ((e = hipGetDevice(&dev)) != hipSuccess ||
      (e = hipGetDeviceProperties(&prop, dev)) != hipSuccess) {
    delete b;
    return hip_fail(e, "hipGetDeviceProperties");
  }
  b->device_cus = prop.multiProcessorCount;
  b->grid = 8 * b->device_cus;
  if (stream) {
    b->stream = (hipStream_t)stream;
  } else {
    if ((e = hipStreamCreate(&b->stream)) != hipSuccess) {
      delete b;
      return hip_fail(e, "hipStreamCreate");
    }
    b->own_stream = true;
  }
  if ((e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess) {
    lbm_batch_destroy(b);
    return hip_fail(e, "hipEventCreate");
  }
  *out = b;
  return LBM_OK;
}

int lbm_batch_destroy(lbm_batch* b) {
  if (!b) return LBM_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (size_t s = 0; s < b->slots.size(); ++s) {
    if (b->slots[s]) lbm_destroy(b->slots[s]);
    if (b->snapshots[s]) (void)hipFree(b->snapshots[s]);
  }
  for (void* p : {(void*)b->d_lats, (void*)b->d_ntiles, (void*)b->d_chunks, (void*)b->d_partials})
    if (p) (void)hipFree(p);
  for (int k = 0; k < kBatchClasses; ++k)
    if (b->d_tiles[k]) (void)hipFree(b->d_tiles[k]);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return LBM_OK;
}

int64_t lbm_batch_estimate_bytes(const lbm_config* cfg, int track_residual) {
  if (!cfg) return -1;
  return member_bytes(*cfg, track_residual);
}

int lbm_batch_add(lbm_batch* b, const lbm_config* cfg, const uint8_t* solid,
                  int track_residual, int32_t* slot) {
  if (int rc = check_batch(b)) return rc;
  if (!cfg || !slot) return fail(LBM_ERR_INVALID, "null argument");
  if (cfg->mode != LBM_MODE_STANDARD)
    return fail(LBM_ERR_UNSUPPORTED, "batches hold standard-mode lattices");
  if (cfg->storage != b->storage)
    return fail(LBM_ERR_INVALID, "lattice storage differs from the batch storage");
  if (int rc = check_config(cfg)) return rc;
  if (b->max_cell_re > 0.0) {
    const double re = cell_reynolds(*cfg);
    if (!(re <= b->max_cell_re))
      return fail(LBM_ERR_UNSTABLE, "cell Reynolds number " + std::to_string(re) +
                                        " above the batch's stability limit " +
                                        std::to_string(b->max_cell_re));
  }
  const int64_t need = member_bytes(*cfg, track_residual);
  if (b->bytes + need > b->budget)
    return fail(LBM_ERR_BUDGET, "HBM budget of the batch would be exceeded");
  lbm_ctx* c = nullptr;
  int rc = lbm_create(cfg, b->stream, &c);
  if (rc) return rc;
  if (solid && (rc = lbm_upload_solid(c, solid))) {
    lbm_destroy(c);
    return rc;
  }
  if ((rc = lbm_init_equilibrium(c, nullptr, nullptr))) {
    lbm_destroy(c);
    return rc;
  }
  float* snap = nullptr;
  if (track_residual) {
    const size_t sb = sizeof(float) * 3 * (size_t)c->grid.n;
    hipError_t e = hipMalloc((void**)&snap, sb);
    if (e == hipSuccess) e = hipMemsetAsync(snap, 0, sb, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
      if (snap) (void)hipFree(snap);
      lbm_destroy(c);
      return hip_fail(e, "hipMalloc(residual snapshot)");
    }
  }
  int32_t s = 0;
  while (s < (int32_t)b->slots.size() && b->slots[s]) ++s;
  if (s == (int32_t)b->slots.size()) {
    b->slots.push_back(nullptr);
    b->snapshots.push_back(nullptr);
    b->slot_bytes.push_back(0);
  }
  b->slots[s] = c;
  b->snapshots[s] = snap;
  b->slot_bytes[s] = need;
  b->bytes += need;
  b->dirty = true;
  *slot = s;
  return LBM_OK;
}

int lbm_batch_set_stability_limit(lbm_batch* b, double max_cell_re) {
  if (int rc = check_batch(b)) return rc;
  if (!(max_cell_re >= 0.0)) return fail(LBM_ERR_INVALID, "stability limit must be >= 0");
  b->max_cell_re = max_cell_re;
  return LBM_OK;
}

int lbm_batch_remove(lbm_batch* b, int32_t slot) {
  if (int rc = check_batch(b)) return rc;
  if (slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot])
    return fail(LBM_ERR_NOTFOUND, "no lattice in that slot");
  (void)hipStreamSynchronize(b->stream);
  lbm_destroy(b->slots[slot]);
  if (b->snapshots[slot]) (void)hipFree(b->snapshots[slot]);
  b->slots[slot] = nullptr;
  b->snapshots[slot] = nullptr;
  b->bytes -= b->slot_bytes[slot];
  b->slot_bytes[slot] = 0;
  b->dirty = true;
  return LBM_OK;
}

int lbm_batch_member(lbm_batch* b, int32_t slot, lbm_ctx** ctx) {
  if (int rc = check_batch(b)) return rc;
  if (!ctx) return fail(LBM_ERR_INVALID, "null output");
  if (slot < 0 || slot >= (int32_t)b->slots.size() || !b->slots[slot])
    return fail(LBM_ERR_NOTFOUND, "no lattice in that slot");
  *ctx = b->slots[slot];
  return LBM_OK;
}

int lbm_batch_info(const lbm_batch* b, int32_t* lattices, int64_t* nodes,
                   int64_t* device_bytes) {
  if (int rc = check_batch(b)) return rc;
  int32_t l = 0;
  int64_t n = 0;
  for (const lbm_ctx* c : b->slots)
    if (c) {
      ++l;
      n += c->grid.n;
    }
  if (lattices) *lattices = l;
  if (nodes) *nodes = n;
  if (device_bytes) *device_bytes = b->bytes;
  return LBM_OK;
}

int lbm_batch_set_launch(lbm_batch* b, int block, int grid) {
  if (int rc = check_batch(b)) return rc;
  if (block != 64 && block != 128 && block != 256 && block != 512)
    return fail(LBM_ERR_INVALID, "block must be 64, 128, 256 or 512");
  if (grid < 0) return fail(LBM_ERR_INVALID, "grid must be >= 0");
  b->block = block;
  b->grid = grid;
  return LBM_OK;
}

int lbm_batch_get_launch(const lbm_batch* b, int* block, int* grid) {
  if (int rc = check_batch(b)) return rc;
  if (block) *block = b->block;
  if (grid) *grid = b->grid;
  return LBM_OK;
}

int lbm_batch_step(lbm_batch* b, int n) {
  if (int rc = check_batch(b)) return rc;
  if (n < 0) return fail(LBM_ERR_INVALID, "negative step count");
  TraceRange tr("lbm_batch_step");
  int rc = sync_tables(b);
  if (rc || n == 0 || b->ntiles == 0) return rc;
  if ((rc = launch_steps(b, n))) return rc;
  advance_members(b, n);
  return LBM_OK;
}

int lbm_batch_step_graph(lbm_batch* b, int n, int chunk) {
  if (int rc = check_batch(b)) return rc;
  if (n < 0 || chunk < 2 || (chunk & 1))
    return fail(LBM_ERR_INVALID, "chunk must be an even step count >= 2");
  TraceRange tr("lbm_batch_step_graph");
  int rc = sync_tables(b);
  if (rc || n == 0 || b->ntiles == 0) return rc;
  const int full = n / chunk;
  if (full > 0) {
    if ((rc = ensure_graph(b, chunk))) return rc;
    for (int k = 0; k < full; ++k) LBM_HIP(hipGraphLaunch(b->graph_exec, b->stream));
  }
  const int rest = n - full * chunk;
  if (rest && (rc = launch_steps(b, rest))) return rc;
  advance_members(b, n);
  return LBM_OK;
}

int lbm_batch_synchronize(lbm_batch* b) {
  if (int rc = check_batch(b)) return rc;
  LBM_HIP(hipStreamSynchronize(b->stream));
  return LBM_OK;
}

int lbm_batch_time_steps(lbm_batch* b, int n, int chunk, float* ms) {
  if (int rc = check_batch(b)) return rc;
  if (!ms) return fail(LBM_ERR_INVALID, "null output");
  if (chunk > 0 && (chunk < 2 || (chunk & 1)))
    return fail(LBM_ERR_INVALID, "chunk must be an even step count >= 2");
  int rc = sync_tables(b);
  if (rc) return rc;
  if (chunk > 0 && b->ntiles > 0 && (rc = ensure_graph(b, chunk))) return rc;
  LBM_HIP(hipEventRecord(b->ev0, b->stream));
  rc = chunk > 0 ? api::lbm_batch_step_graph(b, n, chunk) : api::lbm_batch_step(b, n);
  if (rc) return rc;
  LBM_HIP(hipEventRecord(b->ev1, b->stream));
  LBM_HIP(hipEventSynchronize(b->ev1));
  LBM_HIP(hipEventElapsedTime(ms, b->ev0, b->ev1));
  return LBM_OK;
}

int lbm_batch_autotune(lbm_batch* b, int steps, int use_cache, int* block_out,
                       int* grid_out) {
  if (int rc = check_batch(b)) return rc;
  if (steps < 2) return fail(LBM_ERR_INVALID, "autotune needs >= 2 steps");
  TraceRange tr("lbm_batch_autotune");
  steps += steps & 1;  // even: members end where they started
  const auto key = tune_key(b);
  if (use_cache) {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      b->block = it->second.first;
      b->grid = it->second.second;
      if (block_out) *block_out = b->block;
      if (grid_out) *grid_out = b->grid;
      return LBM_OK;
    }
  }
  float best = 0.0f;
  int best_block = b->block, best_grid = b->grid;
  bool have = false;
  int launched = 0;
  const int cus = b->device_cus;
  for (int block : {64, 128, 256, 512}) {
    for (int mult : {0, 2, 4, 8, 16}) {
      b->block = block;
      b->grid = mult * cus;
      if (b->grid == 0 && block == 64) continue;
      int rc = sync_tables(b);
      if (rc) return rc;
      if ((rc = launch_steps(b, 2))) return rc;  // warm-up (even)
      LBM_HIP(hipEventRecord(b->ev0, b->stream));
      if ((rc = launch_steps(b, steps))) return rc;
      launched += steps + 2;
      LBM_HIP(hipEventRecord(b->ev1, b->stream));
      LBM_HIP(hipEventSynchronize(b->ev1));
      float ms = 0.0f;
      LBM_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
      if (!have || ms < best) {
        best = ms;
        best_block = block;
        best_grid = b->grid;
        have = true;
      }
    }
  }
  b->block = best_block;
  b->grid = best_grid;
  advance_members(b, launched);  // an even count: buffers end where they began
  {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    g_tune_cache[key] = {best_block, best_grid};
  }
  if (block_out) *block_out = best_block;
  if (grid_out) *grid_out = best_grid;
  return sync_tables(b);
}

int lbm_batch_residuals(lbm_batch* b, double* out, int32_t nslots) {
  if (int rc = check_batch(b)) return rc;
  if (!out || nslots < (int32_t)b->slots.size())
    return fail(LBM_ERR_INVALID, "output must hold one value per slot");
  TraceRange tr("lbm_batch_residuals");
  int rc = sync_tables(b);
  if (rc) return rc;
  for (int32_t s = 0; s < nslots; ++s) out[s] = -1.0;
  if (b->nchunks == 0) return LBM_OK;
  if ((rc = launch_batch_residual(device_storage(LBM_MODE_STANDARD, b->storage), b->d_lats,
                                  b->d_chunks, b->nchunks,
                                  b->d_partials, b->stream)))
    return rc;
  std::vector<double> part(kResidualPartials * (size_t)b->nchunks);
  LBM_HIP(hipMemcpyAsync(part.data(), b->d_partials, sizeof(double) * part.size(),
                         hipMemcpyDeviceToHost, b->stream));
  LBM_HIP(hipStreamSynchronize(b->stream));
  const double limit2 = LBM_DIVERGED_SPEED * LBM_DIVERGED_SPEED;
  for (size_t k = 0; k < b->h_lats.size(); ++k) {
    const int32_t s = b->h_lat_slot[k];
    if (!b->snapshots[s]) continue;
    double d2 = 0.0, u2 = 0.0, m2 = 0.0;
    const uint32_t c0 = b->h_lats[k].res_first;
    for (uint32_t c = c0; c < c0 + b->h_res_count[k]; ++c) {
      d2 += part[kResidualPartials * c];
      u2 += part[kResidualPartials * c + 1];
      m2 = std::max(m2, part[kResidualPartials * c + 2]);
    }
    // a non-finite velocity anywhere, or a speed beyond any stable BGK
    // flow, marks the lattice as diverged
    if (!std::isfinite(d2) || !std::isfinite(u2) || !(m2 <= limit2))
      out[s] = NAN;
    else
      out[s] = u2 > 0.0 ? std::sqrt(d2 / u2) : (d2 > 0.0 ? INFINITY : 0.0);
  }
  return LBM_OK;
}

}  // namespace api
}  // namespace lbm
