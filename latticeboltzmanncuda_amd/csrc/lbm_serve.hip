// Continuous-batching solve server (C-ABI lbm_serve_run).
//
// Requests arrive at given times; each is a standard-mode lattice to run
// until its velocity residual |u^n - u^(n-chunk)| / |u^n| drops below its
// tolerance or it reaches its step limit.  The loop:
//   1. moves arrived requests to a FIFO;
//   2. admits from the FIFO head while the batch's HBM budget (both
//      population buffers, geometry, residual snapshot) and the resident
//      limits (lattices, nodes) allow -- requests that arrive while others
//      run join at the next chunk boundary as memory frees up.  A node
//      limit a little above what fills the GPU keeps the throughput and
//      lets early requests finish first instead of sharing the GPU with
//      every later one;
//   3. advances every resident lattice by one chunk: with the batch's
//      persistent kernel (replayed hipGraph), or -- the comparison
//      engine -- with one plain launch per lattice per step;
//   4. measures all residuals in one launch, retires finished lattices
//      (velocity copied to the host, latency = arrival -> result on host)
//      and frees their memory.
// A lattice's arithmetic and its stop decision depend only on its own
// state (per-node updates, residuals summed over fixed node chunks), so a
// request gives the same result alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <deque>
#include <thread>
#include <vector>

#include "lbm_batch.h"
#include "lbm_internal.h"
#include "lbm_store.h"

using namespace lbm;

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

double percentile(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = p * (double)(v.size() - 1);
  const size_t lo = (size_t)std::floor(pos), hi = (size_t)std::ceil(pos);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

// Device staging of retirement readbacks, freed when the server returns
// (each readback synchronises before the next one reuses it).
struct StageBuffer {
  void* p = nullptr;
  size_t bytes = 0;
  ~StageBuffer() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int grow(size_t n) {
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return hip_fail(e, "hipMalloc(serving readback staging)");
    }
    bytes = n;
    return LBM_OK;
  }
};

int validate_request(const lbm_request& r, int storage) {
  if (int rc = check_config(&r.cfg)) return rc;
  if (r.cfg.mode != LBM_MODE_STANDARD)
    return fail(LBM_ERR_UNSUPPORTED, "requests must be standard-mode lattices");
  if (r.cfg.storage != storage)
    return fail(LBM_ERR_INVALID, "request storage differs from the server storage");
  if (r.max_steps < 0 || !(r.arrival_s >= 0.0))
    return fail(LBM_ERR_INVALID, "bad request limits");
  return LBM_OK;
}

// Chooses the launch shape on a scratch batch built from the first
// requests, so the served lattices are not stepped by the tuning runs.
int pretune(const lbm_serve_params& p, const lbm_request* reqs, int32_t n,
            lbm_batch* target) {
  lbm_batch* t = nullptr;
  int rc = lbm_batch_create(p.storage, p.budget_bytes, nullptr, &t);
  if (rc) return rc;
  int added = 0;
  for (int32_t k = 0; k < n && added < 32; ++k) {
    int32_t slot;
    if (lbm_batch_add(t, &reqs[k].cfg, reqs[k].solid, 0, &slot) == LBM_OK) ++added;
  }
  int block = 256, grid = 0;
  if (added > 0) rc = lbm_batch_autotune(t, 10, 1, &block, &grid);
  lbm_batch_destroy(t);
  if (rc) return rc;
  return lbm_batch_set_launch(target, block, grid);
}

}  // namespace

// Entry point of include/lbm.h (exported by lbm_exports.hip).
namespace lbm {
namespace api {

int lbm_serve_run(const lbm_serve_params* p, lbm_request* reqs, int32_t nreq,
                  lbm_serve_stats* stats) {
  if (!p || (!reqs && nreq > 0) || nreq < 0 || !stats)
    return fail(LBM_ERR_INVALID, "null argument");
  if (p->chunk < 2 || (p->chunk & 1))
    return fail(LBM_ERR_INVALID, "chunk must be an even step count >= 2");
  if (p->engine != LBM_ENGINE_BATCHED && p->engine != LBM_ENGINE_PER_LATTICE)
    return fail(LBM_ERR_INVALID, "unknown serving engine");
  if (!(p->max_cell_re >= 0.0)) return fail(LBM_ERR_INVALID, "max_cell_re must be >= 0");
  for (int32_t k = 0; k < nreq; ++k) {
    if (int rc = validate_request(reqs[k], p->storage)) return rc;
    reqs[k].t_admit_s = reqs[k].t_done_s = -1.0;
    reqs[k].steps = 0;
    reqs[k].residual = -1.0;
    reqs[k].status = -1;
  }
  *stats = lbm_serve_stats{};
  lbm_batch* b = nullptr;
  int rc = lbm_batch_create(p->storage, p->budget_bytes, nullptr, &b);
  if (rc) return rc;
  if (p->engine == LBM_ENGINE_BATCHED && p->autotune && (rc = pretune(*p, reqs, nreq, b))) {
    lbm_batch_destroy(b);
    return rc;
  }
  std::vector<int32_t> order(nreq);
  for (int32_t k = 0; k < nreq; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t c) {
    return reqs[a].arrival_s < reqs[c].arrival_s;
  });
  std::deque<int32_t> fifo;
  std::vector<int32_t> slot_req;  // slot -> request (-1 free)
  int32_t resident = 0, next = 0, finished = 0;
  int64_t resident_nodes = 0;
  // algorithmic bytes one step of the resident lattices moves: 2 Q s per
  // node (every population read once and written once)
  double resident_step_bytes = 0.0, pop_bytes = 0.0;
  const double elem = (double)storage_bytes(p->storage);
  auto step_bytes = [&](const lbm_config& cfg) {
    return (double)cfg.nx * cfg.ny * cfg.nz * 2.0 * cfg.lattice * elem;
  };
  std::vector<double> residuals;
  std::vector<float> scratch;
  StageBuffer stage;
  std::vector<lbm_ctx*> members;
  int64_t updates = 0;
  const auto t0 = Clock::now();
  double t_first = -1.0, t_last = 0.0;

  auto retire = [&](int32_t slot, int32_t status, double res) -> int {
    const int32_t r = slot_req[slot];
    lbm_ctx* c = nullptr;
    int e = lbm_batch_member(b, slot, &c);
    if (e) return e;
    const size_t n3 = 3 * (size_t)c->grid.n;
    float* dst = reqs[r].out_u;
    if (!dst) {
      scratch.resize(n3);
      dst = scratch.data();
    }
    // the velocity is staged through one device buffer of the server,
    // grown to the largest lattice retired so far (no allocation per
    // request)
    if (stage.bytes < sizeof(float) * n3) {
      stage.release();
      if ((e = stage.grow(sizeof(float) * n3))) return e;
    }
    if ((e = read_moments(c, stage.p, nullptr, dst))) return e;
    reqs[r].t_done_s = seconds_since(t0);
    reqs[r].residual = res;
    reqs[r].status = status;
    t_last = reqs[r].t_done_s;
    if ((e = lbm_batch_remove(b, slot))) return e;
    slot_req[slot] = -1;
    --resident;
    resident_nodes -= (int64_t)reqs[r].cfg.nx * reqs[r].cfg.ny * reqs[r].cfg.nz;
    resident_step_bytes -= step_bytes(reqs[r].cfg);
    ++finished;
    return LBM_OK;
  };

  double next_report = p->progress_s > 0.0 ? p->progress_s : INFINITY;
  while (finished < nreq) {
    const double now = seconds_since(t0);
    if (now >= next_report) {
      std::fprintf(stderr, "lbm_serve: t=%.1fs finished %d/%d resident %d queued %zu\n", now,
                   finished, nreq, resident, fifo.size() + (size_t)(nreq - next));
      std::fflush(stderr);
      next_report = now + p->progress_s;
    }
    while (next < nreq && reqs[order[next]].arrival_s <= now) fifo.push_back(order[next++]);
    // admission, FIFO
    while (!fifo.empty()) {
      const int32_t r = fifo.front();
      if (p->max_cell_re > 0.0 && !(cell_reynolds(reqs[r].cfg) <= p->max_cell_re)) {
        // a BGK flow this far beyond the stability limit would diverge
        reqs[r].status = 4;
        reqs[r].t_done_s = seconds_since(t0);
        fifo.pop_front();
        ++finished;
        ++stats->unstable;
        continue;
      }
      const int64_t need = lbm_batch_estimate_bytes(&reqs[r].cfg, 1);
      if (need > p->budget_bytes) {
        reqs[r].status = 2;
        reqs[r].t_done_s = seconds_since(t0);
        fifo.pop_front();
        ++finished;
        ++stats->rejected;
        continue;
      }
      if (p->max_resident > 0 && resident >= p->max_resident) break;
      const int64_t nodes = (int64_t)reqs[r].cfg.nx * reqs[r].cfg.ny * reqs[r].cfg.nz;
      if (p->max_resident_nodes > 0 && resident > 0 &&
          resident_nodes + nodes > p->max_resident_nodes)
        break;
      int32_t slot = -1;
      const int e = lbm_batch_add(b, &reqs[r].cfg, reqs[r].solid, 1, &slot);
      if (e == LBM_ERR_BUDGET && resident > 0) break;  // wait for memory to free up
      if (e == LBM_ERR_BUDGET) {
        // nothing resident to wait for: the request with the batch tables
        // does not fit the budget
        reqs[r].status = 2;
        reqs[r].t_done_s = seconds_since(t0);
        fifo.pop_front();
        ++finished;
        ++stats->rejected;
        continue;
      }
      if (e) {
        lbm_batch_destroy(b);
        return e;
      }
      if ((int32_t)slot_req.size() <= slot) slot_req.resize(slot + 1, -1);
      slot_req[slot] = r;
      reqs[r].t_admit_s = seconds_since(t0);
      if (t_first < 0.0) t_first = reqs[r].t_admit_s;
      fifo.pop_front();
      ++resident;
      resident_nodes += nodes;
      resident_step_bytes += step_bytes(reqs[r].cfg);
      stats->max_resident = std::max(stats->max_resident, resident);
      int64_t bytes = 0;
      lbm_batch_info(b, nullptr, nullptr, &bytes);
      stats->peak_bytes = std::max(stats->peak_bytes, bytes);
    }
    if (resident == 0) {
      if (next >= nreq && fifo.empty()) break;
      if (fifo.empty() && next < nreq) {
        const double wait = reqs[order[next]].arrival_s - seconds_since(t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.001)));
      }
      continue;
    }
    // one chunk for everybody resident (the range also covers the
    // residuals and retirements that follow it)
    TraceRange chunk_range("lbm_serve chunk");
    int64_t nodes = 0;
    lbm_batch_info(b, nullptr, &nodes, nullptr);
    if (p->engine == LBM_ENGINE_PER_LATTICE) {
      members.clear();
      for (int32_t s = 0; s < (int32_t)slot_req.size(); ++s) {
        lbm_ctx* c = nullptr;
        if (slot_req[s] >= 0 && lbm_batch_member(b, s, &c) == LBM_OK) members.push_back(c);
      }
      rc = lbm_step_each(members.data(), (int32_t)members.size(), p->chunk);
    } else {
      rc = lbm_batch_step_graph(b, p->chunk, p->chunk);
    }
    if (rc) {
      lbm_batch_destroy(b);
      return rc;
    }
    updates += nodes * (int64_t)p->chunk;
    pop_bytes += resident_step_bytes * p->chunk;
    residuals.assign(slot_req.size(), -1.0);
    if ((rc = lbm_batch_residuals(b, residuals.data(), (int32_t)residuals.size()))) {
      lbm_batch_destroy(b);
      return rc;
    }
    for (int32_t s = 0; s < (int32_t)slot_req.size(); ++s) {
      const int32_t r = slot_req[s];
      if (r < 0) continue;
      reqs[r].steps += p->chunk;
      const double res = residuals[s];
      const bool converged = res >= 0.0 && res < reqs[r].tol;
      const bool limit = reqs[r].steps >= reqs[r].max_steps;
      // a diverged lattice reads NaN (lbm_batch_residuals): retired now,
      // the other lattices' arithmetic does not depend on it
      if (converged || limit || !std::isfinite(res)) {
        const int32_t status = converged ? 0 : (std::isfinite(res) ? 1 : 3);
        if (status == 3) ++stats->diverged;
        if ((rc = retire(s, status, res))) {
          lbm_batch_destroy(b);
          return rc;
        }
      }
    }
  }
  lbm_batch_destroy(b);
  std::vector<double> lat;
  for (int32_t k = 0; k < nreq; ++k)
    if (reqs[k].status == 0 || reqs[k].status == 1 || reqs[k].status == 3)
      lat.push_back(reqs[k].t_done_s - reqs[k].arrival_s);
  stats->completed = (int32_t)lat.size();
  stats->wall_s = t_first < 0.0 ? 0.0 : t_last - t_first;
  stats->lattice_updates = updates;
  stats->population_bytes = pop_bytes;
  stats->mlups = stats->wall_s > 0.0 ? (double)updates / stats->wall_s / 1e6 : 0.0;
  stats->p50_latency_s = percentile(lat, 0.50);
  stats->p99_latency_s = percentile(lat, 0.99);
  double sum = 0.0;
  for (double v : lat) sum += v;
  stats->mean_latency_s = lat.empty() ? 0.0 : sum / (double)lat.size();
  return LBM_OK;
}

}  // namespace api
}  // namespace lbm
