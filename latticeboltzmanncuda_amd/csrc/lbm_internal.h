// Internal state of a lattice context and the launchers each kernel file
// exports to the C-ABI layer.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lbm.h"

namespace lbm {

// A roctx range over a host-side scope -- step calls, hipGraph chunks,
// serving chunks, geometry builds -- shown by `rocprofv3 --marker-trace`
// on the timeline beside the kernels it launched (a no-op call when no
// profiler is attached).
struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

// Division by a runtime constant d for dividends below 2^31:
// q = (umulhi(n, mul) + n) >> shift  (Granlund-Montgomery round-up method).
struct Magic {
  uint32_t mul;
  uint32_t shift;
  __host__ __device__ uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    const uint32_t t = __umulhi(n, mul);
#else
    const uint32_t t = (uint32_t)(((uint64_t)n * mul) >> 32);
#endif
    return (t + n) >> shift;
  }
};

inline Magic make_magic(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  const uint64_t m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
  return Magic{(uint32_t)m, l};
}

// Grid::flags: the periodic axes, and whether the 3-D standard step reads
// every node's link mask directly (lattices with solid nodes) instead of
// its 2-bit class first (the mask only for boundary nodes).
constexpr uint32_t kWrapX = 1u, kWrapY = 2u, kWrapZ = 4u, kReadMasks = 8u;
inline uint32_t wrap_bits(const int periodic[3]) {
  return (periodic[0] ? kWrapX : 0u) | (periodic[1] ? kWrapY : 0u) | (periodic[2] ? kWrapZ : 0u);
}

// Geometry of a lattice as the kernels see it.
struct Grid {
  uint32_t nx, ny, nz;
  uint32_t n;     // nodes
  Magic dz, dy;   // division by nz and by ny
  uint32_t flags;  // kWrapX | kWrapY | kWrapZ (periodic axes) | kReadMasks
};

// Per-lattice constants of the standard step in the compute type C,
// precomputed on the host so that the kernels spend no vector instructions
// on values that are the same for every node.  Index i is the direction.
template <class C>
struct StepCoef {
  C omega;           // 1 / tau
  C half_force[3];   // F / 2 (velocity correction of the Guo scheme)
  C force[3];        // F
  C lid[27];         // 6 w_i rho_w (e_i . u_w) if a link of direction i can
                     // come from beyond the moving face, else 0
  C gw[27];          // (1 - omega / 2) w_i
  C gf3[27];         // 3 (1 - omega / 2) w_i (e_i . F)
  C gf9[27];         // 9 (1 - omega / 2) w_i (e_i . F)
  int forced;        // F != 0 (else the Guo terms are skipped)
};

// Parameters of the standard (BGK) step.
struct StdParams {
  float tau;       // relaxation time
  float force[3];  // Guo body force
  int lid_axis;    // -1: no moving wall
  int lid_sign;    // +1 max face, -1 min face
  float lid_u[3];
  float wall_rho;
  StepCoef<float> cf;   // FP32 compute (FP32 / BF16 storage)
  StepCoef<double> cd;  // FP64 compute (FP64 storage)
  // FP32 compute in FP16 stored units (2^12 x real): cf with every term
  // that is added to a stored value (half_force, lid, gw, gf3, gf9) scaled
  // by 2^12; omega and force as in cf
  StepCoef<float> ch;
};

struct FlowSpec {
  int kind;
  int32_t params[6];
  float induced[27];
  const uint8_t* mask;
};

StdParams make_std_params(const lbm_config& cfg);
// The checks lbm_create applies to a configuration (LBM_OK or an error
// code with lbm_last_error set).
int check_config(const lbm_config* cfg);
// device bytes lbm_create allocates for cfg (populations, geometry, scratch)
int64_t estimate_ctx_bytes(const lbm_config& cfg);
// BGK cell Reynolds number U / nu of a standard-mode configuration (the
// rule of lbm_cell_reynolds); cfg must pass check_config.
double cell_reynolds(const lbm_config& cfg);
// The moments of lbm_moments (rho and / or u, either may be null) staged
// through the device buffer `stage` -- 4 N bytes for rho, then 12 N for u
// (12 N from the start when rho is null) -- and copied to the host.
int read_moments(lbm_ctx* c, void* stage, float* rho, float* u);

// error helpers --------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// Status of an exception caught at the C-ABI boundary, from inside the
// catch block (std::bad_alloc -> LBM_ERR_NOMEM, anything else ->
// LBM_ERR_INVALID), with the message set when the message itself can be
// stored.
int api_exception() noexcept;

#define LBM_HIP(call)                                     \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return ::lbm::hip_fail(e_, #call); \
  } while (0)

// launchers: standard mode (k_standard.hip) -----------------------------
int launch_std_step(int q, int storage, const Grid& g, const void* A, void* B,
                    const uint32_t* mask, const uint32_t* cls,
                    const StdParams& p, int rev, hipStream_t s);
int launch_build_mask(int q, const Grid& g, const uint8_t* solid,
                      const int periodic[3], int lid_face, uint32_t* mask,
                      uint32_t* cls, hipStream_t s);
// Whether a standard-mode lattice keeps the 16-bit copy of its link masks
// (D2Q9 other than FP64), and its refresh from the 32-bit masks.
inline bool has_mask16(const lbm_config& cfg) {
  return cfg.mode == LBM_MODE_STANDARD && cfg.lattice == 9 && cfg.storage != LBM_STORE_F64;
}
int launch_mask16(const Grid& g, uint32_t* mask, hipStream_t s);
int launch_init_eq(int q, int storage, const Grid& g, const double* rho,
                   const double* u, void* f, hipStream_t s);
int launch_std_moments(int q, int storage, const Grid& g, const void* f,
                       const uint32_t* mask, const float force[3], float* rho,
                       float* u, hipStream_t s);
int launch_conservation(int q, int storage, const Grid& g, const void* f,
                        const uint32_t* mask, double* partials, int nblocks,
                        double* out4, hipStream_t s);
int launch_convert_in(int q, int storage, const Grid& g, const void* src,
                      int src_is_f64, void* f, hipStream_t s);
int launch_convert_out(int q, int storage, const Grid& g, const void* f,
                       void* dst, int dst_is_f64, hipStream_t s);

// launchers: reference mode (k_reference.hip) ---------------------------
int launch_ref_stream(const Grid& g, const float* A, float* B, hipStream_t s);
int launch_ref_collide(const Grid& g, float* f, float tau,
                       const FlowSpec* flow, hipStream_t s);
int launch_ref_stream_collide(const Grid& g, const float* A, float* B,
                              float tau, hipStream_t s);
int launch_ref_reflect(const Grid& g, float* f, int32_t nspecial,
                       const int32_t* special, const int32_t* special_rec,
                       const uint8_t* flags, const uint32_t* rec_dirs,
                       const float* rec_w, float* force_partials,
                       int nblocks, float* force, hipStream_t s);
int ref_reflect_blocks(int32_t nspecial);
// Fused stream + reflect + collide + flow (A -> B).  With reflections
// (cls non-null) a gather pass over the (receiver, direction) groups runs
// first; partials (3 per gather wavefront, ref_gather_partials) and force
// may be null to skip the force.
int launch_ref_step_full(const Grid& g, const float* A, float* B, float tau,
                         const FlowSpec& flow, const uint32_t* cls,
                         const uint32_t* inmask, const int32_t* rxbase,
                         int64_t ngrp, const uint32_t* gsrc, const uint32_t* gbeg,
                         const uint32_t* gid, float* gsum, int64_t nent, const uint32_t* eidx,
                         const uint32_t* keep,
                         const float* ew, const uint8_t* es, float* partials,
                         float* force, int rev, hipStream_t s);
int ref_gather_partials(int64_t ngrp);
int launch_ref_flow(const Grid& g, float* f, const FlowSpec& flow,
                    hipStream_t s);
int launch_ref_moments(const Grid& g, const float* f, float* rho, float* u,
                       hipStream_t s);

// launchers: shared (k_common.hip) --------------------------------------
int launch_points_in(int q, int storage, const Grid& g,
                     const lbm_lattice_point* pts, void* f, uint8_t* flags,
                     hipStream_t s);
int launch_points_out(int q, int storage, const Grid& g, const void* f,
                      const uint8_t* flags, lbm_lattice_point* pts,
                      hipStream_t s);
int launch_test_coordinates(const Grid& g, uint32_t* counts, int* ok,
                            hipStream_t s);
int launch_copy_stream(const void* src, void* dst, uint64_t bytes, hipStream_t s);

}  // namespace lbm

// The context behind the opaque lbm_ctx handle.
struct lbm_ctx {
  lbm_config cfg{};
  int q = 0;
  lbm::Grid grid{};
  size_t elem = 4;
  void* f[2] = {nullptr, nullptr};
  int cur = 0;
  uint32_t* mask = nullptr;  // standard mode: link mask per node
  uint32_t* cls = nullptr;   // standard mode: 2-bit node classes
  uint64_t steps_done = 0;   // parity picks the tile order of the next step
  uint64_t params_version = 0;  // bumped when tau or the body force change
  uint8_t* flags = nullptr;  // reference mode
  std::vector<uint8_t> host_flags;

  // reflection records (reference mode)
  int32_t nrec = 0;
  int32_t* rec_node = nullptr;
  uint32_t* rec_dirs = nullptr;
  float* rec_w = nullptr;
  std::vector<int32_t> host_rec_node;
  std::vector<uint32_t> host_rec_dirs;  // 14 packed words per record
  std::vector<float> host_rec_w;        // 81 weights per record
  // nodes with a nonzero flag byte, and their record (or -1)
  int32_t nspecial = 0;
  int32_t* special = nullptr;
  int32_t* special_rec = nullptr;
  bool special_dirty = true;

  // flow injection
  lbm::FlowSpec flow{};
  uint8_t* flow_mask = nullptr;

  // force register and reduction scratch
  float* force = nullptr;
  float* force_partials = nullptr;
  int force_partials_cap = 0;

  // fused reference step (stream + reflect + collide + flow in one pass):
  // a gather pass sums, for every (receiving node t, direction j), the
  // streamed value A_j(x) of the reflecting node x = t - e_j and the
  // reflected slot values W[s] F_{s/3}(x) in ascending s; the fused pass
  // then takes that sum as f_j(t)
  bool ref_geom = false;           // any flagged node
  uint32_t* ref_cls = nullptr;     // 2-bit: 0 plain, 1 receiver, 2 internal
  uint32_t* ref_keep = nullptr;    // bit per tile: the gather reads it (cached stores)
  uint32_t* ref_inmask = nullptr;  // per node: bit j = a reflected link arrives in direction j
  int32_t* ref_rxbase = nullptr;   // per node: first group of a receiver (groups ascend in j)
  int64_t ref_ngrp = 0;            // groups (t, j)
  // the gather's threads take the groups by reflecting node, then direction
  uint32_t* ref_gsrc = nullptr;    // per thread: element index of its group's A_j(x)
  uint32_t* ref_gbeg = nullptr;    // per thread: first entry; the last ends at ref_nent
  uint32_t* ref_gid = nullptr;     // per thread: its group (slot of gsum)
  float* ref_gsum = nullptr;       // per group: the gathered f_j(t) of the current step
  int64_t ref_nent = 0;            // entries (one per reflected slot)
  uint32_t* ref_eidx = nullptr;    // per entry: element index of F_k(x) = A_k(x - e_k), ~0u outside
  float* ref_ew = nullptr;         // per entry: slot weight W[s]
  uint8_t* ref_es = nullptr;       // per entry: slot s (its force term comes from a table)
  float* ref_partials = nullptr;   // force partials of the gather pass, 3 per wavefront
  int ref_partials_cap = 0;
  double* red = nullptr;
  int red_cap = 0;
  double* red_out = nullptr;
  // device staging of lbm_moments (rho and u), allocated by its first call
  // and kept, so later calls allocate nothing
  void* stage = nullptr;
  size_t stage_bytes = 0;

  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int device = 0;
  int64_t device_bytes = 0;
};
