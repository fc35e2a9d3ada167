/*
 * lbm.h -- C-ABI of the MI355X (gfx950) lattice-Boltzmann library
 * (liblbm_hip.so).  Plain C types and pointers only; no HIP or C++ types
 * cross this boundary.  Every entry point returns an lbm_status; the
 * message of the last failure on the calling thread is lbm_last_error().
 *
 * Conventions kept from the reference (CharlieNicholls/LatticeBoltzmannCuda):
 *   - node index n = z + nz*(y + ny*x)   (test/CoordinatesTest.cpp:80-82,
 *     simulation/source/Lattice.cpp:359); 2-D lattices use nz = 1;
 *   - D3Q27 velocity order and inverse map of
 *     simulation/source/Constants.cpp:6-7; D3Q19 is its first 19 entries;
 *     D2Q9 is its z = 0 subset in the same order;
 *   - 144-byte LatticePoint host records (simulation/include/LatticePoint.h:3-17)
 *     as the exchange format of Lattice::load_data / retrieve_data.
 *
 * Device layout (not visible through the ABI): blocked structure of
 * arrays -- tiles of 64 consecutive nodes, inside a tile one contiguous run
 * of 64 values per lattice direction -- double-buffered.
 *
 * The reference's launch boundary is namespace RunCudaFunctions
 * (simulation/include/SimCalcFuncs.cuh:13-26, SimCalcFuncs.cu:218-249),
 * C++ linkage, taking (dim3 blocks, dim3 threads, LatticeData, ...).  Each
 * replacement below names the function it stands in for.
 */
#ifndef LBM_H
#define LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 1

/* ------------------------------------------------------------ status */
enum lbm_status {
  LBM_OK = 0,
  LBM_ERR_INVALID = 1,     /* bad argument or configuration            */
  LBM_ERR_HIP = 2,         /* a HIP runtime call failed                */
  LBM_ERR_NOMEM = 3,       /* device or host allocation failed         */
  LBM_ERR_STATE = 4,       /* call not valid in the context's mode     */
  LBM_ERR_UNSUPPORTED = 5, /* combination not built                    */
  LBM_ERR_BUDGET = 6,      /* HBM budget of a batch would be exceeded  */
  LBM_ERR_NOTFOUND = 7,    /* unknown handle / slot                    */
  LBM_ERR_UNSTABLE = 8     /* lattice beyond the admission stability limit */
};

/* ------------------------------------------------------------ enums */
enum lbm_lattice { LBM_D2Q9 = 9, LBM_D3Q19 = 19, LBM_D3Q27 = 27 };

/* Population storage.  F16/BF16 store the shifted deviation f_i - w_i;
 * collision is always computed in FP32 for them and in FP64 for F64. */
enum lbm_storage {
  LBM_STORE_F32 = 0,
  LBM_STORE_F64 = 1,
  LBM_STORE_F16 = 2,
  LBM_STORE_BF16 = 3
};

/* REFERENCE mode reproduces the reference kernels (D3Q27, F32 only):
 *   zero-fill pull streaming, SPLIT velocity u = sum f e /|e|^2 / rho,
 *   FP64-evaluated equilibrium, reflection tables, flow injection.
 * STANDARD mode is BGK with half-way bounce-back, periodic axes, one
 * optional moving wall and Guo body forcing. */
enum lbm_mode { LBM_MODE_STANDARD = 0, LBM_MODE_REFERENCE = 1 };

/* Flow-injection criteria (reference FlowCriterion.cu:5-16 are POINT(1,1,1)
 * and PLANE(axis 0, 0)); the criterion is data, not a device function. */
enum lbm_flow_kind {
  LBM_FLOW_NONE = 0,
  LBM_FLOW_POINT = 1, /* params: x, y, z                         */
  LBM_FLOW_PLANE = 2, /* params: axis (0..2), coordinate          */
  LBM_FLOW_BOX = 3,   /* params: x0, y0, z0, x1, y1, z1 inclusive */
  LBM_FLOW_MASK = 4   /* per-node byte mask, nonzero = inject     */
};

/* per-node flag byte of reference mode (LatticePoint::isInternal /
 * isReflected) */
#define LBM_FLAG_INTERNAL 1u
#define LBM_FLAG_REFLECTED 2u

/* standard-mode link mask word: bit i = link i bounces back (its source
 * x - e_i is solid or beyond a wall); LID = node borders the moving wall;
 * SOLID = node is solid and is not updated. */
#define LBM_MASK_SOLID (1u << 31)
#define LBM_MASK_LID (1u << 30)

/* A lattice whose speed |u| exceeds this anywhere (lattice units; the
 * speed of sound is 1/sqrt(3) ~ 0.577, and stable BGK flows stay far
 * below it) counts as diverged, as does a non-finite velocity. */
#define LBM_DIVERGED_SPEED 0.5

/* lid_face: -1 = none, else 2*axis + side (side 0 = min face, 1 = max) */
typedef struct lbm_config {
  int32_t nx, ny, nz;
  int32_t lattice; /* enum lbm_lattice */
  int32_t storage; /* enum lbm_storage */
  int32_t mode;    /* enum lbm_mode    */
  float tau;       /* relaxation time (FluidData::m_characteristicTimescale) */
  int32_t periodic[3];
  int32_t lid_face;
  float lid_u[3];
  float wall_rho;  /* density used in the moving-wall term, usually 1 */
  float force[3];  /* Guo body force per node (standard mode) */
} lbm_config;

/* Byte-compatible with the reference's LatticePoint (144 bytes on x86-64):
 * int x,y,z @0; float f[27] @12; bool isReflected @120; bool isInternal
 * @121; two pointers @128/@136.  The library never dereferences the
 * pointer fields; downloads write them as NULL. */
typedef struct lbm_lattice_point {
  int32_t x, y, z;
  float particle_distribution[27];
  uint8_t isReflected;
  uint8_t isInternal;
  uint8_t pad_[6];
  void* reflections;
  void* d_reflections;
} lbm_lattice_point;

typedef struct lbm_ctx lbm_ctx;

/* ------------------------------------------------------------ library */
const char* lbm_last_error(void);
int lbm_abi_version(void);
int lbm_device_info(int* device_count, char* name, size_t name_len);

/* ------------------------------------------------------------ lattice */
/* Lattice::Lattice / createExtent / allocateLatticeArray
 * (simulation/source/Lattice.cpp:22-54).  Populations start at zero (the
 * reference leaves them uninitialised).  stream == NULL gives the context
 * its own HIP stream (default priority); otherwise the hipStream_t passed
 * as void* is used and not destroyed. */
int lbm_create(const lbm_config* cfg, void* stream, lbm_ctx** out);
/* Lattice::~Lattice (Lattice.cpp:41-44) */
int lbm_destroy(lbm_ctx* ctx);
int lbm_get_config(const lbm_ctx* ctx, lbm_config* out);
int64_t lbm_node_count(const lbm_ctx* ctx);
/* device bytes held by the context */
int64_t lbm_device_bytes(const lbm_ctx* ctx);
int lbm_set_tau(lbm_ctx* ctx, float tau);
int lbm_set_body_force(lbm_ctx* ctx, const float force[3]);
int lbm_synchronize(lbm_ctx* ctx);

/* ------------------------------------------------------------ exchange */
/* Lattice::load_data (Lattice.cpp:73-85): count must equal nx*ny*nz.
 * Copies f[0..Q-1] (converted to the storage type) and, in reference
 * mode, isInternal/isReflected into the flag bytes. */
int lbm_upload_points(lbm_ctx* ctx, const lbm_lattice_point* pts,
                      int64_t count);
/* Lattice::retrieve_data (Lattice.cpp:87-113): x,y,z from the index,
 * populations (entries >= Q are 0), flags; pointer fields NULL. */
int lbm_download_points(lbm_ctx* ctx, lbm_lattice_point* pts, int64_t count);
/* Populations as Q planes of N values (plane i holds direction i). */
int lbm_set_populations_f32(lbm_ctx* ctx, const float* f);
int lbm_get_populations_f32(lbm_ctx* ctx, float* f);
int lbm_set_populations_f64(lbm_ctx* ctx, const double* f);
int lbm_get_populations_f64(lbm_ctx* ctx, double* f);
/* f_i = feq_i(rho, u); rho (N values) and u (3 planes of N) may be NULL
 * for rho = 1 and u = 0. */
int lbm_init_equilibrium(lbm_ctx* ctx, const double* rho, const double* u);

/* ------------------------------------------------------------ geometry */
/* Standard mode obstacle/boundary upload: one byte per node (nonzero =
 * solid); the link masks are built on the device from it and from the
 * periodic / lid configuration.  solid == NULL means no obstacles. */
int lbm_upload_solid(lbm_ctx* ctx, const uint8_t* solid);
int lbm_get_mask(lbm_ctx* ctx, uint32_t* mask);
/* Reference mode flags (isInternal/isReflected of Lattice::insertModel,
 * Lattice.cpp:228-260). */
int lbm_upload_flags(lbm_ctx* ctx, const uint8_t* flags);
int lbm_get_flags(lbm_ctx* ctx, uint8_t* flags);
/* Reflection records of Lattice::preProcessModel (Lattice.cpp:323-426):
 * for record r, node[r] is the node index, dirs[14*r..] the 5-bit packed
 * slot directions (slot s in word s/6, bits 5*(s%6)) and weights[81*r..]
 * the slot weights.  Replaces the per-node d_reflections pointers.  A
 * record's node must carry LBM_FLAG_INTERNAL when lbm_reflect runs. */
int lbm_upload_reflections(lbm_ctx* ctx, int32_t count, const int32_t* node,
                           const uint32_t* dirs, const float* weights);

/* ------------------------------------------------------------ flow */
/* Lattice::setFlowData (Lattice.cpp:428-436) with FlowData
 * (simulation/include/FlowSurface.h:9-46): params has up to 6 ints,
 * induced 27 floats, mask N bytes for LBM_FLOW_MASK (else NULL). */
int lbm_set_flow(lbm_ctx* ctx, int kind, const int32_t* params,
                 const float* induced, const uint8_t* mask);

/* ------------------------------------------------------------ stages */
/* RunCudaFunctions::run_calculate_streaming (SimCalcFuncs.cu:220-223),
 * as driven by Lattice::simulateStreaming (Lattice.cpp:115-154). */
int lbm_stream(lbm_ctx* ctx);
/* run_calculate_collision (SimCalcFuncs.cu:225-228), with the given tau. */
int lbm_collide(lbm_ctx* ctx, float tau);
/* run_calculate_reflections / run_calculate_reflections_data
 * (SimCalcFuncs.cu:230-238) as driven by Lattice::simulateReflections
 * (Lattice.cpp:163-210); with_force accumulates into the force register. */
int lbm_reflect(lbm_ctx* ctx, int with_force);
/* run_generate_flow (SimCalcFuncs.cu:245-248) / Lattice::simulateFlow. */
int lbm_flow(lbm_ctx* ctx);
/* run_prime_points (SimCalcFuncs.cu:240-243): coordinates are implied by
 * the index in this layout, so this only validates the context. */
int lbm_prime_points(lbm_ctx* ctx);
/* Lattice::setReflectionData / retrieve_reflection_data
 * (Lattice.cpp:61-71,438-443): the device force accumulator. */
int lbm_set_force(lbm_ctx* ctx, const float in[3]);
int lbm_get_force(lbm_ctx* ctx, float out[3]);

/* ------------------------------------------------------------ stepping */
/* n fused steps.  Standard mode: BGK pull collide+stream (D2Q9 lattices
 * with 16-bit storage larger than the Infinity Cache are updated 4 or 2
 * nodes per thread; the environment variable LBM_GROUP_STEP=0 / 1 turns
 * that off / on for any size -- the results are the same bytes).
 * Reference mode: Lattice::simulateLattice (Lattice.cpp:222-226), stream
 * then collide. */
int lbm_step(lbm_ctx* ctx, int n);
/* n steps of each of count standard-mode lattices, one lbm_step launch
 * per lattice per step (the unbatched way to step many lattices; no
 * counterpart in the reference). */
int lbm_step_each(lbm_ctx* const* ctxs, int32_t count, int n);
/* Reference mode: n steps of the LatticeBoltzmannCuda.cpp:85-93 loop
 * (stream, reflect, collide, flow). */
int lbm_step_reference(lbm_ctx* ctx, int n, int with_force);
/* hipEvent-timed n steps (lbm_step) on the context's stream. */
int lbm_time_steps(lbm_ctx* ctx, int n, float* ms);
/* hipEvent-timed n reference-mode full steps. */
int lbm_time_steps_reference(lbm_ctx* ctx, int n, int with_force, float* ms);

/* ------------------------------------------------------------ readback */
/* rho (N floats) and u (3 planes of N floats), either may be NULL.
 * Reference mode uses the SPLIT velocity of SimCalcFuncs.cu:102-120;
 * standard mode u = (sum f e - F/2)/rho of the stored state; solid nodes
 * read rho = 0, u = 0. */
int lbm_moments(lbm_ctx* ctx, float* rho, float* u);
/* FP64 totals over non-solid nodes: out[0] = mass, out[1..3] = momentum
 * sum f e. */
int lbm_conservation(lbm_ctx* ctx, double out[4]);
/* No counterpart in the reference; for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  lbm_moments without any allocation:
 * the same kernels and results, staged in a fixed 64 MiB device array of
 * the library (one per GPU, held until the process ends) instead of a
 * buffer it allocates.  Lattices needing more than that -- 16 bytes per
 * node for rho and u (4,194,304 nodes), 12 for u alone, 4 for rho alone --
 * are refused with LBM_ERR_INVALID.  Calls on one GPU share the array and
 * are serialised by a lock held until the readback has completed. */
int lbm_moments_noalloc(lbm_ctx* ctx, float* rho, float* u);

/* ------------------------------------------------------------ test helper */
/* Replaces RunCudaTestFunctions::run_test_coordinates
 * (test/CudaTestHelper.cu:12-31): a device kernel checks that the launch
 * mapping visits every node once and that index <-> (x,y,z) round-trips. */
int lbm_test_coordinates(lbm_ctx* ctx, int* ok);

/* ------------------------------------------------------------ batches */
/* No counterpart in the reference (it steps one lattice with one launch
 * per stage).  A batch holds many standard-mode lattices of one storage
 * type (D2Q9 / D3Q19 / D3Q27 may be mixed) that stay resident on the GPU;
 * one persistent kernel per step updates all of them, and chunks of steps
 * can be replayed from a captured hipGraph.  Admission is refused with
 * LBM_ERR_BUDGET when the batch's device memory would exceed its budget. */
typedef struct lbm_batch lbm_batch;
int lbm_batch_create(int storage, int64_t budget_bytes, void* stream,
                     lbm_batch** out);
int lbm_batch_destroy(lbm_batch* b);
/* device bytes lbm_batch_add would reserve for cfg */
int64_t lbm_batch_estimate_bytes(const lbm_config* cfg, int track_residual);
/* Adds a lattice (solid may be NULL), initialised to rho = 1, u = 0;
 * track_residual keeps a velocity snapshot for lbm_batch_residuals. */
int lbm_batch_add(lbm_batch* b, const lbm_config* cfg, const uint8_t* solid,
                  int track_residual, int32_t* slot);
int lbm_batch_remove(lbm_batch* b, int32_t slot);
/* BGK cell Reynolds number U / nu of a standard-mode configuration, nu =
 * (tau - 1/2) / 3 in lattice units, U the larger of the moving wall's
 * speed |u_w| and, with a body force F, the centre-line speed of a
 * Poiseuille channel across the narrowest walled extent W, |F| W^2 /
 * (8 nu) (W = the largest extent when no axis has walls).  BGK flows
 * driven at U ~ 0.1 turn unstable as U / nu approaches ~17 (lid-driven
 * cavities, measured); a negative value means an invalid cfg.  No
 * counterpart in the reference, which has no stability check. */
double lbm_cell_reynolds(const lbm_config* cfg);
/* Admission limit of a batch: lbm_batch_add refuses a lattice whose
 * lbm_cell_reynolds exceeds max_cell_re with LBM_ERR_UNSTABLE; 0 (the
 * default) admits every lattice. */
int lbm_batch_set_stability_limit(lbm_batch* b, double max_cell_re);
/* The member's context, for the per-lattice calls (init, populations,
 * moments, conservation).  Do not lbm_step or lbm_destroy a member. */
int lbm_batch_member(lbm_batch* b, int32_t slot, lbm_ctx** ctx);
int lbm_batch_info(const lbm_batch* b, int32_t* lattices, int64_t* nodes,
                   int64_t* device_bytes);
/* Launch shape: block in {64,128,256,512} nodes per tile; grid = number of
 * persistent workgroups, 0 = one workgroup per tile. */
int lbm_batch_set_launch(lbm_batch* b, int block, int grid);
int lbm_batch_get_launch(const lbm_batch* b, int* block, int* grid);
/* Times every (block, grid) candidate over `steps` steps of the resident
 * set and keeps the fastest; results are cached per (storage, lattice mix,
 * log2 nodes) when use_cache.  The members advance by the steps run. */
int lbm_batch_autotune(lbm_batch* b, int steps, int use_cache, int* block,
                       int* grid);
/* n steps of every member: plain launches, or full chunks of `chunk`
 * (even) steps replayed from a captured hipGraph plus plain launches for
 * the rest. */
int lbm_batch_step(lbm_batch* b, int n);
int lbm_batch_step_graph(lbm_batch* b, int n, int chunk);
int lbm_batch_synchronize(lbm_batch* b);
/* hipEvent-timed n steps (chunk 0: plain launches). */
int lbm_batch_time_steps(lbm_batch* b, int n, int chunk, float* ms);
/* out[slot] = |u - u_prev|_2 / |u|_2 over fluid nodes of every tracked
 * member since its previous call (the snapshot is then replaced); -1 for
 * empty or untracked slots; NaN for a member that diverged: a non-finite
 * velocity at any fluid node, or a speed above LBM_DIVERGED_SPEED.
 * nslots >= the number of slots. */
int lbm_batch_residuals(lbm_batch* b, double* out, int32_t nslots);

/* ------------------------------------------------------------ serving */
/* No counterpart in the reference.  Continuous batching of solve requests
 * on one GPU: requests arrive at arrival_s (seconds after the call
 * starts), wait in FIFO order, are admitted into one batch while the HBM
 * budget (population buffers, geometry, residual snapshot) and resident
 * limit allow, advance together in chunks of `chunk` steps, and retire
 * when their residual |u - u_prev|_2 / |u|_2 (over one chunk) is below tol
 * or after max_steps (rounded up to whole chunks).  The final velocity (3
 * planes of N floats) is copied to out_u (or a scratch buffer) before
 * t_done_s is taken.  status: 0 converged, 1 step limit, 2 rejected (does
 * not fit the budget), 3 diverged (retired at the chunk where its velocity
 * turned non-finite or faster than LBM_DIVERGED_SPEED; the other
 * requests are unaffected), 4 rejected as unstable (cell Reynolds number
 * above the server's max_cell_re; not run). */
typedef struct lbm_request {
  lbm_config cfg;       /* standard mode, storage = the server's */
  const uint8_t* solid; /* one byte per node, or NULL */
  double arrival_s;
  int32_t max_steps;
  double tol;
  float* out_u;         /* 3*N floats, or NULL */
  double t_admit_s, t_done_s; /* outputs, seconds after the start */
  int32_t steps;
  double residual;
  int32_t status;
} lbm_request;

typedef struct lbm_serve_params {
  int32_t storage;
  int64_t budget_bytes;
  int32_t chunk;        /* even */
  int32_t max_resident; /* 0 = no limit */
  int32_t autotune;     /* pick the launch shape on a scratch batch first */
  int32_t engine;       /* LBM_ENGINE_BATCHED or LBM_ENGINE_PER_LATTICE */
  double progress_s;    /* > 0: a status line on stderr this often */
  int64_t max_resident_nodes; /* 0 = no limit; otherwise admit only while
                                 the resident lattices hold fewer nodes
                                 (a lattice alone is always admitted) */
  double max_cell_re;   /* > 0: requests whose lbm_cell_reynolds exceeds
                           it are rejected with status 4; 0 = no limit */
} lbm_serve_params;

/* how a server advances its resident lattices by one chunk */
enum {
  LBM_ENGINE_BATCHED = 0,     /* one persistent launch per step, hipGraph chunks */
  LBM_ENGINE_PER_LATTICE = 1  /* one plain launch per lattice per step */
};

typedef struct lbm_serve_stats {
  double wall_s;            /* first admission to last result */
  int64_t lattice_updates;  /* sum over chunks of resident nodes x chunk */
  double mlups;             /* lattice_updates / wall_s / 1e6 */
  double p50_latency_s, p99_latency_s, mean_latency_s;
  int32_t completed, rejected, max_resident;
  int64_t peak_bytes;
  int32_t unstable;         /* rejected by max_cell_re (not in rejected) */
  int32_t diverged;         /* retired with status 3 (in completed) */
  double population_bytes;  /* algorithmic HBM bytes of the steps: sum
                               over chunks and resident lattices of nodes
                               x 2 Q x storage bytes x chunk */
} lbm_serve_stats;

int lbm_serve_run(const lbm_serve_params* params, lbm_request* requests,
                  int32_t count, lbm_serve_stats* stats);

/* ------------------------------------------------------------ geometry */
/* Marks the nodes inside a closed triangle mesh, on the GPU: inside[n] = 1
 * when node (x, y, z), at position (x, y, z) * spacing (formed in float),
 * lies on a face or has generalized winding number |W| > 1/2 -- the
 * decisions of the host loop it replaces (Lattice::insertModel,
 * Lattice.cpp:228-260, over Model::isPointInsideModel, Model.cpp:51-63).
 * vertices: 3 floats each; faces: 3 vertex indices each; inside: one byte
 * per node on the host; count (may be NULL): the number of inside nodes. */
int lbm_voxelize_mesh(lbm_ctx* ctx, const float* vertices, int32_t nv,
                      const int32_t* faces, int32_t nf, float spacing,
                      uint8_t* inside, int64_t* count);

/* The reflection tables of preProcessModel (Lattice.cpp:323-426, 450-477)
 * built on the GPU from the mesh and the context's current flags
 * (LBM_FLAG_INTERNAL as uploaded): one record per internal node with at
 * least one eligible link, in node order -- node[k], 14 packed direction
 * words dirs[14k..], 81 weights weights[81k..] -- the values the host
 * builder gives.  count receives the number of records; it is an error
 * (with count set) when that exceeds max_records. */
int lbm_build_reflections(lbm_ctx* ctx, const float* vertices, int32_t nv,
                          const int32_t* faces, int32_t nf, float spacing,
                          int32_t max_records, int32_t* node, uint32_t* dirs,
                          float* weights, int32_t* count);

/* ------------------------------------------------------------ snapshots */
/* No counterpart in the reference (its state leaves the GPU only as the
 * 144-byte record stream of retrieve_data / load_data, Lattice.cpp:73-113):
 * writes everything needed to resume -- configuration, step count, the
 * current populations bit for bit, the standard mode's link mask and class
 * map or the reference mode's flags, reflection records, flow injection
 * and force register -- to one binary file. */
int lbm_snapshot_save(lbm_ctx* ctx, const char* path);
/* Restores a snapshot into a context created with the same extents,
 * lattice, storage and mode; tau, body force, moving wall, periodicity,
 * flow and force come from the snapshot.  Nothing changes on error. */
int lbm_snapshot_load(lbm_ctx* ctx, const char* path);

/* ------------------------------------------------------------ measurement */
/* No counterpart in the reference: copies a `bytes`-sized device buffer
 * into another with 16-byte accesses, iters times, and reports the
 * read+write bandwidth in GB/s -- the practical HBM ceiling of the GPU,
 * used by the benchmark beside the vendor peak. */
int lbm_measure_copy_bandwidth(int64_t bytes, int iters, double* gbs);

/* ------------------------------------------------------------ setup */
/* Random spheres of the channel benchmark: draws (cx, cy, cz, r) from
 * MT19937-64(seed) as u = (x >> 11) * 2^-53, c = u * n_axis,
 * r = rmin + (rmax - rmin) u, until at least frac of the nodes are solid;
 * node (x,y,z) is solid when |(x,y,z) - c|^2 <= r^2 (x distance periodic
 * when periodic_x).  Returns the sphere count (< 0 on error); writes up to
 * max_spheres*4 doubles when spheres != NULL. */
int64_t lbm_random_spheres(int32_t nx, int32_t ny, int32_t nz, uint64_t seed,
                           double frac, double rmin, double rmax,
                           int32_t periodic_x, uint8_t* solid,
                           int64_t max_spheres, double* spheres);

#ifdef __cplusplus
}
#endif
#endif /* LBM_H */
