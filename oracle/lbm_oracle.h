/*
 * lbm_oracle.h -- CPU restatement of the reference lattice-Boltzmann path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  Nothing in the product
 * (latticeboltzmanncuda_amd/, include/) links, loads or calls it.
 *
 * Three groups of functions:
 *
 *  1. Reference mode (D3Q27, node index n = z + nz*(y + ny*x), float SoA
 *     populations f[q*N + n]).  Each function restates one reference kernel
 *     or host routine, cited by file:line of CharlieNicholls/LatticeBoltzmannCuda:
 *       orc_ref_stream      simulation/source/SimCalcFuncs.cu:79-96
 *       orc_ref_collide     simulation/source/SimCalcFuncs.cu:98-156
 *       orc_ref_reflect     simulation/source/SimCalcFuncs.cu:158-198,
 *                           simulation/source/Lattice.cpp:163-210
 *       orc_ref_flow        simulation/source/SimCalcFuncs.cu:200-215,
 *                           simulation/source/FlowCriterion.cu:5-16
 *       orc_ref_run_main    LatticeBoltzmannCuda.cpp:85-93 (S,R,C,F order)
 *       orc_ref_run_lattice simulation/source/Lattice.cpp:222-226 (S then C)
 *   The force accumulator is kept twice: in float with one addition per
 *   term (the reference's atomicAdd semantics, SimCalcFuncs.cu:181-183) and
 *   in FP64 per pass (the exact sum, which the GPU register tracks).
 *
 *  2. Geometry (host setup of the reference):
 *       orc_mesh_*              simulation/source/Model.cpp:22-91 (CGAL
 *                               semantics restated without CGAL: boundary
 *                               counts as inside, first ray hit, face normal)
 *       orc_rotate_vector       simulation/source/Geometry.cpp:22-42
 *       orc_distribute_vector   simulation/source/Lattice.cpp:262-321
 *       orc_insert_model        simulation/source/Lattice.cpp:228-260
 *       orc_preprocess_model    simulation/source/Lattice.cpp:323-426,450-477
 *       orc_compress_ints       simulation/source/Utils.cpp:3-28
 *
 *  3. Standard mode (no reference counterpart; defined by the build's
 *     benchmark configurations): D2Q9/D3Q19/D3Q27 BGK pull with half-way
 *     bounce-back, one optional moving wall, periodic axes and Guo forcing,
 *     in float, double, and float compute with shifted FP16/BF16 storage;
 *     MT19937-64 random-sphere obstacles.
 *
 * Parity pins (what fixes each function; the tests live in tests/):
 *   orc_ref_stream/collide  CalculationsTest.cpp:19-105 golden values, mass
 *                           invariants, brute-force permutation count.
 *   orc_ref_reflect         CalculationsTest.cpp:107-197 golden values and
 *                           the closed-form force.  The force term for slots
 *                           s >= 27 reads past the reference's tables
 *                           (SimCalcFuncs.cu:181-183): it is DEFINED here as 0
 *                           and is "parity unpinned".
 *   orc_ref_flow            CalculationsTest.cpp:199-277.
 *   orc_mesh_*              ModelTest.cpp:5-74, CoordinatesTest.cpp:137-169;
 *                           points within 1e-6 of a face and rays through
 *                           mesh edges are "parity unpinned" (CGAL float
 *                           kernel decisions).
 *   orc_distribute_vector,  CoordinatesTest.cpp:171-229.
 *   orc_preprocess_model
 *   orc_rotate_vector       GeometryTest.cpp:11-32.
 *   orc_compress_ints       CalculationsTest.cpp:279-284.
 *   standard mode           analytic Poiseuille / Taylor-Green decay,
 *                           mass and momentum conservation, equilibrium
 *                           fixed point, f32 vs f64 agreement.
 *   orc_mt64                std::mt19937_64 10000th output (C++ [rand.predef]).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- tables ---------------- */
/* D3Q27 velocity set and inverse map (simulation/source/Constants.cpp:6-7). */
extern const int ORC_E27[27][3];
extern const int ORC_INV27[27];
/* 1/|e|^2 (simulation/source/SimCalcFuncs.cu:14). */
extern const float ORC_SPLIT27[27];

void orc_set_threads(int n);
int orc_get_threads(void);

/* ---------------- reference mode ---------------- */
/* flag bits of the per-node flag byte */
#define ORC_FLAG_INTERNAL 1u
#define ORC_FLAG_REFLECTED 2u

/* flow criterion kinds */
#define ORC_FLOW_NONE 0
#define ORC_FLOW_POINT 1 /* params: x, y, z            */
#define ORC_FLOW_PLANE 2 /* params: axis, coordinate   */
#define ORC_FLOW_BOX 3   /* params: x0,y0,z0,x1,y1,z1 (inclusive) */

void orc_ref_stream(int nx, int ny, int nz, const float* L, float* T);
void orc_ref_collide(int nx, int ny, int nz, float* f, float tau);
void orc_ref_reflect(int nx, int ny, int nz, const float* L, float* T,
                     const uint8_t* flags, int nrec, const int32_t* rec_node,
                     const uint32_t* rec_dirs, const float* rec_w,
                     float* force /* nullable, 3 floats, accumulated */,
                     double* force64 /* nullable, FP64 sum of the pass */);
void orc_ref_flow(int nx, int ny, int nz, float* f, int kind,
                  const int32_t* params, const float* induced);
/* Whole time loop of LatticeBoltzmannCuda.cpp:85-93: per step stream,
 * reflect (with force when force != NULL), collide, flow. */
void orc_ref_run_main(int nx, int ny, int nz, float* f, const uint8_t* flags,
                      int nrec, const int32_t* rec_node,
                      const uint32_t* rec_dirs, const float* rec_w, float tau,
                      int flow_kind, const int32_t* flow_params,
                      const float* induced, float* force, double* force64,
                      int steps);
/* Lattice::simulateLattice repeated: stream then collide. */
void orc_ref_run_lattice(int nx, int ny, int nz, float* f, float tau,
                         int steps);
/* rho and the SPLIT velocity of SimCalcFuncs.cu:102-120; u is 3 planes. */
void orc_ref_moments(int nx, int ny, int nz, const float* f, float* rho,
                     float* u);

/* ---------------- geometry ---------------- */
typedef struct orc_mesh orc_mesh;
orc_mesh* orc_mesh_load(const char* path); /* NULL if the file is missing */
void orc_mesh_free(orc_mesh* m);
int orc_mesh_nverts(const orc_mesh* m);
int orc_mesh_nfaces(const orc_mesh* m);
int orc_mesh_valid(const orc_mesh* m);  /* triangle mesh and closed */
int orc_mesh_closed(const orc_mesh* m);
void orc_mesh_bbox(const orc_mesh* m, float out[6]);
int orc_mesh_inside(const orc_mesh* m, float x, float y, float z);
void orc_mesh_reflection_vector(const orc_mesh* m, const float p1[3],
                                const float p2[3], float out[3]);
void orc_rotate_vector(const float axis[3], const float v[3], float angle,
                       float out[3]);
void orc_distribute_vector(const float v[3], float score[27], int32_t idx[27]);
/* returns 1 when flags were written, 0 when the model lies outside the box
 * (the reference silently ignores it, Lattice.cpp:236-244). */
int orc_insert_model(const orc_mesh* m, int nx, int ny, int nz, float spacing,
                     uint8_t* internal);
/* Builds reflection records (node index, 14 packed direction words, 81
 * weights) for every internal node with a fluid neighbour.  Sets
 * ORC_FLAG_REFLECTED in flags.  Returns the record count or -1 when
 * max_rec is too small. */
int orc_preprocess_model(const orc_mesh* m, int nx, int ny, int nz,
                         float spacing, uint8_t* flags, int max_rec,
                         int32_t* rec_node, uint32_t* rec_dirs, float* rec_w);
unsigned orc_compress_ints(const int* a, int num, int bits);

/* ---------------- standard mode ---------------- */
#define ORC_MASK_SOLID (1u << 31)
#define ORC_MASK_LID (1u << 30)

/* lattice: 9, 19 or 27.  2-D lattices use nz == 1 and ignore z. */
int orc_std_q_valid(int q);
void orc_std_tables(int q, int32_t* e /* q*3 */, int32_t* inv /* q */,
                    double* w /* q */);
/* periodic[a] != 0 wraps axis a, otherwise links leaving the box bounce
 * back half-way.  lid_face: -1 none, else 2*axis + (0 for min, 1 for max). */
void orc_std_build_mask(int q, int nx, int ny, int nz, const uint8_t* solid,
                        const int32_t periodic[3], int lid_face,
                        uint32_t* mask);
typedef struct orc_std_params {
  float tau;
  float force[3];
  int32_t lid_face;
  float lid_u[3];
  float wall_rho;
  int32_t periodic[3];
} orc_std_params;
void orc_std_init_eq_f64(int q, int nx, int ny, int nz, const double* rho,
                         const double* u, double* f);
/* storage: 0 float f_i, computed on f_i; 2 fp16 shifted, 3 bf16 shifted,
 * 4 float shifted: values f_i - w_i, computed in the shifted form and
 * rounded to the storage type after every step (the GPU library's
 * standard-mode FP32 is 4).  fA is in/out (float).  Solid nodes are set to
 * the rest state w_i by every step. */
void orc_std_run_f32(int q, int nx, int ny, int nz, float* fA,
                     const uint32_t* mask, const orc_std_params* p, int steps,
                     int storage);
void orc_std_run_f64(int q, int nx, int ny, int nz, double* fA,
                     const uint32_t* mask, const orc_std_params* p, int steps);
/* rho and the Guo velocity u = (sum f e - F/2)/rho of the stored
 * post-collision state; solid nodes give rho = 0, u = 0. */
void orc_std_moments_f64(int q, int nx, int ny, int nz, const double* f,
                         const uint32_t* mask, const float force[3],
                         double* rho, double* u);
/* float <-> storage rounding used by the shifted storage emulation */
float orc_round_fp16(float x);
float orc_round_bf16(float x);
/* the stored (unscaled) 16-bit value of the shifted deviation v of
 * direction `direction` with stochastic rounding under the node hash
 * `hash`, storage 2 = FP16 (times 2^12), 3 = BF16 -- as forced lattices
 * store it; orc_sr_mul is the direction's multiplier in the rounding hash */
float orc_store16_sr(int storage, float v, int direction, uint32_t hash);
uint32_t orc_sr_mul(int direction);

/* ---------------- random spheres ---------------- */
uint64_t orc_mt64_nth(uint64_t seed, int64_t n); /* n-th output, 1-based */
/* Adds spheres drawn from MT19937-64(seed) until the solid fraction is at
 * least frac.  Draw order per sphere: cx, cy, cz, r with u = (x>>11)*2^-53,
 * c = u*n_axis, r = rmin + (rmax-rmin)*u.  Node (x,y,z) is solid when
 * dx^2+dy^2+dz^2 <= r^2 (dx periodic when periodic_x).  Writes up to
 * max_spheres*4 doubles (cx,cy,cz,r) and returns the sphere count. */
int64_t orc_random_spheres(int nx, int ny, int nz, uint64_t seed, double frac,
                           double rmin, double rmax, int periodic_x,
                           uint8_t* solid, int64_t max_spheres,
                           double* spheres);

#ifdef __cplusplus
}
#endif
#endif
