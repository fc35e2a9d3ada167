// The C-linkage exports of include/lbm.h.  Each entry point's body lives in
// namespace lbm::api (lbm_capi.hip, lbm_batch.hip, lbm_serve.hip,
// lbm_snapshot.hip, k_voxel.hip); the export below calls it inside a
// try block, so no C++ exception (a std::vector or std::string that cannot
// allocate, for example) crosses the C-ABI: it becomes a status code --
// LBM_ERR_NOMEM for std::bad_alloc, LBM_ERR_INVALID otherwise -- with the
// message in lbm_last_error(), and a negative status for the entry points
// that return a count.
#include "lbm_internal.h"

// X(return type, name, parameter list, argument list)
#define LBM_API_LIST(X)                                                                         \
  X(int, lbm_device_info, (int* a, char* b, size_t c), (a, b, c))                              \
  X(int, lbm_create, (const lbm_config* a, void* b, lbm_ctx** c), (a, b, c))                   \
  X(int, lbm_destroy, (lbm_ctx* a), (a))                                                       \
  X(int, lbm_get_config, (const lbm_ctx* a, lbm_config* b), (a, b))                            \
  X(int64_t, lbm_node_count, (const lbm_ctx* a), (a))                                          \
  X(int64_t, lbm_device_bytes, (const lbm_ctx* a), (a))                                        \
  X(int, lbm_set_tau, (lbm_ctx* a, float b), (a, b))                                           \
  X(int, lbm_set_body_force, (lbm_ctx* a, const float b[3]), (a, b))                           \
  X(int, lbm_synchronize, (lbm_ctx* a), (a))                                                   \
  X(int, lbm_upload_points, (lbm_ctx* a, const lbm_lattice_point* b, int64_t c), (a, b, c))    \
  X(int, lbm_download_points, (lbm_ctx* a, lbm_lattice_point* b, int64_t c), (a, b, c))        \
  X(int, lbm_set_populations_f32, (lbm_ctx* a, const float* b), (a, b))                        \
  X(int, lbm_get_populations_f32, (lbm_ctx* a, float* b), (a, b))                              \
  X(int, lbm_set_populations_f64, (lbm_ctx* a, const double* b), (a, b))                       \
  X(int, lbm_get_populations_f64, (lbm_ctx* a, double* b), (a, b))                             \
  X(int, lbm_init_equilibrium, (lbm_ctx* a, const double* b, const double* c), (a, b, c))      \
  X(int, lbm_upload_solid, (lbm_ctx* a, const uint8_t* b), (a, b))                             \
  X(int, lbm_get_mask, (lbm_ctx* a, uint32_t* b), (a, b))                                      \
  X(int, lbm_upload_flags, (lbm_ctx* a, const uint8_t* b), (a, b))                             \
  X(int, lbm_get_flags, (lbm_ctx* a, uint8_t* b), (a, b))                                      \
  X(int, lbm_upload_reflections,                                                               \
    (lbm_ctx* a, int32_t b, const int32_t* c, const uint32_t* d, const float* e),              \
    (a, b, c, d, e))                                                                           \
  X(int, lbm_set_flow,                                                                         \
    (lbm_ctx* a, int b, const int32_t* c, const float* d, const uint8_t* e), (a, b, c, d, e))  \
  X(int, lbm_stream, (lbm_ctx* a), (a))                                                        \
  X(int, lbm_collide, (lbm_ctx* a, float b), (a, b))                                           \
  X(int, lbm_reflect, (lbm_ctx* a, int b), (a, b))                                             \
  X(int, lbm_flow, (lbm_ctx* a), (a))                                                          \
  X(int, lbm_prime_points, (lbm_ctx* a), (a))                                                  \
  X(int, lbm_set_force, (lbm_ctx* a, const float b[3]), (a, b))                                \
  X(int, lbm_get_force, (lbm_ctx* a, float b[3]), (a, b))                                      \
  X(int, lbm_step, (lbm_ctx* a, int b), (a, b))                                                \
  X(int, lbm_step_each, (lbm_ctx* const* a, int32_t b, int c), (a, b, c))                      \
  X(int, lbm_step_reference, (lbm_ctx* a, int b, int c), (a, b, c))                            \
  X(int, lbm_time_steps, (lbm_ctx* a, int b, float* c), (a, b, c))                             \
  X(int, lbm_time_steps_reference, (lbm_ctx* a, int b, int c, float* d), (a, b, c, d))         \
  X(int, lbm_moments, (lbm_ctx* a, float* b, float* c), (a, b, c))                             \
  X(int, lbm_moments_noalloc, (lbm_ctx* a, float* b, float* c), (a, b, c))                     \
  X(int, lbm_conservation, (lbm_ctx* a, double b[4]), (a, b))                                  \
  X(int, lbm_test_coordinates, (lbm_ctx* a, int* b), (a, b))                                   \
  X(int, lbm_measure_copy_bandwidth, (int64_t a, int b, double* c), (a, b, c))                 \
  X(int64_t, lbm_random_spheres,                                                               \
    (int32_t a, int32_t b, int32_t c, uint64_t d, double e, double f, double g, int32_t h,     \
     uint8_t* i, int64_t j, double* k),                                                        \
    (a, b, c, d, e, f, g, h, i, j, k))                                                         \
  X(int, lbm_batch_create, (int a, int64_t b, void* c, lbm_batch** d), (a, b, c, d))           \
  X(int, lbm_batch_destroy, (lbm_batch* a), (a))                                               \
  X(int64_t, lbm_batch_estimate_bytes, (const lbm_config* a, int b), (a, b))                   \
  X(double, lbm_cell_reynolds, (const lbm_config* a), (a))                                     \
  X(int, lbm_batch_set_stability_limit, (lbm_batch* a, double b), (a, b))                      \
  X(int, lbm_batch_add,                                                                        \
    (lbm_batch* a, const lbm_config* b, const uint8_t* c, int d, int32_t* e), (a, b, c, d, e)) \
  X(int, lbm_batch_remove, (lbm_batch* a, int32_t b), (a, b))                                  \
  X(int, lbm_batch_member, (lbm_batch* a, int32_t b, lbm_ctx** c), (a, b, c))                  \
  X(int, lbm_batch_info, (const lbm_batch* a, int32_t* b, int64_t* c, int64_t* d),             \
    (a, b, c, d))                                                                              \
  X(int, lbm_batch_set_launch, (lbm_batch* a, int b, int c), (a, b, c))                        \
  X(int, lbm_batch_get_launch, (const lbm_batch* a, int* b, int* c), (a, b, c))                \
  X(int, lbm_batch_step, (lbm_batch* a, int b), (a, b))                                        \
  X(int, lbm_batch_step_graph, (lbm_batch* a, int b, int c), (a, b, c))                        \
  X(int, lbm_batch_synchronize, (lbm_batch* a), (a))                                           \
  X(int, lbm_batch_time_steps, (lbm_batch* a, int b, int c, float* d), (a, b, c, d))           \
  X(int, lbm_batch_autotune, (lbm_batch* a, int b, int c, int* d, int* e), (a, b, c, d, e))    \
  X(int, lbm_batch_residuals, (lbm_batch* a, double* b, int32_t c), (a, b, c))                 \
  X(int, lbm_serve_run,                                                                        \
    (const lbm_serve_params* a, lbm_request* b, int32_t c, lbm_serve_stats* d), (a, b, c, d))  \
  X(int, lbm_voxelize_mesh,                                                                    \
    (lbm_ctx* a, const float* b, int32_t c, const int32_t* d, int32_t e, float f, uint8_t* g,  \
     int64_t* h),                                                                              \
    (a, b, c, d, e, f, g, h))                                                                  \
  X(int, lbm_build_reflections,                                                                \
    (lbm_ctx* a, const float* b, int32_t c, const int32_t* d, int32_t e, float f, int32_t g,   \
     int32_t* h, uint32_t* i, float* j, int32_t* k),                                           \
    (a, b, c, d, e, f, g, h, i, j, k))                                                         \
  X(int, lbm_snapshot_save, (lbm_ctx* a, const char* b), (a, b))                               \
  X(int, lbm_snapshot_load, (lbm_ctx* a, const char* b), (a, b))

namespace lbm {
namespace api {
#define LBM_API_DECLARE(R, name, params, args) R name params;
LBM_API_LIST(LBM_API_DECLARE)
#undef LBM_API_DECLARE
}  // namespace api

namespace {
// status codes are positive; entry points returning a count report an
// error as the negated status
template <class R>
R error_value(int code) {
  return sizeof(R) == sizeof(int64_t) ? (R) - (int64_t)code : (R)code;
}
}  // namespace
}  // namespace lbm

extern "C" {
#define LBM_API_EXPORT(R, name, params, args) \
  R name params {                             \
    try {                                     \
      return lbm::api::name args;             \
    } catch (...) {                           \
      return lbm::error_value<R>(lbm::api_exception()); \
    }                                         \
  }
LBM_API_LIST(LBM_API_EXPORT)
#undef LBM_API_EXPORT
}  // extern "C"
