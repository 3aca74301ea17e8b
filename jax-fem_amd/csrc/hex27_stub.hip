// HEX27 kernels: placeholder entry points until the MFMA kernels land.
#include "common.hpp"
#include "hipfem.h"
namespace hf {
#define HEX27_TODO(name, OUT)                                                                         \
  int name(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*, \
           OUT*, hipStream_t) {                                                                       \
    set_error(#name ": HEX27 kernels are not built in this library version");                       \
    return -1;                                                                                        \
  }
HEX27_TODO(hex27_residual_f64, double)
HEX27_TODO(hex27_residual_f32, float)
HEX27_TODO(hex27_jacobian_f64, double)
HEX27_TODO(hex27_jacobian_f32, float)
int hex27_param_vjp(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                    const double*, double*, hipStream_t) {
  set_error("hex27_param_vjp: HEX27 kernels are not built in this library version");
  return -1;
}
}  // namespace hf
