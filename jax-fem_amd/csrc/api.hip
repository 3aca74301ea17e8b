// C-ABI front for the cell kernels; error plumbing.
#include "common.hpp"
#include "hipfem.h"

namespace hf {
static thread_local char g_err[1024] = {0};
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* get_error() { return g_err; }

int hex8_box_geometry(const double*, const int32_t*, void*, hipStream_t);
int hex8_residual_box_f64(const hf_material*, const void*, const int32_t*, int64_t, const double*, const double*,
                          double*, hipStream_t);
int hex8_residual_f64(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                      double*, hipStream_t);
int hex8_residual_f32(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                      float*, hipStream_t);
int hex8_jacobian_f64(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                      double*, hipStream_t);
int hex8_jacobian_f32(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                      float*, hipStream_t);
int hex8_jacobian_upper_f32(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                            float*, hipStream_t);
int hex8_param_vjp(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                   const double*, double*, hipStream_t);
int hex27_residual_f64(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                       double*, hipStream_t);
int hex27_residual_f32(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                       float*, hipStream_t);
int hex27_jacobian_f64(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                       double*, hipStream_t);
int hex27_jacobian_f32(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                       float*, hipStream_t);
int hex27_param_vjp(const hf_material*, const double*, const int32_t*, int64_t, const double*, const double*,
                    const double*, double*, hipStream_t);
}  // namespace hf

using namespace hf;

extern "C" int hf_last_error(char* buf, size_t n) {
  const size_t len = strlen(g_err);
  if (buf && n) {
    strncpy(buf, g_err, n - 1);
    buf[n - 1] = 0;
  }
  return (int)len;
}

extern "C" const char* hf_version(void) { return "libhipfem 0.1 (gfx950)"; }

static int check_args(int ele, const hf_material* m, const double* p, const int32_t* c, const double* u,
                      const void* out) {
  HF_REQUIRE(m && p && c && u && out, "cell kernel: null argument");
  HF_REQUIRE(ele == 8 || ele == 27, "cell kernel: element with %d nodes not supported", ele);
  HF_REQUIRE(m->kind >= 0 && m->kind <= 2, "cell kernel: unknown material %d", m->kind);
  HF_REQUIRE(!(m->kind == HF_MAT_POISSON && ele == 27), "cell kernel: HEX27 Poisson not supported");
  if (ele == 8)
    HF_REQUIRE(m->quad1d == 0 || m->quad1d == 2, "cell kernel: HEX8 kernels use 2x2x2 Gauss points "
               "(quadrature_order 2 or 3); got %d per direction", m->quad1d);
  else
    HF_REQUIRE(m->quad1d == 3, "cell kernel: HEX27 kernels use 3x3x3 Gauss points (quadrature_order 4 or 5); "
               "got %d per direction", m->quad1d);
  return 0;
}

extern "C" int hf_cell_residual_f64(int32_t ele, const hf_material* m, const double* p, const int32_t* c,
                                    int64_t n, const double* u, const double* prm, double* o, hf_stream_t s) {
  if (int e = check_args(ele, m, p, c, u, o)) return e;
  return ele == 8 ? hex8_residual_f64(m, p, c, n, u, prm, o, (hipStream_t)s)
                  : hex27_residual_f64(m, p, c, n, u, prm, o, (hipStream_t)s);
}
extern "C" size_t hf_hex8_box_geometry_bytes(void) { return (8 * 8 * 3 + 8) * sizeof(double); }
extern "C" int hf_hex8_box_geometry(const double* p, const int32_t* c, int64_t n, double* geo, hf_stream_t s) {
  HF_REQUIRE(p && c && geo, "hf_hex8_box_geometry: null argument");
  HF_REQUIRE(n > 0, "hf_hex8_box_geometry: no cells");
  return hex8_box_geometry(p, c, geo, (hipStream_t)s);
}
extern "C" int hf_cell_residual_box_f64(const hf_material* m, const double* geo, const int32_t* c, int64_t n,
                                        const double* u, const double* prm, double* o, hf_stream_t s) {
  HF_REQUIRE(m && geo && c && u && o, "hf_cell_residual_box_f64: null argument");
  HF_REQUIRE(((uintptr_t)c & 15) == 0, "hf_cell_residual_box_f64: cells not 16-byte aligned");
  HF_REQUIRE(m->kind >= 0 && m->kind <= 2, "hf_cell_residual_box_f64: unknown material %d", m->kind);
  HF_REQUIRE(m->quad1d == 0 || m->quad1d == 2, "hf_cell_residual_box_f64: HEX8 kernels use 2x2x2 Gauss points; "
             "got %d per direction", m->quad1d);
  return hex8_residual_box_f64(m, geo, c, n, u, prm, o, (hipStream_t)s);
}
extern "C" int hf_cell_residual_f32(int32_t ele, const hf_material* m, const double* p, const int32_t* c,
                                    int64_t n, const double* u, const double* prm, float* o, hf_stream_t s) {
  if (int e = check_args(ele, m, p, c, u, o)) return e;
  return ele == 8 ? hex8_residual_f32(m, p, c, n, u, prm, o, (hipStream_t)s)
                  : hex27_residual_f32(m, p, c, n, u, prm, o, (hipStream_t)s);
}
extern "C" int hf_cell_jacobian_f64(int32_t ele, const hf_material* m, const double* p, const int32_t* c,
                                    int64_t n, const double* u, const double* prm, double* o, hf_stream_t s) {
  if (int e = check_args(ele, m, p, c, u, o)) return e;
  return ele == 8 ? hex8_jacobian_f64(m, p, c, n, u, prm, o, (hipStream_t)s)
                  : hex27_jacobian_f64(m, p, c, n, u, prm, o, (hipStream_t)s);
}
extern "C" int hf_cell_jacobian_f32(int32_t ele, const hf_material* m, const double* p, const int32_t* c,
                                    int64_t n, const double* u, const double* prm, float* o, hf_stream_t s) {
  if (int e = check_args(ele, m, p, c, u, o)) return e;
  return ele == 8 ? hex8_jacobian_f32(m, p, c, n, u, prm, o, (hipStream_t)s)
                  : hex27_jacobian_f32(m, p, c, n, u, prm, o, (hipStream_t)s);
}
extern "C" int hf_hex8_jacobian_upper_f32(const hf_material* m, const double* p, const int32_t* c, int64_t n,
                                          const double* u, const double* prm, float* o, hf_stream_t s) {
  if (int e = check_args(8, m, p, c, u, o)) return e;
  return hex8_jacobian_upper_f32(m, p, c, n, u, prm, o, (hipStream_t)s);
}
extern "C" int hf_cell_param_vjp_f64(int32_t ele, const hf_material* m, const double* p, const int32_t* c,
                                     int64_t n, const double* u, const double* prm, const double* lam,
                                     double* grad, hf_stream_t s) {
  if (int e = check_args(ele, m, p, c, u, grad)) return e;
  HF_REQUIRE(prm && lam, "hf_cell_param_vjp_f64: params and lam are required");
  return ele == 8 ? hex8_param_vjp(m, p, c, n, u, prm, lam, grad, (hipStream_t)s)
                  : hex27_param_vjp(m, p, c, n, u, prm, lam, grad, (hipStream_t)s);
}
