// Shared helpers for libhipfem (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdarg>
#include <string>

namespace hf {

// Per-thread last error text; read through hf_last_error().
void set_error(const char* fmt, ...);
const char* get_error();

#define HF_CHECK_HIP(expr)                                                        \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      ::hf::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,                \
                      hipGetErrorString(_e));                                     \
      return -2;                                                                  \
    }                                                                             \
  } while (0)

#define HF_REQUIRE(cond, ...)                                                     \
  do {                                                                            \
    if (!(cond)) {                                                                \
      ::hf::set_error(__VA_ARGS__);                                               \
      return -1;                                                                  \
    }                                                                             \
  } while (0)

#define HF_LAUNCH_CHECK()                                                         \
  do {                                                                            \
    hipError_t _e = hipGetLastError();                                            \
    if (_e != hipSuccess) {                                                       \
      ::hf::set_error("%s:%d launch -> %s", __FILE__, __LINE__,                   \
                      hipGetErrorString(_e));                                     \
      return -2;                                                                  \
    }                                                                             \
  } while (0)

constexpr int kWave = 64;

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Wave64 sum of a double (DPP-free butterfly through shuffles).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Material ids understood by the cell kernels (mirrors include/hipfem.h).
enum MaterialKind : int { MAT_POISSON = 0, MAT_LINEAR_ELASTIC = 1, MAT_NEO_HOOKEAN = 2 };
// How the per-quadrature parameter p enters the moduli.
enum ParamMode : int { P_NONE = 0, P_SCALE_E = 1, P_SIMP = 2, P_MU = 3, P_SCALE_K = 4 };

}  // namespace hf
