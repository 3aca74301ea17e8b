// Element-agnostic cell kernels: the reference's laplace and mass kernels
// (jax_fem/problem.py:189-236) and their element Jacobians (jacfwd of them,
// problem.py:262-266) for a constitutive map that is only known pointwise.
//
// The pointwise law is evaluated by the caller at the quadrature points (the
// stress P(u_grad, ...) and its tangent C = dP/du_grad, the mass term m(u, x, ...)
// and dm/du); these kernels do the element contractions around it:
//   u_grad[c,q,i,j] = sum_a u[cells[c,a], i] dN_a/dx_j (c, q)
//   u_q[c,q,i]      = sum_a u[cells[c,a], i] N_a(q)
//   r[c,a,i]        = sum_q JxW[c,q] (sum_j P[c,q,i,j] dN_a/dx_j + m[c,q,i] N_a(q))
//   K[c,a,b,i,k]    = sum_q JxW[c,q] (sum_{j,l} dN_a/dx_j C[c,q,i,j,k,l] dN_b/dx_l
//                                      + N_a(q) Mt[c,q,i,k] N_b(q))
// with the geometry (physical shape gradients and JxW, fe.py:112-141) given per
// (cell, quadrature point), so any element type and dimension the set-up
// tabulates runs through them.  Element residuals are (ncell, nn, vec); element
// Jacobians block-major (ncell, nn, nn, vec, vec), the layout hf_assemble reads.
#include "common.hpp"
#include "hipfem.h"

namespace hf {
namespace {

constexpr int kGenBlock = 256;

inline int gen_grid(int64_t n) {
  const int64_t g = ceil_div(n, kGenBlock);
  return (int)(g < 1 ? 1 : (g > 65535 * 8 ? 65535 * 8 : g));
}

template <int VEC, int DIM>
__global__ void __launch_bounds__(kGenBlock) ugrad_kernel(const int32_t* __restrict__ cells, int64_t ncell, int nn,
                                                          int nq, const double* __restrict__ u,
                                                          const double* __restrict__ sg,
                                                          const double* __restrict__ sv, double* __restrict__ ug,
                                                          double* __restrict__ uq) {
  const int64_t total = ncell * nq;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / nq;
    const int q = (int)(t - c * nq);
    double g[VEC][DIM] = {};
    double v[VEC] = {};
    for (int a = 0; a < nn; ++a) {
      const int64_t node = cells[c * nn + a];
      double ua[VEC];
#pragma unroll
      for (int i = 0; i < VEC; ++i) ua[i] = u[node * VEC + i];
      const double* ga = sg + ((t * nn) + a) * DIM;
#pragma unroll
      for (int i = 0; i < VEC; ++i)
#pragma unroll
        for (int j = 0; j < DIM; ++j) g[i][j] += ua[i] * ga[j];
      if (uq) {
        const double na = sv[q * nn + a];
#pragma unroll
        for (int i = 0; i < VEC; ++i) v[i] += ua[i] * na;
      }
    }
#pragma unroll
    for (int i = 0; i < VEC; ++i)
#pragma unroll
      for (int j = 0; j < DIM; ++j) ug[(t * VEC + i) * DIM + j] = g[i][j];
    if (uq)
#pragma unroll
      for (int i = 0; i < VEC; ++i) uq[t * VEC + i] = v[i];
  }
}

template <int VEC, int DIM>
__global__ void __launch_bounds__(kGenBlock) residual_kernel(int64_t ncell, int nn, int nq,
                                                             const double* __restrict__ sg,
                                                             const double* __restrict__ sv,
                                                             const double* __restrict__ jxw,
                                                             const double* __restrict__ P,
                                                             const double* __restrict__ M,
                                                             double* __restrict__ out) {
  const int64_t total = ncell * nn;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / nn;
    const int a = (int)(t - c * nn);
    double r[VEC] = {};
    for (int q = 0; q < nq; ++q) {
      const int64_t cq = c * nq + q;
      const double w = jxw[cq];
      if (P) {
        const double* ga = sg + (cq * nn + a) * DIM;
#pragma unroll
        for (int i = 0; i < VEC; ++i) {
          double s = 0.0;
#pragma unroll
          for (int j = 0; j < DIM; ++j) s += P[(cq * VEC + i) * DIM + j] * ga[j];
          r[i] += s * w;
        }
      }
      if (M) {
        const double na = sv[q * nn + a] * w;
#pragma unroll
        for (int i = 0; i < VEC; ++i) r[i] += M[cq * VEC + i] * na;
      }
    }
#pragma unroll
    for (int i = 0; i < VEC; ++i) out[t * VEC + i] = r[i];
  }
}

template <typename T, int VEC, int DIM>
__global__ void __launch_bounds__(kGenBlock) jacobian_kernel(int64_t ncell, int nn, int nq,
                                                             const double* __restrict__ sg,
                                                             const double* __restrict__ sv,
                                                             const double* __restrict__ jxw,
                                                             const double* __restrict__ C,
                                                             const double* __restrict__ Mt, T* __restrict__ out) {
  const int64_t pairs = (int64_t)nn * nn;
  const int64_t total = ncell * pairs;
  constexpr int VD = VEC * DIM;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / pairs;
    const int ab = (int)(t - c * pairs);
    const int a = ab / nn, b = ab - (ab / nn) * nn;
    double K[VEC][VEC] = {};
    for (int q = 0; q < nq; ++q) {
      const int64_t cq = c * nq + q;
      const double w = jxw[cq];
      if (C) {
        const double* ga = sg + (cq * nn + a) * DIM;
        const double* gb = sg + (cq * nn + b) * DIM;
        double gaw[DIM], gbv[DIM];
#pragma unroll
        for (int j = 0; j < DIM; ++j) {
          gaw[j] = ga[j] * w;
          gbv[j] = gb[j];
        }
        const double* Cq = C + cq * VD * VD;
#pragma unroll
        for (int i = 0; i < VEC; ++i)
#pragma unroll
          for (int k = 0; k < VEC; ++k) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < DIM; ++j)
#pragma unroll
              for (int l = 0; l < DIM; ++l) s += gaw[j] * Cq[(i * DIM + j) * VD + k * DIM + l] * gbv[l];
            K[i][k] += s;
          }
      }
      if (Mt) {
        const double nab = sv[q * nn + a] * sv[q * nn + b] * w;
#pragma unroll
        for (int i = 0; i < VEC; ++i)
#pragma unroll
          for (int k = 0; k < VEC; ++k) K[i][k] += Mt[(cq * VEC + i) * VEC + k] * nab;
      }
    }
    T* o = out + t * VEC * VEC;
#pragma unroll
    for (int i = 0; i < VEC; ++i)
#pragma unroll
      for (int k = 0; k < VEC; ++k) o[i * VEC + k] = (T)K[i][k];
  }
}

int check_shape(int vec, int dim, int nn, int nq) {
  HF_REQUIRE(vec >= 1 && vec <= 3, "generic cell kernels: vec=%d (1..3)", vec);
  HF_REQUIRE(dim == 2 || dim == 3, "generic cell kernels: dim=%d (2 or 3)", dim);
  HF_REQUIRE(nn >= 1 && nn <= 64 && nq >= 1 && nq <= 1024, "generic cell kernels: %d nodes, %d points", nn, nq);
  return 0;
}

#define GEN_DISPATCH(KERNEL_CALL)                                         \
  do {                                                                    \
    if (vec == 1 && dim == 2) { KERNEL_CALL(1, 2); }                      \
    else if (vec == 1 && dim == 3) { KERNEL_CALL(1, 3); }                 \
    else if (vec == 2 && dim == 2) { KERNEL_CALL(2, 2); }                 \
    else if (vec == 2 && dim == 3) { KERNEL_CALL(2, 3); }                 \
    else if (vec == 3 && dim == 2) { KERNEL_CALL(3, 2); }                 \
    else { KERNEL_CALL(3, 3); }                                           \
  } while (0)

template <typename T>
int jacobian_impl(int64_t ncell, int32_t nn, int32_t nq, int32_t vec, int32_t dim, const double* sg,
                  const double* sv, const double* jxw, const double* C, const double* Mt, T* out, hipStream_t s) {
  if (check_shape(vec, dim, nn, nq)) return -1;
  HF_REQUIRE(sg && jxw && out && (C || Mt) && (!Mt || sv), "hf_generic_jacobian: null argument");
  if (ncell == 0) return 0;
  const int g = gen_grid(ncell * nn * nn);
#define JAC(V, D) jacobian_kernel<T, V, D><<<g, kGenBlock, 0, s>>>(ncell, nn, nq, sg, sv, jxw, C, Mt, out)
  GEN_DISPATCH(JAC);
#undef JAC
  HF_LAUNCH_CHECK();
  return 0;
}

}  // namespace
}  // namespace hf

using namespace hf;

extern "C" int hf_generic_ugrad_f64(const int32_t* cells, int64_t ncell, int32_t nn, int32_t nq, int32_t vec,
                                    int32_t dim, const double* u, const double* shape_grads,
                                    const double* shape_vals, double* u_grad, double* u_quad, hf_stream_t stream) {
  if (check_shape(vec, dim, nn, nq)) return -1;
  HF_REQUIRE(cells && u && shape_grads && u_grad && (!u_quad || shape_vals), "hf_generic_ugrad_f64: null argument");
  if (ncell == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const int g = gen_grid(ncell * nq);
#define UG(V, D) \
  ugrad_kernel<V, D><<<g, kGenBlock, 0, s>>>(cells, ncell, nn, nq, u, shape_grads, shape_vals, u_grad, u_quad)
  GEN_DISPATCH(UG);
#undef UG
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_generic_residual_f64(int64_t ncell, int32_t nn, int32_t nq, int32_t vec, int32_t dim,
                                       const double* shape_grads, const double* shape_vals, const double* jxw,
                                       const double* P, const double* M, double* elem_res, hf_stream_t stream) {
  if (check_shape(vec, dim, nn, nq)) return -1;
  HF_REQUIRE(shape_grads && jxw && elem_res && (!M || shape_vals), "hf_generic_residual_f64: null argument");
  if (ncell == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  const int g = gen_grid(ncell * nn);
#define RES(V, D) \
  residual_kernel<V, D><<<g, kGenBlock, 0, s>>>(ncell, nn, nq, shape_grads, shape_vals, jxw, P, M, elem_res)
  GEN_DISPATCH(RES);
#undef RES
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_generic_jacobian_f64(int64_t ncell, int32_t nn, int32_t nq, int32_t vec, int32_t dim,
                                       const double* shape_grads, const double* shape_vals, const double* jxw,
                                       const double* C, const double* Mt, double* elem_jac, hf_stream_t stream) {
  return jacobian_impl<double>(ncell, nn, nq, vec, dim, shape_grads, shape_vals, jxw, C, Mt, elem_jac,
                               (hipStream_t)stream);
}

extern "C" int hf_generic_jacobian_f32(int64_t ncell, int32_t nn, int32_t nq, int32_t vec, int32_t dim,
                                       const double* shape_grads, const double* shape_vals, const double* jxw,
                                       const double* C, const double* Mt, float* elem_jac, hf_stream_t stream) {
  return jacobian_impl<float>(ncell, nn, nq, vec, dim, shape_grads, shape_vals, jxw, C, Mt, elem_jac,
                              (hipStream_t)stream);
}
