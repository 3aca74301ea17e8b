// Small dense algebra and constitutive laws shared by the cell kernels.
#pragma once
#include "common.hpp"
#include "hipfem.h"

namespace hf {

template <typename T>
__device__ __forceinline__ T det3(const T A[3][3]) {
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
         A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}

// inv = A^-1, returns det(A).
template <typename T>
__device__ __forceinline__ T inv3(const T A[3][3], T inv[3][3]) {
  const T c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const T c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const T c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const T det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  const T id = (T)1 / det;
  inv[0][0] = c00 * id;
  inv[1][0] = c01 * id;
  inv[2][0] = c02 * id;
  inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  return det;
}

template <typename T>
__device__ __forceinline__ void matvec3(const T A[3][3], const T x[3], T y[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) y[i] = A[i][0] * x[0] + A[i][1] * x[1] + A[i][2] * x[2];
}

// Register-array element select with a lane-varying index (no scratch).
template <typename T, int NN>
__device__ __forceinline__ T select_node(const T g[NN][3], int a, int j) {
  T v = g[0][j];
#pragma unroll
  for (int n = 1; n < NN; ++n) v = (a == n) ? g[n][j] : v;
  return v;
}
template <typename T>
__device__ __forceinline__ T select8(const T g[8][3], int a, int j) {
  return select_node<T, 8>(g, a, j);
}

template <typename T, int NN, int VEC>
__device__ __forceinline__ void grad_u(const T U[NN][VEC], const T g[NN][3], T ug[VEC][3]) {
#pragma unroll
  for (int i = 0; i < VEC; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T s = 0;
#pragma unroll
      for (int n = 0; n < NN; ++n) s += U[n][i] * g[n][j];
      ug[i][j] = s;
    }
}

// Moduli (m1, m2) and their derivatives w.r.t. the per-quadrature parameter:
// (mu, lambda) linear elastic, (mu, kappa) Neo-Hookean, (k, 0) Poisson.
struct Moduli {
  double m1, m2, d1, d2;
};

__device__ __forceinline__ void moduli_from_E(const hf_material& m, double E, double& m1, double& m2) {
  const double nu = m.nu;
  m1 = E / (2.0 * (1.0 + nu));
  m2 = (m.kind == MAT_LINEAR_ELASTIC) ? E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
                                      : E / (3.0 * (1.0 - 2.0 * nu));
}

__device__ __forceinline__ Moduli moduli(const hf_material& m, double p, bool has_p) {
  Moduli r{0, 0, 0, 0};
  if (m.kind == MAT_POISSON) {
    if (has_p && m.param_mode == P_SCALE_K) {
      r.m1 = m.k * p;
      r.d1 = m.k;
    } else {
      r.m1 = m.k;
    }
    return r;
  }
  if (!has_p || m.param_mode == P_NONE) {
    moduli_from_E(m, m.E, r.m1, r.m2);
    return r;
  }
  if (m.param_mode == P_SCALE_E) {
    moduli_from_E(m, m.E * p, r.m1, r.m2);
    moduli_from_E(m, m.E, r.d1, r.d2);
  } else if (m.param_mode == P_SIMP) {
    const double pp = pow(p, m.penal - 1.0);
    const double E = m.Emin + (m.E0 - m.Emin) * pp * p;
    const double dE = (m.E0 - m.Emin) * m.penal * pp;
    moduli_from_E(m, E, r.m1, r.m2);
    moduli_from_E(m, 1.0, r.d1, r.d2);
    r.d1 *= dE;
    r.d2 *= dE;
  } else if (m.param_mode == P_MU) {
    r.m1 = p;
    r.m2 = m.kappa;
    r.d1 = 1.0;
    r.d2 = 0.0;
  }
  return r;
}

template <typename T>
struct NHState {
  T F[3][3], H[3][3];  // F = I + grad u, H = F^-T
  T J, I1, c;          // det F, tr(F^T F), J^-2/3
};

template <typename T>
__device__ __forceinline__ T inv_cbrt_sq(T J) {
  // J^(-2/3)
  if constexpr (sizeof(T) == 8) {
    const double cb = cbrt(J);
    return 1.0 / (cb * cb);
  } else {
    const float cb = cbrtf(J);
    return 1.0f / (cb * cb);
  }
}

template <typename T>
__device__ __forceinline__ void nh_state(const T ug[3][3], NHState<T>& s) {
  s.I1 = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      s.F[i][j] = ug[i][j] + (i == j ? (T)1 : (T)0);
      s.I1 += s.F[i][j] * s.F[i][j];
    }
  T inv[3][3];
  s.J = inv3(s.F, inv);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) s.H[i][j] = inv[j][i];
  s.c = inv_cbrt_sq(s.J);
}

// P = m1 * P1 + m2 * P2
template <typename T, int MAT, int VEC>
__device__ __forceinline__ void material_stress_parts(const T ug[VEC][3], T P1[VEC][3], T P2[VEC][3]) {
  if constexpr (MAT == MAT_POISSON) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      P1[0][j] = ug[0][j];
      P2[0][j] = 0;
    }
  } else if constexpr (MAT == MAT_LINEAR_ELASTIC) {
    const T tr = ug[0][0] + ug[1][1] + ug[2][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        P1[i][j] = ug[i][j] + ug[j][i];
        P2[i][j] = (i == j) ? tr : (T)0;
      }
  } else {
    NHState<T> s;
    nh_state<T>(ug, s);
    const T a = s.I1 * (T)(1.0 / 3.0);
    const T b = s.J * (s.J - (T)1);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        P1[i][j] = s.c * (s.F[i][j] - a * s.H[i][j]);
        P2[i][j] = b * s.H[i][j];
      }
  }
}

// Full tangent C[((i*3+J)*3+k)*3+L] = dP_iJ / dF_kL, scaled by w (vec = 3).
//   linear elastic (m1 = mu, m2 = lambda): mu (d_ik d_JL + d_iL d_Jk) + lambda d_iJ d_kL
//   Neo-Hookean (m1 = mu, m2 = kappa; c = J^-2/3, H = F^-T):
//     mu c [d_ik d_JL - 2/3 H_kL F_iJ - 2/3 F_kL H_iJ + 2/9 I1 H_kL H_iJ + 1/3 I1 H_iL H_kJ]
//     + kappa [(2J - 1) J H_kL H_iJ - (J^2 - J) H_iL H_kJ]
// Contracted with g_a, g_b this is the closed form the HEX8 Jacobian kernel uses.
template <typename T, int MAT>
__device__ __forceinline__ void material_tangent(const T ug[3][3], T m1, T m2, T w, T* __restrict__ C) {
  if constexpr (MAT == MAT_LINEAR_ELASTIC) {
    const T mu = m1 * w, la = m2 * w;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int J = 0; J < 3; ++J)
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int L = 0; L < 3; ++L)
            C[((i * 3 + J) * 3 + k) * 3 + L] = mu * ((T)(i == k && J == L) + (T)(i == L && J == k)) +
                                               la * (T)(i == J && k == L);
  } else {
    NHState<T> s;
    nh_state<T>(ug, s);
    const T s1 = m1 * s.c * w, s2 = m2 * w;
    const T a1 = (T)(-2.0 / 3.0) * s1, a2 = (T)(2.0 / 9.0) * s1 * s.I1 + s2 * ((T)2 * s.J - (T)1) * s.J;
    const T a3 = (T)(1.0 / 3.0) * s1 * s.I1 - s2 * s.J * (s.J - (T)1);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int J = 0; J < 3; ++J)
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int L = 0; L < 3; ++L)
            C[((i * 3 + J) * 3 + k) * 3 + L] = s1 * (T)(i == k && J == L) +
                                               a1 * (s.H[k][L] * s.F[i][J] + s.F[k][L] * s.H[i][J]) +
                                               a2 * s.H[k][L] * s.H[i][J] + a3 * s.H[i][L] * s.H[k][J];
  }
}

template <typename T, int MAT, int VEC>
__device__ __forceinline__ void material_stress(const T ug[VEC][3], T m1, T m2, T P[VEC][3]) {
  T P1[VEC][3], P2[VEC][3];
  material_stress_parts<T, MAT, VEC>(ug, P1, P2);
#pragma unroll
  for (int i = 0; i < VEC; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) P[i][j] = m1 * P1[i][j] + m2 * P2[i][j];
}

}  // namespace hf
