// HEX27 cell kernels: element residual, element Jacobian, parameter VJP, on
// 3x3x3 Gauss points (quadrature_order 4 or 5, basis.py:64 rule (order+2)//2).
//
// Math as the HEX8 kernels (jax_fem/problem.py:189-214 laplace kernel,
// :262-266 jacfwd; jax_fem/fe.py:112-141 geometry), with the triquadratic
// Lagrange basis on [0,1]^3 through the nodes 0, 1/2, 1 per direction, nodes in
// VTK order (the reference's HEX27 re_order, basis.py:60-61, maps them onto
// basix).  Quadrature index q = (i*3 + j)*3 + k with x from i.
//
// Residual (one wavefront per cell): lane q < 27 builds J = sum_n x_n (x)
// dN_n/dxi, grad u = (sum_n u_n (x) dN_n/dxi) J^-1, P and S_q = w P J^-T;
// lane (n, i) then sums r_n[i] = sum_q sum_e dN_n/dxi_e(q) S_q[i][e].
//
// Jacobian, K[(a,i),(b,k)] = sum_q sum_{J,L} G_a[q][J] C_q[iJkL] G_b[q][L] with C_q
// the full tangent (w included) and G the physical gradients.
//   FP64 (one 256-thread workgroup per cell, everything in LDS):
//     M_k[(a,i)][(q,L)] = sum_J G_a[q][J] C_q[iJkL], K_k = M_k G on the vector ALU.
//   FP32 on MFMA (hex27_jacobian_mfma: one wavefront per cell, 4 cells per
//   workgroup, no workgroup barrier): for each (i,k) the 27 x 27 block
//     K^{ik}[a][b] = sum_r A[a][r] V^{ik}[r][b],  r = (q, J),
//     A[a][r] = G[q][a][J],  V^{ik}[r][b] = sum_L C_q[iJkL] G[q][b][L],
//   is one GEMM on v_mfma_f32_32x32x2f32 (27 padded to 32, K = 84 in 42 steps);
//   the tangent is symmetric, so the six blocks i <= k are formed and the other
//   three written transposed.
//   Lane l stands for node n = l % 32 as both the row a of A and the column b
//   of V, and half-wave h = l / 32 owns the quadrature points q = t + 14h
//   (t = 0..13): step s feeds r = (t = s / 3, J = s % 3), so every operand is a
//   compile-time register of the lane's own gradients Gl[t][.] (42 floats),
//   and only the 3 tangent entries C_q[iJk.] come from LDS, a broadcast per
//   half-wave.  No index division or bound test in the K loop.
// 2 * 3 * 81 * 81 * 27 = 1.06 MFLOP per cell of algorithmic GEMM work (the
// six blocks executed: 0.93 MFLOP on MFMA with the padding), 0.12 MFLOP forming
// V.  Output block-major (cell, a, b, i, k).
#include "common.hpp"
#include "element_math.hpp"

namespace hf {


// Lattice position (0, 1, 2 = reference coordinate 0, 1/2, 1) of each VTK node.
__constant__ int8_t c_h27_node[27][3] = {
    {0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2}, {0, 2, 2}, {1, 0, 0},
    {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2}, {1, 2, 2}, {0, 1, 2}, {0, 0, 1}, {2, 0, 1},
    {2, 2, 1}, {0, 2, 1}, {0, 1, 1}, {2, 1, 1}, {1, 0, 1}, {1, 2, 1}, {1, 1, 0}, {1, 1, 2}, {1, 1, 1}};

// 3-point Gauss-Legendre on [0,1]: 1/2 -+ sqrt(3/5)/2, weights 5/18, 8/18, 5/18.
__host__ __device__ constexpr double g3_point(int i) {
  return i == 0 ? 0.5 - 0.38729833462074168852 : (i == 1 ? 0.5 : 0.5 + 0.38729833462074168852);
}
__host__ __device__ constexpr double g3_weight(int i) { return i == 1 ? 8.0 / 18.0 : 5.0 / 18.0; }

// quadratic Lagrange through 0, 1/2, 1 and its derivative
template <typename T>
__host__ __device__ constexpr T lag2(int a, T t) {
  return a == 0 ? (T)2 * (t - (T)0.5) * (t - (T)1) : (a == 1 ? (T)(-4) * t * (t - (T)1) : (T)2 * t * (t - (T)0.5));
}
template <typename T>
__host__ __device__ constexpr T dlag2(int a, T t) {
  return a == 0 ? (T)4 * t - (T)3 : (a == 1 ? (T)4 - (T)8 * t : (T)4 * t - (T)1);
}

// The reference gradients dN_n/dxi at the 27 Gauss points, [q][n][e], evaluated
// in FP64 at compile time (a kernel reads three entries instead of evaluating
// three products of Lagrange factors with their branches per node and point).
struct H27Grad {
  double v[27][27][3];
};
constexpr H27Grad make_h27_grad() {
  constexpr int8_t pos[27][3] = {{0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2},
                                 {0, 2, 2}, {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2},
                                 {1, 2, 2}, {0, 1, 2}, {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1}, {0, 1, 1},
                                 {2, 1, 1}, {1, 0, 1}, {1, 2, 1}, {1, 1, 0}, {1, 1, 2}, {1, 1, 1}};
  H27Grad g{};
  for (int q = 0; q < 27; ++q) {
    const double tx = g3_point(q / 9), ty = g3_point((q / 3) % 3), tz = g3_point(q % 3);
    for (int n = 0; n < 27; ++n) {
      const int a = pos[n][0], b = pos[n][1], c = pos[n][2];
      const double vx = lag2<double>(a, tx), vy = lag2<double>(b, ty), vz = lag2<double>(c, tz);
      g.v[q][n][0] = dlag2<double>(a, tx) * vy * vz;
      g.v[q][n][1] = vx * dlag2<double>(b, ty) * vz;
      g.v[q][n][2] = vx * vy * dlag2<double>(c, tz);
    }
  }
  return g;
}
__constant__ H27Grad c_h27_grad = make_h27_grad();

// reference gradient dN_n/dxi at quadrature point q
template <typename T>
__device__ __forceinline__ void ref_grad27(int n, int q, T dN[3]) {
  const double* g = c_h27_grad.v[q][n];
  dN[0] = (T)g[0];
  dN[1] = (T)g[1];
  dN[2] = (T)g[2];
}

__device__ __forceinline__ double g3_qweight(int q) {
  return g3_weight(q / 9) * g3_weight((q / 3) % 3) * g3_weight(q % 3);
}

// Geometry and displacement gradient at q from LDS node data: inv = (dx/dxi)^-1,
// ug = grad u, returns w = det * weight.
template <typename T>
__device__ __forceinline__ T hex27_point(int q, const T (*X)[3], const T (*U)[3], T inv[3][3], T ug[3][3]) {
  T Jm[3][3], Hr[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int e = 0; e < 3; ++e) Jm[d][e] = Hr[d][e] = (T)0;
  // partly unrolled: fully unrolled, the 81 gradient loads are hoisted and the
  // FP64 kernels run at one wave per SIMD
#pragma unroll 3
  for (int n = 0; n < 27; ++n) {
    T dN[3];
    ref_grad27<T>(n, q, dN);
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        Jm[d][e] += X[n][d] * dN[e];
        Hr[d][e] += U[n][d] * dN[e];
      }
  }
  const T det = inv3(Jm, inv);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) ug[i][j] = Hr[i][0] * inv[0][j] + Hr[i][1] * inv[1][j] + Hr[i][2] * inv[2][j];
  return det * (T)g3_qweight(q);
}

template <typename T>
__device__ __forceinline__ void hex27_load(const double* __restrict__ pts, const int32_t* __restrict__ cells,
                                           const double* __restrict__ v, int64_t c, int t, T (*X)[3], T (*U)[3]) {
  if (t < 27) {
    const int64_t nid = cells[c * 27 + t];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      X[t][d] = (T)pts[3 * nid + d];
      U[t][d] = (T)v[3 * nid + d];
    }
  }
}

// ---------------------------------------------------------------------------
template <typename T, int MAT>
__global__ void __launch_bounds__(64, 4) hex27_residual_kernel(const double* __restrict__ pts,
                                                            const int32_t* __restrict__ cells, int64_t ncell,
                                                            const double* __restrict__ u,
                                                            const double* __restrict__ params, hf_material mat,
                                                            T* __restrict__ out) {
  __shared__ T sX[27][3], sU[27][3], sS[27][9];
  const int64_t c = blockIdx.x;
  const int t = threadIdx.x;
  hex27_load<T>(pts, cells, u, c, t, sX, sU);
  __syncthreads();
  if (t < 27) {
    const int q = t;
    T inv[3][3], ug[3][3];
    const T w = hex27_point<T>(q, sX, sU, inv, ug);
    const Moduli md = moduli(mat, params ? params[c * 27 + q] : 1.0, params != nullptr);
    T P[3][3];
    material_stress<T, MAT, 3>(ug, (T)md.m1, (T)md.m2, P);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int e = 0; e < 3; ++e) sS[q][i * 3 + e] = w * (P[i][0] * inv[e][0] + P[i][1] * inv[e][1] + P[i][2] * inv[e][2]);
  }
  __syncthreads();
  for (int o = t; o < 81; o += 64) {
    const int n = o / 3, i = o - 3 * (o / 3);
    T r = 0;
#pragma unroll 3
    for (int q = 0; q < 27; ++q) {
      T dN[3];
      ref_grad27<T>(n, q, dN);
      r += dN[0] * sS[q][i * 3] + dN[1] * sS[q][i * 3 + 1] + dN[2] * sS[q][i * 3 + 2];
    }
    out[c * 81 + o] = r;
  }
}

// ---------------------------------------------------------------------------
constexpr int kMStride = 85;  // row stride of M_k in LDS (81 columns + padding)

template <typename T, int MAT>
__global__ void __launch_bounds__(256) hex27_jacobian_kernel(const double* __restrict__ pts,
                                                             const int32_t* __restrict__ cells, int64_t ncell,
                                                             const double* __restrict__ u,
                                                             const double* __restrict__ params, hf_material mat,
                                                             T* __restrict__ out) {
  __shared__ T sX[27][3], sU[27][3];
  __shared__ T sC[27][81];    // tangent per quadrature point, w included
  __shared__ T sInv[27][9];
  __shared__ T sG[27][27][3];  // physical gradients [q][node][l]
  __shared__ T sM[81 * kMStride];
  const int64_t c = blockIdx.x;
  const int t = threadIdx.x;
  hex27_load<T>(pts, cells, u, c, t, sX, sU);
  __syncthreads();
  if (t < 27) {
    const int q = t;
    T inv[3][3], ug[3][3];
    const T w = hex27_point<T>(q, sX, sU, inv, ug);
    const Moduli md = moduli(mat, params ? params[c * 27 + q] : 1.0, params != nullptr);
    material_tangent<T, MAT>(ug, (T)md.m1, (T)md.m2, w, &sC[q][0]);
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
      for (int l = 0; l < 3; ++l) sInv[q][e * 3 + l] = inv[e][l];
  }
  __syncthreads();
  for (int p = t; p < 27 * 27; p += 256) {
    const int q = p / 27, b = p - 27 * (p / 27);
    T dN[3];
    ref_grad27<T>(b, q, dN);
#pragma unroll
    for (int l = 0; l < 3; ++l) sG[q][b][l] = dN[0] * sInv[q][l] + dN[1] * sInv[q][3 + l] + dN[2] * sInv[q][6 + l];
  }
  __syncthreads();
  T* o = out + c * (27 * 27 * 9);
  for (int k = 0; k < 3; ++k) {
    for (int p = t; p < 81 * 81; p += 256) {
      const int row = p / 81, col = p - 81 * (p / 81);
      const int a = row / 3, i = row - 3 * (row / 3), q = col / 3, l = col - 3 * (col / 3);
      const T* Cq = &sC[q][0];
      sM[row * kMStride + col] = sG[q][a][0] * Cq[((i * 3 + 0) * 3 + k) * 3 + l] +
                                 sG[q][a][1] * Cq[((i * 3 + 1) * 3 + k) * 3 + l] +
                                 sG[q][a][2] * Cq[((i * 3 + 2) * 3 + k) * 3 + l];
    }
    __syncthreads();
    {
      for (int p = t; p < 81 * 27; p += 256) {
        const int row = p / 27, b = p - 27 * (p / 27);
        const T* mr = sM + row * kMStride;
        T s = 0;
        for (int q = 0; q < 27; ++q)
          s += mr[3 * q] * sG[q][b][0] + mr[3 * q + 1] * sG[q][b][1] + mr[3 * q + 2] * sG[q][b][2];
        const int a = row / 3, i = row - 3 * (row / 3);
        o[((a * 27 + b) * 3 + i) * 3 + k] = s;
      }
    }
    __syncthreads();  // sM is rewritten for the next k
  }
}

// ---------------------------------------------------------------------------
// FP32 element Jacobian on MFMA, one wavefront per cell.
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kH27Waves = 4;   // cells (wavefronts) per workgroup
constexpr int kH27T = 14;      // quadrature points per half-wave

// VTK node lattice positions (c_h27_node) as a constexpr function: folds to a
// constant in unrolled loops, so the 1-D factor tables are indexed at compile time.
__host__ __device__ constexpr int h27_pos(int m, int d) {
  constexpr int8_t tbl[27][3] = {{0, 0, 0}, {2, 0, 0}, {2, 2, 0}, {0, 2, 0}, {0, 0, 2}, {2, 0, 2}, {2, 2, 2},
                                 {0, 2, 2}, {1, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 1, 0}, {1, 0, 2}, {2, 1, 2},
                                 {1, 2, 2}, {0, 1, 2}, {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1}, {0, 1, 1},
                                 {2, 1, 1}, {1, 0, 1}, {1, 2, 1}, {1, 1, 0}, {1, 1, 2}, {1, 1, 1}};
  return tbl[m][d];
}

struct H27Wave {          // one wavefront's LDS
  float X[27][3], U[27][3];
  float C[2 * kH27T][81];  // tangent per point (w included); point 27 is zero padding
  float Inv[27][9];
};

__device__ __forceinline__ void h27_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int MAT>
__global__ void __launch_bounds__(64 * kH27Waves, 2) hex27_jacobian_mfma(const double* __restrict__ pts,
                                                                      const int32_t* __restrict__ cells,
                                                                      int64_t ncell, const double* __restrict__ u,
                                                                      const double* __restrict__ params,
                                                                      hf_material mat, float* __restrict__ out) {
  __shared__ H27Wave sw[kH27Waves];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t c = (int64_t)blockIdx.x * kH27Waves + wv;
  if (c >= ncell) return;  // wave-uniform; the kernel has no workgroup barrier
  H27Wave& S = sw[wv];
  const int h = lane >> 5, n = lane & 31;

  // nodes of the cell
  if (lane < 27) {
    const int64_t nid = cells[c * 27 + lane];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      S.X[lane][d] = (float)pts[3 * nid + d];
      S.U[lane][d] = (float)u[3 * nid + d];
    }
  }
  S.C[27][lane] = 0.f;                       // the padding point: entries 0..63
  if (lane < 81 - 64) S.C[27][64 + lane] = 0.f;  // and 64..80
  h27_wave_sync();

  // geometry and tangent at point q = n: half 0 sums x (x) dN, half 1 u (x) dN
  {
    const int q = n < 27 ? n : 26;
    const float tq[3] = {(float)g3_point(q / 9), (float)g3_point((q / 3) % 3), (float)g3_point(q % 3)};
    float P1[3][3], D1[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        P1[d][a] = lag2<float>(a, tq[d]);
        D1[d][a] = dlag2<float>(a, tq[d]);
      }
    const float* src = h ? &S.U[0][0] : &S.X[0][0];
    float acc[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int e = 0; e < 3; ++e) acc[d][e] = 0.f;
#pragma unroll
    for (int m = 0; m < 27; ++m) {
      const int a = h27_pos(m, 0), b = h27_pos(m, 1), cc = h27_pos(m, 2);
      const float dN0 = D1[0][a] * P1[1][b] * P1[2][cc];
      const float dN1 = P1[0][a] * D1[1][b] * P1[2][cc];
      const float dN2 = P1[0][a] * P1[1][b] * D1[2][cc];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const float v = src[m * 3 + d];
        acc[d][0] += v * dN0;
        acc[d][1] += v * dN1;
        acc[d][2] += v * dN2;
      }
    }
    float Jm[3][3], Hr[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const float other = __shfl_xor(acc[d][e], 32, 64);
        Jm[d][e] = h ? other : acc[d][e];
        Hr[d][e] = h ? acc[d][e] : other;
      }
    float inv[3][3], ug[3][3];
    const float det = inv3(Jm, inv);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) ug[i][j] = Hr[i][0] * inv[0][j] + Hr[i][1] * inv[1][j] + Hr[i][2] * inv[2][j];
    if (n < 27) {
      if (h == 0) {
        const float w = det * (float)g3_qweight(q);
        const Moduli md = moduli(mat, params ? params[c * 27 + q] : 1.0, params != nullptr);
        material_tangent<float, MAT>(ug, (float)md.m1, (float)md.m2, w, &S.C[q][0]);
      } else {
#pragma unroll
        for (int e = 0; e < 3; ++e)
#pragma unroll
          for (int l = 0; l < 3; ++l) S.Inv[q][e * 3 + l] = inv[e][l];
      }
    }
  }
  h27_wave_sync();

  // this lane's node's physical gradients at its half's points: Gl[t][L] = G[t + 14h][n][L]
  float Gl[kH27T][3];
  {
    const int nn = n < 27 ? n : 0;
    float lv[3][3], ld[3][3];  // [direction][Gauss point] for the node's lattice position
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int a = c_h27_node[nn][d];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        lv[d][p] = lag2<float>(a, (float)g3_point(p));
        ld[d][p] = dlag2<float>(a, (float)g3_point(p));
      }
    }
    const bool node_ok = n < 27;
#pragma unroll
    for (int t = 0; t < kH27T; ++t) {
      // the point index of each half is a compile-time constant: select between them
      const int q0 = t, q1 = t + kH27T < 27 ? t + kH27T : 26;
      const float fx = h ? lv[0][q1 / 9] : lv[0][q0 / 9], gx = h ? ld[0][q1 / 9] : ld[0][q0 / 9];
      const float fy = h ? lv[1][(q1 / 3) % 3] : lv[1][(q0 / 3) % 3];
      const float gy = h ? ld[1][(q1 / 3) % 3] : ld[1][(q0 / 3) % 3];
      const float fz = h ? lv[2][q1 % 3] : lv[2][q0 % 3], gz = h ? ld[2][q1 % 3] : ld[2][q0 % 3];
      const float dN0 = gx * fy * fz, dN1 = fx * gy * fz, dN2 = fx * fy * gz;
      const int q = h ? q1 : q0;
      const bool ok = node_ok && (t + kH27T * h) < 27;
      const float* iv = &S.Inv[q][0];
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        const float g = dN0 * iv[l] + dN1 * iv[3 + l] + dN2 * iv[6 + l];
        Gl[t][l] = ok ? g : 0.f;
      }
    }
  }

  // The tangent is symmetric (C_iJkL = C_kLiJ for a hyperelastic law), so
  // K^{ik} = (K^{ki})^T: six GEMMs (i <= k) stay in accumulators and the three
  // blocks below the diagonal are read transposed through LDS, so that each 3x3
  // block (a, b) of the output is written at once: 9 consecutive words per lane,
  // 27 lanes -> 972 contiguous bytes per row a (full cache lines while hot in L2)
  const float* Cbase = &S.C[kH27T * h][0];
  constexpr int kI[6] = {0, 0, 0, 1, 1, 2}, kK[6] = {0, 1, 2, 1, 2, 2};
  f32x16 acc[6];
#pragma unroll
  for (int gk = 0; gk < 6; ++gk) {
    const int i = kI[gk], k = kK[gk];
    const float* Cik = Cbase + i * 27 + k * 3;   // + t * 81 + J * 9 + L
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[gk][r] = 0.f;
#pragma unroll
    for (int s2 = 0; s2 < 3 * kH27T; ++s2) {
      const int t = s2 / 3, J = s2 - 3 * (s2 / 3);
      const float* cq = Cik + t * 81 + J * 9;
      const float bv = cq[0] * Gl[t][0] + cq[1] * Gl[t][1] + cq[2] * Gl[t][2];
      acc[gk] = __builtin_amdgcn_mfma_f32_32x32x2f32(Gl[t][J], bv, acc[gk], 0, 0, 0);
    }
  }
  // D[m][n]: lane holds column n, rows m = 8 (r / 4) + 4 h + r % 4.  The blocks
  // (0,1), (0,2), (1,2) go to the wave's tangent region (no longer read), 27 x 27 each
  h27_wave_sync();
  float* T = &S.C[0][0];
  static_assert(3 * 27 * 27 <= 2 * kH27T * 81, "transposed blocks exceed the tangent region");
  if (n < 27) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = 8 * (r / 4) + 4 * h + (r % 4);
      if (m < 27) {
        T[m * 27 + n] = acc[1][r];
        T[729 + m * 27 + n] = acc[2][r];
        T[1458 + m * 27 + n] = acc[4][r];
      }
    }
  }
  h27_wave_sync();
  if (n < 27) {
    float* o = out + c * (27 * 27 * 9) + (int64_t)n * 9;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = 8 * (r / 4) + 4 * h + (r % 4);
      if (m < 27) {
        float* blk = o + m * (27 * 9);
        blk[0] = acc[0][r];
        blk[1] = acc[1][r];
        blk[2] = acc[2][r];
        blk[3] = T[n * 27 + m];          // (1, 0): K^{01}[n][m]
        blk[4] = acc[3][r];
        blk[5] = acc[4][r];
        blk[6] = T[729 + n * 27 + m];    // (2, 0): K^{02}[n][m]
        blk[7] = T[1458 + n * 27 + m];   // (2, 1): K^{12}[n][m]
        blk[8] = acc[5][r];
      }
    }
  }
}

// grad[c,q] = - w sum_ij (grad lam)_ij dP_ij/dp
template <int MAT>
__global__ void __launch_bounds__(64, 4) hex27_param_vjp_kernel(const double* __restrict__ pts,
                                                             const int32_t* __restrict__ cells, int64_t ncell,
                                                             const double* __restrict__ u,
                                                             const double* __restrict__ params, hf_material mat,
                                                             const double* __restrict__ lam,
                                                             double* __restrict__ grad) {
  __shared__ double sX[27][3], sU[27][3], sL[27][3];
  const int64_t c = blockIdx.x;
  const int t = threadIdx.x;
  hex27_load<double>(pts, cells, u, c, t, sX, sU);
  if (t < 27) {
    const int64_t nid = cells[c * 27 + t];
#pragma unroll
    for (int d = 0; d < 3; ++d) sL[t][d] = lam[3 * nid + d];
  }
  __syncthreads();
  if (t < 27) {
    const int q = t;
    double inv[3][3], ug[3][3], inv2[3][3], lg[3][3];
    const double w = hex27_point<double>(q, sX, sU, inv, ug);
    hex27_point<double>(q, sX, sL, inv2, lg);
    const Moduli md = moduli(mat, params[c * 27 + q], true);
    double P1[3][3], P2[3][3];
    material_stress_parts<double, MAT, 3>(ug, P1, P2);
    double s = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) s += lg[i][j] * (md.d1 * P1[i][j] + md.d2 * P2[i][j]);
    grad[c * 27 + q] = -w * s;
  }
}

// ---------------------------------------------------------------------------
template <typename T>
static int launch_hex27_residual(const hf_material* m, const double* p, const int32_t* c, int64_t n,
                                 const double* u, const double* prm, T* o, hipStream_t s) {
  if (n == 0) return 0;
  if (m->kind == MAT_LINEAR_ELASTIC)
    hex27_residual_kernel<T, MAT_LINEAR_ELASTIC><<<(unsigned)n, 64, 0, s>>>(p, c, n, u, prm, *m, o);
  else
    hex27_residual_kernel<T, MAT_NEO_HOOKEAN><<<(unsigned)n, 64, 0, s>>>(p, c, n, u, prm, *m, o);
  HF_LAUNCH_CHECK();
  return 0;
}

template <typename T, bool MFMA>
static int launch_hex27_jacobian(const hf_material* m, const double* p, const int32_t* c, int64_t n,
                                 const double* u, const double* prm, T* o, hipStream_t s) {
  if (n == 0) return 0;
  if constexpr (MFMA) {
    const unsigned grid = (unsigned)ceil_div(n, kH27Waves);
    if (m->kind == MAT_LINEAR_ELASTIC)
      hex27_jacobian_mfma<MAT_LINEAR_ELASTIC><<<grid, 64 * kH27Waves, 0, s>>>(p, c, n, u, prm, *m, o);
    else
      hex27_jacobian_mfma<MAT_NEO_HOOKEAN><<<grid, 64 * kH27Waves, 0, s>>>(p, c, n, u, prm, *m, o);
  } else {
    if (m->kind == MAT_LINEAR_ELASTIC)
      hex27_jacobian_kernel<T, MAT_LINEAR_ELASTIC><<<(unsigned)n, 256, 0, s>>>(p, c, n, u, prm, *m, o);
    else
      hex27_jacobian_kernel<T, MAT_NEO_HOOKEAN><<<(unsigned)n, 256, 0, s>>>(p, c, n, u, prm, *m, o);
  }
  HF_LAUNCH_CHECK();
  return 0;
}

int hex27_residual_f64(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                       const double* prm, double* o, hipStream_t s) {
  return launch_hex27_residual<double>(m, p, c, n, u, prm, o, s);
}
int hex27_residual_f32(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                       const double* prm, float* o, hipStream_t s) {
  return launch_hex27_residual<float>(m, p, c, n, u, prm, o, s);
}
int hex27_jacobian_f64(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                       const double* prm, double* o, hipStream_t s) {
  return launch_hex27_jacobian<double, false>(m, p, c, n, u, prm, o, s);
}
int hex27_jacobian_f32(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                       const double* prm, float* o, hipStream_t s) {
  return launch_hex27_jacobian<float, true>(m, p, c, n, u, prm, o, s);
}
int hex27_param_vjp(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                    const double* prm, const double* lam, double* grad, hipStream_t s) {
  if (n == 0) return 0;
  if (m->kind == MAT_LINEAR_ELASTIC)
    hex27_param_vjp_kernel<MAT_LINEAR_ELASTIC><<<(unsigned)n, 64, 0, s>>>(p, c, n, u, prm, *m, lam, grad);
  else
    hex27_param_vjp_kernel<MAT_NEO_HOOKEAN><<<(unsigned)n, 64, 0, s>>>(p, c, n, u, prm, *m, lam, grad);
  HF_LAUNCH_CHECK();
  return 0;
}

}  // namespace hf
