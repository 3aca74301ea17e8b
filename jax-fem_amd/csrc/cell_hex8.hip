// HEX8 cell kernels: element residual, element Jacobian, parameter VJP.
//
// Math (reference: jax_fem/problem.py:189-214 laplace kernel, :262-266 jacfwd;
// jax_fem/fe.py:112-141 geometry):
//   J = sum_n x_n (x) dN_n/dxi, g_n = dN_n/dxi . J^-1, w = det(J) * w_ref
//   u_grad[i][j] = sum_n u[n][i] g_n[j],  r[n][i] = sum_q w sum_j P[i][j] g_n[j]
//   K[a i][b k]  = sum_q w g_a . A[i,:,k,:] . g_b, contracted in closed form:
//     linear elastic  w (lam g_a[i] g_b[k] + mu (d_ik g_a.g_b + g_a[k] g_b[i]))
//     Neo-Hookean     (H = F^-T, c = J^-2/3, s1 = mu c w, s2 = kappa w)
//        alpha_i Hg_b[k] + beta_i Fg_b[k] + Hg_b[i] gamma_k + d_ik s1 g_a.g_b
//        alpha = s1 (-2/3 Fg_a + 2/9 I1 Hg_a) + s2 (2J-1) J Hg_a
//        beta  = -2/3 s1 Hg_a,  gamma = (s1 I1/3 - s2 J (J-1)) Hg_a
// Geometry is recomputed per quadrature point from the node coordinates (no
// stored shape gradients: 0 bytes of per-cell geometry traffic).
//
// Thread mapping (wave64):
//   residual / param VJP: one thread per cell (24 DOF in registers).
//   Jacobian: 8 consecutive lanes per cell; lane q first forms quadrature point
//   q's shared data in LDS, then lane a contracts row a over the 8 points and
//   writes its 8 x 3 x 3 row of blocks (288 contiguous bytes in FP32): the 8
//   lanes of a cell write one contiguous 2304-byte element matrix.
#include "common.hpp"
#include "element_math.hpp"
#include "hipfem.h"
#include "krylov_common.hpp"

namespace hf {

// Reference gradient of mesh-order HEX8 node n at Gauss point q along d.
// Node reference coordinates in VTK order; Gauss points (1 -+ 1/sqrt3)/2 with
// quad index q = (i*2 + j)*2 + k (x from i) as in the oracle's tensor rule.
__host__ __device__ constexpr int hex8_node_bit(int n, int d) {
  // x: 0,1,1,0,0,1,1,0   y: 0,0,1,1,0,0,1,1   z: 0,0,0,0,1,1,1,1
  return d == 0 ? ((n & 3) == 1 || (n & 3) == 2) : d == 1 ? ((n & 3) >= 2) : (n >= 4);
}
__host__ __device__ constexpr double hex8_gauss(int i) {
  return i == 0 ? 0.21132486540518711775 : 0.78867513459481288225;
}
__host__ __device__ constexpr int hex8_qbit(int q, int d) {
  return d == 0 ? (q >> 2) & 1 : d == 1 ? (q >> 1) & 1 : q & 1;
}
__host__ __device__ constexpr double hex8_l(int bit, double t) { return bit ? t : 1.0 - t; }
__host__ __device__ constexpr double hex8_dl(int bit) { return bit ? 1.0 : -1.0; }
__host__ __device__ constexpr double hex8_dN(int q, int n, int d) {
  double v = 1.0;
  for (int e = 0; e < 3; ++e) {
    const int b = hex8_node_bit(n, e);
    const double t = hex8_gauss(hex8_qbit(q, e));
    v *= (e == d) ? hex8_dl(b) : hex8_l(b, t);
  }
  return v;
}

// ---------------------------------------------------------------------------
// Per-quadrature geometry: physical gradients g[n][3] and weight w.
template <typename T>
__device__ __forceinline__ void hex8_geometry(int q, const T X[8][3], T g[8][3], T& w) {
  T Jm[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      T s = 0;
#pragma unroll
      for (int n = 0; n < 8; ++n) s += X[n][d] * (T)hex8_dN(q, n, e);
      Jm[d][e] = s;
    }
  T inv[3][3];
  const T det = inv3(Jm, inv);  // inv = (dxi/dx)
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T s = 0;
#pragma unroll
      for (int e = 0; e < 3; ++e) s += (T)hex8_dN(q, n, e) * inv[e][j];
      g[n][j] = s;
    }
  w = det * (T)0.125;
}

template <typename T, int VEC>
__device__ __forceinline__ void load_cell(const double* __restrict__ pts, const int32_t* __restrict__ cells,
                                          const double* __restrict__ u, int64_t c, T X[8][3], T U[8][VEC]) {
  const int4* cp = reinterpret_cast<const int4*>(cells + c * 8);
  const int4 c0 = cp[0], c1 = cp[1];
  const int nid[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
  for (int n = 0; n < 8; ++n) {
#pragma unroll
    for (int d = 0; d < 3; ++d) X[n][d] = (T)pts[3 * (int64_t)nid[n] + d];
#pragma unroll
    for (int i = 0; i < VEC; ++i) U[n][i] = (T)u[VEC * (int64_t)nid[n] + i];
  }
}

// ---------------------------------------------------------------------------
template <typename T, int MAT, int VEC>
__global__ void __launch_bounds__(256) hex8_residual_kernel(
    const double* __restrict__ pts, const int32_t* __restrict__ cells, int64_t ncell,
    const double* __restrict__ u, const double* __restrict__ params, hf_material mat,
    T* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  T X[8][3], U[8][VEC];
  load_cell<T, VEC>(pts, cells, u, c, X, U);
  T r[8][VEC];
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int i = 0; i < VEC; ++i) r[n][i] = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    T g[8][3], w;
    hex8_geometry<T>(q, X, g, w);
    T ug[VEC][3];
    grad_u<T, 8, VEC>(U, g, ug);
    Moduli md = moduli(mat, params ? params[c * 8 + q] : 1.0, params != nullptr);
    T P[VEC][3];
    material_stress<T, MAT, VEC>(ug, (T)md.m1, (T)md.m2, P);
#pragma unroll
    for (int n = 0; n < 8; ++n)
#pragma unroll
      for (int i = 0; i < VEC; ++i)
        r[n][i] += w * (P[i][0] * g[n][0] + P[i][1] * g[n][1] + P[i][2] * g[n][2]);
  }
  T* o = out + c * (8 * VEC);
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int i = 0; i < VEC; ++i) o[n * VEC + i] = r[n][i];
}

// Element Jacobian K[a][b][i][k] = dR_(a,i) / du_(b,k), split into what a
// quadrature point shares between the 8 rows of its cell (geometry, state,
// F g_b and H g_b) and the contraction of one row.
template <typename T>
struct QpShared {
  T g[8][3];    // physical shape gradients
  T fg[8][3];   // F g_b (Neo-Hookean)
  T hg[8][3];   // H g_b (Neo-Hookean)
  T s1, s2, J, I1;  // NH: mu c w, kappa w, J, I1; linear elastic: mu w, lam w; Poisson: k w
};

template <typename T, int MAT, int VEC>
__device__ __forceinline__ void qp_shared(const T X[8][3], const T U[8][VEC], const double* __restrict__ prm,
                                          const hf_material& mat, int q, QpShared<T>& d) {
  T g[8][3], w;
  hex8_geometry<T>(q, X, g, w);
  Moduli md = moduli(mat, prm ? prm[q] : 1.0, prm != nullptr);
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int j = 0; j < 3; ++j) d.g[n][j] = g[n][j];
  if constexpr (MAT == MAT_NEO_HOOKEAN) {
    T ug[3][3];
    grad_u<T, 8, 3>(U, g, ug);
    NHState<T> s;
    nh_state<T>(ug, s);
    d.s1 = (T)md.m1 * s.c * w;
    d.s2 = (T)md.m2 * w;
    d.J = s.J;
    d.I1 = s.I1;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      matvec3(s.F, g[n], d.fg[n]);
      matvec3(s.H, g[n], d.hg[n]);
    }
  } else if constexpr (MAT == MAT_LINEAR_ELASTIC) {
    d.s1 = (T)md.m1 * w;
    d.s2 = (T)md.m2 * w;
  } else {
    d.s1 = (T)md.m1 * w;
  }
}

template <typename T, int MAT, int VEC>
__device__ __forceinline__ void k_row_accumulate(const QpShared<T>& d, int a, T K[8][VEC][VEC]) {
  T ga[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) ga[j] = d.g[a][j];
  if constexpr (MAT == MAT_NEO_HOOKEAN) {
    const T s1 = d.s1, s2 = d.s2;
    T Fga[3], Hga[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      Fga[j] = d.fg[a][j];
      Hga[j] = d.hg[a][j];
    }
    T al[3], be[3], ga_[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      al[i] = s1 * ((T)(-2.0 / 3.0) * Fga[i] + (T)(2.0 / 9.0) * d.I1 * Hga[i]) + s2 * ((T)2 * d.J - (T)1) * d.J * Hga[i];
      be[i] = (T)(-2.0 / 3.0) * s1 * Hga[i];
      ga_[i] = (s1 * d.I1 * (T)(1.0 / 3.0) - s2 * d.J * (d.J - (T)1)) * Hga[i];
    }
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const T gab = s1 * (ga[0] * d.g[b][0] + ga[1] * d.g[b][1] + ga[2] * d.g[b][2]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
          K[b][i][k] += al[i] * d.hg[b][k] + be[i] * d.fg[b][k] + d.hg[b][i] * ga_[k] + (i == k ? gab : (T)0);
    }
  } else if constexpr (MAT == MAT_LINEAR_ELASTIC) {
    const T lw = d.s2, mw = d.s1;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const T gab = mw * (ga[0] * d.g[b][0] + ga[1] * d.g[b][1] + ga[2] * d.g[b][2]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k)
          K[b][i][k] += lw * ga[i] * d.g[b][k] + mw * ga[k] * d.g[b][i] + (i == k ? gab : (T)0);
    }
  } else {  // Poisson, VEC == 1
    const T kw = d.s1;
#pragma unroll
    for (int b = 0; b < 8; ++b) K[b][0][0] += kw * (ga[0] * d.g[b][0] + ga[1] * d.g[b][1] + ga[2] * d.g[b][2]);
  }
}

// Element Jacobians, 8 cells per wavefront: lane (cell, q) forms quadrature
// point q's shared data in LDS, then lane (cell, a) contracts row a over the 8
// points (each point's geometry and state computed once per cell, not once per
// row).  Output block-major (cell, a, b, i, k): the 8 lanes of a cell write one
// contiguous element matrix.
template <typename T>
constexpr int jac_waves() { return sizeof(T) == 4 ? 2 : 1; }   // LDS: 64 QpShared per wave

// UPPER: only the blocks (a, b) whose column node id is >= the row node's (the
// diagonal and upper blocks a lattice stencil keeps), cell-innermost:
// out[((a * 8 + b) * VEC*VEC + VEC*i + k) * ncell + c], so that threads of
// consecutive nodes read consecutive cells; the other blocks are not written.
template <typename T, int MAT, int VEC, bool UPPER>
__global__ void __launch_bounds__(64 * jac_waves<T>()) hex8_jacobian_kernel(
    const double* __restrict__ pts, const int32_t* __restrict__ cells, int64_t ncell,
    const double* __restrict__ u, const double* __restrict__ params, hf_material mat,
    T* __restrict__ out) {
  __shared__ QpShared<T> sh[jac_waves<T>()][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ci = lane >> 3, r = lane & 7;
  const int64_t c = ((int64_t)blockIdx.x * jac_waves<T>() + wv) * 8 + ci;
  if (c < ncell) {
    T X[8][3], U[8][VEC];
    load_cell<T, VEC>(pts, cells, u, c, X, U);
    qp_shared<T, MAT, VEC>(X, U, params ? params + c * 8 : nullptr, mat, r, sh[wv][lane]);
  }
  wave_lds_sync();
  if (c >= ncell) return;
  T K[8][VEC][VEC];
#pragma unroll
  for (int b = 0; b < 8; ++b)
#pragma unroll
    for (int i = 0; i < VEC; ++i)
#pragma unroll
      for (int k = 0; k < VEC; ++k) K[b][i][k] = 0;
#pragma unroll 1
  for (int q = 0; q < 8; ++q) k_row_accumulate<T, MAT, VEC>(sh[wv][ci * 8 + q], r, K);
  if constexpr (UPPER) {
    const int32_t* cp = cells + c * 8;
    const int32_t nr = cp[r];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      if (cp[b] < nr) continue;
      T* o = out + (int64_t)((r * 8 + b) * VEC * VEC) * ncell + c;
#pragma unroll
      for (int i = 0; i < VEC; ++i)
#pragma unroll
        for (int k = 0; k < VEC; ++k) o[(int64_t)(i * VEC + k) * ncell] = K[b][i][k];
    }
  } else {
    T* o = out + (c * 8 + r) * (8 * VEC * VEC);
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int i = 0; i < VEC; ++i)
#pragma unroll
        for (int k = 0; k < VEC; ++k) o[(b * VEC + i) * VEC + k] = K[b][i][k];
  }
}

// grad[c,q] = - w sum_ij (sum_a lam[a][i] g_a[j]) dP_ij/dp
template <int MAT, int VEC>
__global__ void __launch_bounds__(256) hex8_param_vjp_kernel(
    const double* __restrict__ pts, const int32_t* __restrict__ cells, int64_t ncell,
    const double* __restrict__ u, const double* __restrict__ params, hf_material mat,
    const double* __restrict__ lam, double* __restrict__ grad) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double X[8][3], U[8][VEC], L[8][VEC];
  load_cell<double, VEC>(pts, cells, u, c, X, U);
  {
    const int32_t* cp = cells + c * 8;
#pragma unroll
    for (int n = 0; n < 8; ++n)
#pragma unroll
      for (int i = 0; i < VEC; ++i) L[n][i] = lam[VEC * (int64_t)cp[n] + i];
  }
#pragma unroll 1
  for (int q = 0; q < 8; ++q) {
    double g[8][3], w;
    hex8_geometry<double>(q, X, g, w);
    double ug[VEC][3], lg[VEC][3];
    grad_u<double, 8, VEC>(U, g, ug);
    grad_u<double, 8, VEC>(L, g, lg);
    Moduli md = moduli(mat, params[c * 8 + q], true);
    double P1[VEC][3], P2[VEC][3];
    material_stress_parts<double, MAT, VEC>(ug, P1, P2);
    double s = 0;
#pragma unroll
    for (int i = 0; i < VEC; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) s += lg[i][j] * (md.d1 * P1[i][j] + md.d2 * P2[i][j]);
    grad[c * 8 + q] = -w * s;
  }
}

// ---------------------------------------------------------------------------
// Element Jacobians assembled straight into the sliced-ELL matrix values (HEX8,
// vec 3), without the element-matrix array: 8 lanes per matrix row (node n of a
// slice row), lane c of the group takes the c-th cell holding n (the cells are the
// contributions of n's diagonal block, in cell order), computes row a = (n's local
// index) of that cell's Jacobian and adds its 8 blocks into the row's accumulators
// in LDS, one cell after the other in ascending cell order - the summation order
// of the two-stage path (hf_cell_jacobian + hf_assemble), which gives the same bits.
struct FusedArgs {
  const int32_t *rowptr, *colidx, *seg, *contrib, *diag, *slice_ptr, *sell_row;
  const uint8_t* cslot;  // per (cell, a): slots of the 8 column nodes in row cells[c,a]
  int64_t npos;  // slice rows (nslices * R)
  int R, max_row;
};

// HEX8 assembly of element matrices into the SELL values, 8 lanes per slice row
// (one per cell of the node, the cells in ascending order as the pattern's
// contribution segments list them): each lane reads its cell's row a of
// blocks (288 contiguous bytes in FP32), the row's slots are accumulated in
// LDS in 8 rounds of ascending cell index (the summation order of
// hf_assemble_*), then written column-major to the row's SELL slots.
template <typename T>
__global__ void __launch_bounds__(256) hex8_assemble_rows(FusedArgs fa, const T* __restrict__ elem,
                                                          T* __restrict__ vals) {
  extern __shared__ unsigned char hf_smem[];
  T* acc = reinterpret_cast<T*>(hf_smem);  // [32 rows][max_row][9], row-major blocks
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ln = lane >> 3, cl = lane & 7;
  const int64_t q = ((int64_t)blockIdx.x * 4 + wv) * 8 + ln;
  T* my = acc + (int64_t)(wv * 8 + ln) * fa.max_row * 9;
  const int32_t n = q < fa.npos ? fa.sell_row[q] : -1;
  int32_t len = 0, ncl = 0, s0 = 0;
  if (n >= 0) {
    len = fa.rowptr[n + 1] - fa.rowptr[n];
    const int32_t db = fa.diag[n];
    if (db >= 0) {
      s0 = fa.seg[db];
      ncl = fa.seg[db + 1] - s0;
    }
  }
  for (int e = cl; e < len * 9; e += 8) my[e] = (T)0;
  wave_lds_sync();
  for (int cbase = 0; __ballot(ncl > cbase) != 0; cbase += 8) {
    const int ci = cbase + cl;
    const bool mine = ci < ncl;
    T K[8][9];
    int jslot[8];
    if (mine) {
      const int32_t id = fa.contrib[s0 + ci];
      const int64_t c = id / 64;
      const int a = (id / 8) % 8;
      const uint2 sl = *reinterpret_cast<const uint2*>(fa.cslot + (c * 8 + a) * 8);
#pragma unroll
      for (int b = 0; b < 8; ++b) jslot[b] = (int)(((b < 4 ? sl.x : sl.y) >> (8 * (b & 3))) & 0xffu);
      const T* src = elem + (c * 8 + a) * 72;
#pragma unroll
      for (int b = 0; b < 8; ++b)
#pragma unroll
        for (int e = 0; e < 9; ++e) K[b][e] = src[b * 9 + e];
    }
    for (int round = 0; round < 8; ++round) {  // cells in ascending order
      if (mine && cl == round) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          if (jslot[b] < len) {
            T* dst = my + jslot[b] * 9;
#pragma unroll
            for (int e = 0; e < 9; ++e) dst[e] += K[b][e];
          }
        }
      }
      wave_lds_sync();
    }
  }
  if (q < fa.npos) {
    const int64_t sl = q / fa.R;
    const int r = (int)(q - sl * fa.R);
    const int32_t t0 = fa.slice_ptr[sl], t1 = fa.slice_ptr[sl + 1];
    for (int32_t j = cl; j < t1 - t0; j += 8) {
      T* dst = vals + ((int64_t)(t0 + j) * fa.R + r) * 9;
#pragma unroll
      for (int e = 0; e < 9; ++e) dst[(e % 3) * 3 + e / 3] = j < len ? my[j * 9 + e] : (T)0;  // column-major
    }
  }
}

template <typename T>
static int launch_hex8_assemble_rows(const hf_pattern* p, const T* elem, T* vals, hipStream_t s) {
  HF_REQUIRE(pattern_vec(p) == 3 && pattern_nn(p) == 8, "hf_assemble_hex8: needs a HEX8 vec-3 pattern");
  FusedArgs fa;
  fa.rowptr = pattern_rowptr(p);
  fa.colidx = pattern_colidx(p);
  fa.cslot = pattern_cslot(p);
  HF_REQUIRE(fa.cslot != nullptr, "hf_assemble_hex8: the pattern has no HEX8 slot table");
  fa.seg = pattern_seg(p);
  fa.contrib = pattern_contrib(p);
  fa.diag = pattern_diag(p);
  fa.slice_ptr = pattern_slice_ptr(p);
  fa.sell_row = pattern_sell_row(p);
  fa.R = pattern_R(p);
  fa.npos = pattern_nslices(p) * fa.R;
  fa.max_row = pattern_max_row_blocks(p);
  const size_t smem = (size_t)32 * fa.max_row * 9 * sizeof(T);
  HF_REQUIRE(smem <= 96 * 1024, "hf_assemble_hex8: rows of %d blocks exceed the LDS budget", fa.max_row);
  const unsigned grid = (unsigned)ceil_div(fa.npos, 32);
  hex8_assemble_rows<T><<<grid, 256, smem, s>>>(fa, elem, vals);
  HF_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------------------
// Host dispatch
template <typename T>
static int launch_hex8_residual(const hf_material* mat, const double* pts, const int32_t* cells,
                                int64_t ncell, const double* u, const double* params, T* out,
                                hipStream_t s) {
  if (ncell == 0) return 0;
  const dim3 blk(256), grd((unsigned)ceil_div(ncell, 256));
  switch (mat->kind) {
    case MAT_POISSON:
      hex8_residual_kernel<T, MAT_POISSON, 1><<<grd, blk, 0, s>>>(pts, cells, ncell, u, params, *mat, out);
      break;
    case MAT_LINEAR_ELASTIC:
      hex8_residual_kernel<T, MAT_LINEAR_ELASTIC, 3><<<grd, blk, 0, s>>>(pts, cells, ncell, u, params, *mat, out);
      break;
    case MAT_NEO_HOOKEAN:
      hex8_residual_kernel<T, MAT_NEO_HOOKEAN, 3><<<grd, blk, 0, s>>>(pts, cells, ncell, u, params, *mat, out);
      break;
    default:
      HF_REQUIRE(false, "unknown material kind %d", mat->kind);
  }
  HF_LAUNCH_CHECK();
  return 0;
}

template <typename T, bool UPPER = false>
static int launch_hex8_jacobian(const hf_material* mat, const double* pts, const int32_t* cells,
                                int64_t ncell, const double* u, const double* params, T* out,
                                hipStream_t s) {
  if (ncell == 0) return 0;
  const dim3 blk(64 * jac_waves<T>()), grd((unsigned)ceil_div(ncell, 8 * jac_waves<T>()));
  switch (mat->kind) {
    case MAT_POISSON:
      hex8_jacobian_kernel<T, MAT_POISSON, 1, UPPER><<<grd, blk, 0, s>>>(pts, cells, ncell, u, params, *mat, out);
      break;
    case MAT_LINEAR_ELASTIC:
      hex8_jacobian_kernel<T, MAT_LINEAR_ELASTIC, 3, UPPER><<<grd, blk, 0, s>>>(pts, cells, ncell, u, params, *mat,
                                                                                out);
      break;
    case MAT_NEO_HOOKEAN:
      hex8_jacobian_kernel<T, MAT_NEO_HOOKEAN, 3, UPPER><<<grd, blk, 0, s>>>(pts, cells, ncell, u, params, *mat,
                                                                             out);
      break;
    default:
      HF_REQUIRE(false, "unknown material kind %d", mat->kind);
  }
  HF_LAUNCH_CHECK();
  return 0;
}

int hex8_param_vjp(const hf_material* mat, const double* pts, const int32_t* cells, int64_t ncell,
                   const double* u, const double* params, const double* lam, double* grad,
                   hipStream_t s) {
  if (ncell == 0) return 0;
  const dim3 blk(256), grd((unsigned)ceil_div(ncell, 256));
  switch (mat->kind) {
    case MAT_POISSON:
      hex8_param_vjp_kernel<MAT_POISSON, 1><<<grd, blk, 0, s>>>(pts, cells, ncell, u, params, *mat, lam, grad);
      break;
    case MAT_LINEAR_ELASTIC:
      hex8_param_vjp_kernel<MAT_LINEAR_ELASTIC, 3><<<grd, blk, 0, s>>>(pts, cells, ncell, u, params, *mat, lam, grad);
      break;
    case MAT_NEO_HOOKEAN:
      hex8_param_vjp_kernel<MAT_NEO_HOOKEAN, 3><<<grd, blk, 0, s>>>(pts, cells, ncell, u, params, *mat, lam, grad);
      break;
    default:
      HF_REQUIRE(false, "unknown material kind %d", mat->kind);
  }
  HF_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------------------
// Box lattices of identical cells (a box_mesh: every cell a translate of cell 0,
// nodes in one corner order): the physical shape gradients and weights at the
// 8 points are the same for every cell, so they are formed once from cell 0
// (hex8_box_geometry, the same arithmetic as hex8_geometry) and the residual
// kernel reads them as uniform values (scalar loads) instead of forming the
// 3x3 Jacobian, its inverse and the gradients per cell and point.
struct Hex8BoxGeom {
  double g[8][8][3];   // [q][n][j]
  double w[8];
};

__global__ void hex8_box_geometry_kernel(const double* __restrict__ pts, const int32_t* __restrict__ cells,
                                         Hex8BoxGeom* __restrict__ geo) {
  const int q = threadIdx.x;
  if (q >= 8) return;
  double X[8][3];
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int64_t nid = cells[n];
#pragma unroll
    for (int d = 0; d < 3; ++d) X[n][d] = pts[3 * nid + d];
  }
  double g[8][3], w;
  hex8_geometry<double>(q, X, g, w);
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int j = 0; j < 3; ++j) geo->g[q][n][j] = g[n][j];
  geo->w[q] = w;
}

template <typename T, int MAT, int VEC>
__global__ void __launch_bounds__(256) hex8_residual_box_kernel(const Hex8BoxGeom* __restrict__ geo,
                                                                const int32_t* __restrict__ cells, int64_t ncell,
                                                                const double* __restrict__ u,
                                                                const double* __restrict__ params, hf_material mat,
                                                                T* __restrict__ out) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int4* cp = reinterpret_cast<const int4*>(cells + c * 8);
  const int4 c0 = cp[0], c1 = cp[1];
  const int nid[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
  T U[8][VEC];
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int i = 0; i < VEC; ++i) U[n][i] = (T)u[VEC * (int64_t)nid[n] + i];
  T r[8][VEC];
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int i = 0; i < VEC; ++i) r[n][i] = 0;
#pragma unroll 1
  for (int q = 0; q < 8; ++q) {
    T g[8][3];
#pragma unroll
    for (int n = 0; n < 8; ++n)
#pragma unroll
      for (int j = 0; j < 3; ++j) g[n][j] = (T)geo->g[q][n][j];
    const T w = (T)geo->w[q];
    T ug[VEC][3];
    grad_u<T, 8, VEC>(U, g, ug);
    Moduli md = moduli(mat, params ? params[c * 8 + q] : 1.0, params != nullptr);
    T P[VEC][3];
    material_stress<T, MAT, VEC>(ug, (T)md.m1, (T)md.m2, P);
#pragma unroll
    for (int n = 0; n < 8; ++n)
#pragma unroll
      for (int i = 0; i < VEC; ++i)
        r[n][i] += w * (P[i][0] * g[n][0] + P[i][1] * g[n][1] + P[i][2] * g[n][2]);
  }
  T* o = out + c * (8 * VEC);
#pragma unroll
  for (int n = 0; n < 8; ++n)
#pragma unroll
    for (int i = 0; i < VEC; ++i) o[n * VEC + i] = r[n][i];
}

int hex8_box_geometry(const double* p, const int32_t* c, void* geo, hipStream_t s) {
  hex8_box_geometry_kernel<<<1, 64, 0, s>>>(p, c, static_cast<Hex8BoxGeom*>(geo));
  HF_LAUNCH_CHECK();
  return 0;
}

int hex8_residual_box_f64(const hf_material* mat, const void* geo, const int32_t* cells, int64_t ncell,
                          const double* u, const double* params, double* out, hipStream_t s) {
  if (ncell == 0) return 0;
  const Hex8BoxGeom* G = static_cast<const Hex8BoxGeom*>(geo);
  const dim3 blk(256), grd((unsigned)ceil_div(ncell, 256));
  switch (mat->kind) {
    case MAT_POISSON:
      hex8_residual_box_kernel<double, MAT_POISSON, 1><<<grd, blk, 0, s>>>(G, cells, ncell, u, params, *mat, out);
      break;
    case MAT_LINEAR_ELASTIC:
      hex8_residual_box_kernel<double, MAT_LINEAR_ELASTIC, 3><<<grd, blk, 0, s>>>(G, cells, ncell, u, params, *mat,
                                                                                  out);
      break;
    case MAT_NEO_HOOKEAN:
      hex8_residual_box_kernel<double, MAT_NEO_HOOKEAN, 3><<<grd, blk, 0, s>>>(G, cells, ncell, u, params, *mat,
                                                                               out);
      break;
    default:
      HF_REQUIRE(false, "unknown material kind %d", mat->kind);
  }
  HF_LAUNCH_CHECK();
  return 0;
}

int hex8_residual_f64(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                      const double* prm, double* o, hipStream_t s) {
  return launch_hex8_residual<double>(m, p, c, n, u, prm, o, s);
}
int hex8_residual_f32(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                      const double* prm, float* o, hipStream_t s) {
  return launch_hex8_residual<float>(m, p, c, n, u, prm, o, s);
}
int hex8_jacobian_f64(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                      const double* prm, double* o, hipStream_t s) {
  return launch_hex8_jacobian<double>(m, p, c, n, u, prm, o, s);
}
int hex8_jacobian_f32(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                      const double* prm, float* o, hipStream_t s) {
  return launch_hex8_jacobian<float>(m, p, c, n, u, prm, o, s);
}
int hex8_jacobian_upper_f32(const hf_material* m, const double* p, const int32_t* c, int64_t n, const double* u,
                            const double* prm, float* o, hipStream_t s) {
  return launch_hex8_jacobian<float, true>(m, p, c, n, u, prm, o, s);
}

}  // namespace hf

extern "C" int hf_assemble_hex8_f32(const hf_pattern* p, const float* elem, float* vals, hf_stream_t s) {
  HF_REQUIRE(p && elem && vals, "hf_assemble_hex8_f32: null argument");
  return hf::launch_hex8_assemble_rows<float>(p, elem, vals, (hipStream_t)s);
}
extern "C" int hf_assemble_hex8_f64(const hf_pattern* p, const double* elem, double* vals, hf_stream_t s) {
  HF_REQUIRE(p && elem && vals, "hf_assemble_hex8_f64: null argument");
  return hf::launch_hex8_assemble_rows<double>(p, elem, vals, (hipStream_t)s);
}
