// Topology-optimisation side of the adjoint: the sensitivity / density filters
// and the MMA subproblem (jax_fem/mma.py:27-66, :114-410) on the device, FP64.
//
// Filters: H is held as a fixed-width neighbour table (ELL): row i lists k
// cells nbr[i*k + t] with weights w[i*k + t] = max(rmin - |c_i - c_j|, 0)
// (k = 20 nearest centroids, built once on the host).  One thread per row.
//
// MMA subproblem (Svanberg 1987; the primal-dual interior-point solver of his
// 2007 notes, the reference's ``subsolv``): the n-sized state (x, xsi, eta and
// their Newton directions) stays in HBM; the m-sized state (y, lam, mu, s) and
// the scalars z, zet live in one device struct.  Every pass over n is one
// kernel that also produces fixed-order per-workgroup partials; a one-workgroup
// kernel finishes the sums, solves the reduced (m+1)x(m+1) Newton system and
// updates the small state.  The host drives the three loops of the method
// (epsi levels, Newton steps, residual-norm backtracking), reading back two
// numbers per residual evaluation.
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "hipfem.h"

namespace hf {
namespace {

constexpr int kMmaBlock = 256;
constexpr int kMmaWaves = kMmaBlock / kWave;
constexpr int kMmaMaxGrid = 1024;
constexpr int kMmaMaxM = 4;

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Per-workgroup partials of NV values: entry j is a sum, or a max where (maxmask >> j) & 1.
template <int NV>
__device__ void wg_partials(double (&v)[NV], unsigned maxmask, double* __restrict__ partials) {
  __shared__ double sm[kMmaWaves][NV];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const double r = ((maxmask >> j) & 1u) ? wave_max(v[j]) : wave_sum(v[j]);
    if (lane == 0) sm[w][j] = r;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double r = sm[0][j];
      for (int k = 1; k < kMmaWaves; ++k) r = ((maxmask >> j) & 1u) ? fmax(r, sm[k][j]) : r + sm[k][j];
      partials[(int64_t)blockIdx.x * NV + j] = r;
    }
  }
}

// Totals of nparts partials (same sum / max mask), fixed order; every thread gets them.
template <int NV>
__device__ void wg_totals(const double* __restrict__ partials, int nparts, unsigned maxmask, double (&tot)[NV]) {
  __shared__ double sm[kMmaWaves][NV];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  double acc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = ((maxmask >> j) & 1u) ? -INFINITY : 0.0;
  for (int b = threadIdx.x; b < nparts; b += kMmaBlock)
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const double p = partials[(int64_t)b * NV + j];
      acc[j] = ((maxmask >> j) & 1u) ? fmax(acc[j], p) : acc[j] + p;
    }
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const double r = ((maxmask >> j) & 1u) ? wave_max(acc[j]) : wave_sum(acc[j]);
    if (lane == 0) sm[w][j] = r;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double r = sm[0][j];
    for (int k = 1; k < kMmaWaves; ++k) r = ((maxmask >> j) & 1u) ? fmax(r, sm[k][j]) : r + sm[k][j];
    tot[j] = r;
  }
  __syncthreads();
}

// The subproblem's small state and data.
struct MmaSmall {
  double y[kMmaMaxM], lam[kMmaMaxM], mu[kMmaMaxM], s[kMmaMaxM], z, zet;
  double yo[kMmaMaxM], lamo[kMmaMaxM], muo[kMmaMaxM], so[kMmaMaxM], zo, zeto;
  double dy[kMmaMaxM], dlam[kMmaMaxM], dmu[kMmaMaxM], ds[kMmaMaxM], dz, dzet;
  double a[kMmaMaxM], b[kMmaMaxM], c[kMmaMaxM], d[kMmaMaxM], a0;
  double steg;
};

struct MmaResult {        // written by the residual finish, read by the host
  double norm, maxabs;
};

// n-sized arrays of one subproblem.
struct MmaVecs {
  int64_t n;
  double *low, *upp, *alfa, *beta, *p0, *q0, *P, *Q;       // P, Q: (m, n) row-major
  double *x, *xsi, *eta, *xo, *xsio, *etao, *dx, *dxsi, *deta;
};

__device__ inline void chunk(int64_t n, int64_t& i0, int64_t& i1) {
  const int64_t per = ceil_div(n, gridDim.x);
  i0 = (int64_t)blockIdx.x * per;
  i1 = i0 + per < n ? i0 + per : n;
}

// Asymptotes, move limits and the convex approximation (mma.py:135-201); x, xsi,
// eta at their starting values (mma.py:208-222).  Partials: P/(U-x) + Q/(x-L) per row.
template <int M>
__global__ void __launch_bounds__(kMmaBlock) mma_setup(MmaVecs v, int iter, const double* __restrict__ xval,
                                                       const double* __restrict__ xold1,
                                                       const double* __restrict__ xold2,
                                                       const double* __restrict__ xmin, const double* __restrict__ xmax,
                                                       const double* __restrict__ df0dx,
                                                       const double* __restrict__ dfdx, double move,
                                                       double* __restrict__ partials) {
  int64_t i0, i1;
  chunk(v.n, i0, i1);
  double loc[M];
#pragma unroll
  for (int k = 0; k < M; ++k) loc[k] = 0;
  for (int64_t j = i0 + threadIdx.x; j < i1; j += kMmaBlock) {
    const double xv = xval[j], box = xmax[j] - xmin[j];
    double lo, up;
    if (iter <= 2) {
      lo = xv - 0.5 * box;
      up = xv + 0.5 * box;
    } else {
      const double x1 = xold1[j];
      const double trend = (xv - x1) * (x1 - xold2[j]);
      const double f = trend > 0 ? 1.2 : (trend < 0 ? 0.7 : 1.0);
      lo = xv - f * (x1 - v.low[j]);
      up = xv + f * (v.upp[j] - x1);
      lo = fmin(fmax(lo, xv - 10.0 * box), xv - 0.01 * box);
      up = fmax(fmin(up, xv + 10.0 * box), xv + 0.01 * box);
    }
    v.low[j] = lo;
    v.upp[j] = up;
    const double al = fmax(fmax(lo + 0.1 * (xv - lo), xv - move * box), xmin[j]);
    const double be = fmin(fmin(up - 0.1 * (up - xv), xv + move * box), xmax[j]);
    v.alfa[j] = al;
    v.beta[j] = be;
    const double inv_box = 1.0 / fmax(box, 1e-5);
    const double ux1 = up - xv, xl1 = xv - lo;
    const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    const double g = df0dx[j];
    const double pos = fmax(g, 0.0), neg = fmax(-g, 0.0);
    const double reg = 0.001 * (pos + neg) + 1e-5 * inv_box;
    v.p0[j] = (pos + reg) * ux2;
    v.q0[j] = (neg + reg) * xl2;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double gk = dfdx[(int64_t)k * v.n + j];
      const double pk = fmax(gk, 0.0), qk = fmax(-gk, 0.0);
      const double rk = 0.001 * (pk + qk) + 1e-5 * inv_box;
      const double Pk = (pk + rk) * ux2, Qk = (qk + rk) * xl2;
      v.P[(int64_t)k * v.n + j] = Pk;
      v.Q[(int64_t)k * v.n + j] = Qk;
      loc[k] += Pk / ux1 + Qk / xl1;
    }
    const double x0 = 0.5 * (al + be);
    v.x[j] = x0;
    v.xsi[j] = fmax(1.0 / (x0 - al), 1.0);
    v.eta[j] = fmax(1.0 / (be - x0), 1.0);
  }
  wg_partials<M>(loc, 0u, partials);
}

template <int M>
__global__ void __launch_bounds__(kMmaBlock) mma_setup_finish(MmaSmall* sm, const double* __restrict__ partials,
                                                              int nparts, const double* __restrict__ fval) {
  double tot[M];
  wg_totals<M>(partials, nparts, 0u, tot);
  if (threadIdx.x != 0) return;
  for (int k = 0; k < M; ++k) {
    sm->b[k] = tot[k] - fval[k];
    sm->y[k] = 1.0;
    sm->lam[k] = 1.0;
    sm->s[k] = 1.0;
    sm->mu[k] = fmax(1.0, 0.5 * sm->c[k]);
  }
  sm->z = 1.0;
  sm->zet = 1.0;
}

// Residual of the perturbed KKT system (mma.py:224-244, :337-364): the n-sized
// parts here.  With trial set, first x = xo + steg dx (xsi, eta likewise) and the
// multipliers at the trial point.  Partials: [sum of squares, max |.|, gvec_k].
template <int M>
__global__ void __launch_bounds__(kMmaBlock) mma_residual(MmaVecs v, const MmaSmall* __restrict__ sm, double epsi,
                                                          int trial, double* __restrict__ partials) {
  int64_t i0, i1;
  chunk(v.n, i0, i1);
  double lam[M];
  const double steg = sm->steg;
#pragma unroll
  for (int k = 0; k < M; ++k) lam[k] = trial ? sm->lamo[k] + steg * sm->dlam[k] : sm->lam[k];
  double loc[2 + M];
#pragma unroll
  for (int k = 0; k < 2 + M; ++k) loc[k] = 0;
  for (int64_t j = i0 + threadIdx.x; j < i1; j += kMmaBlock) {
    double x, xsi, eta;
    if (trial) {
      x = v.xo[j] + steg * v.dx[j];
      xsi = v.xsio[j] + steg * v.dxsi[j];
      eta = v.etao[j] + steg * v.deta[j];
      v.x[j] = x;
      v.xsi[j] = xsi;
      v.eta[j] = eta;
    } else {
      x = v.x[j];
      xsi = v.xsi[j];
      eta = v.eta[j];
    }
    const double ux1 = v.upp[j] - x, xl1 = x - v.low[j];
    double plam = v.p0[j], qlam = v.q0[j];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double Pk = v.P[(int64_t)k * v.n + j], Qk = v.Q[(int64_t)k * v.n + j];
      plam += Pk * lam[k];
      qlam += Qk * lam[k];
      loc[2 + k] += Pk / ux1 + Qk / xl1;
    }
    const double rex = plam / (ux1 * ux1) - qlam / (xl1 * xl1) - xsi + eta;
    const double rexsi = xsi * (x - v.alfa[j]) - epsi;
    const double reeta = eta * (v.beta[j] - x) - epsi;
    loc[0] += rex * rex + rexsi * rexsi + reeta * reeta;
    loc[1] = fmax(loc[1], fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
  }
  wg_partials<2 + M>(loc, 2u, partials);
}

// The m-sized residual parts; with trial, the small state moves to old + steg d
// first and steg halves afterwards (mma.py:326-370).
template <int M>
__global__ void __launch_bounds__(kMmaBlock) mma_residual_finish(MmaSmall* sm, const double* __restrict__ partials,
                                                                 int nparts, double epsi, int trial,
                                                                 MmaResult* out) {
  double tot[2 + M];
  wg_totals<2 + M>(partials, nparts, 2u, tot);
  if (threadIdx.x != 0) return;
  const double steg = sm->steg;
  if (trial) {
    for (int k = 0; k < M; ++k) {
      sm->y[k] = sm->yo[k] + steg * sm->dy[k];
      sm->lam[k] = sm->lamo[k] + steg * sm->dlam[k];
      sm->mu[k] = sm->muo[k] + steg * sm->dmu[k];
      sm->s[k] = sm->so[k] + steg * sm->ds[k];
    }
    sm->z = sm->zo + steg * sm->dz;
    sm->zet = sm->zeto + steg * sm->dzet;
    sm->steg = 0.5 * steg;
  }
  double ss = tot[0], mx = tot[1];
  double alam = 0;
  for (int k = 0; k < M; ++k) alam += sm->a[k] * sm->lam[k];
  const double rez = sm->a0 - sm->zet - alam;
  const double rezet = sm->zet * sm->z - epsi;
  ss += rez * rez + rezet * rezet;
  mx = fmax(mx, fmax(fabs(rez), fabs(rezet)));
  for (int k = 0; k < M; ++k) {
    const double rey = sm->c[k] + sm->d[k] * sm->y[k] - sm->mu[k] - sm->lam[k];
    const double relam = tot[2 + k] - sm->a[k] * sm->z - sm->y[k] + sm->s[k] - sm->b[k];
    const double remu = sm->mu[k] * sm->y[k] - epsi;
    const double res = sm->lam[k] * sm->s[k] - epsi;
    ss += rey * rey + relam * relam + remu * remu + res * res;
    mx = fmax(mx, fmax(fmax(fabs(rey), fabs(relam)), fmax(fabs(remu), fabs(res))));
  }
  out->norm = sqrt(ss);
  out->maxabs = mx;
}

template <int M>
struct NewtonNV {
  static constexpr int NA = M * (M + 1) / 2;
  static constexpr int NV = 2 * M + NA;   // gvec, GG (delx/diagx), GG diag(1/diagx) GG^T (upper)
};

// Newton step, first pass over n (mma.py:250-283): partials of gvec, of
// GG (delx / diagx) and of the reduced matrix GG diag(1/diagx) GG^T.
template <int M>
__global__ void __launch_bounds__(kMmaBlock) mma_newton_a(MmaVecs v, const MmaSmall* __restrict__ sm, double epsi,
                                                          double* __restrict__ partials) {
  using NV = NewtonNV<M>;
  int64_t i0, i1;
  chunk(v.n, i0, i1);
  double lam[M];
#pragma unroll
  for (int k = 0; k < M; ++k) lam[k] = sm->lam[k];
  double loc[NV::NV];
#pragma unroll
  for (int k = 0; k < NV::NV; ++k) loc[k] = 0;
  for (int64_t j = i0 + threadIdx.x; j < i1; j += kMmaBlock) {
    const double x = v.x[j], al = v.alfa[j], be = v.beta[j];
    const double ux1 = v.upp[j] - x, xl1 = x - v.low[j];
    const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    double plam = v.p0[j], qlam = v.q0[j], GG[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double Pk = v.P[(int64_t)k * v.n + j], Qk = v.Q[(int64_t)k * v.n + j];
      plam += Pk * lam[k];
      qlam += Qk * lam[k];
      loc[k] += Pk / ux1 + Qk / xl1;
      GG[k] = Pk / ux2 - Qk / xl2;
    }
    const double delx = plam / ux2 - qlam / xl2 - epsi / (x - al) + epsi / (be - x);
    const double diagx = 2.0 * (plam / (ux1 * ux2) + qlam / (xl1 * xl2)) + v.xsi[j] / (x - al) + v.eta[j] / (be - x);
    const double dxd = delx / diagx;
#pragma unroll
    for (int k = 0; k < M; ++k) loc[M + k] += GG[k] * dxd;
    int e = 2 * M;
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int l = k; l < M; ++l) loc[e++] += GG[k] * GG[l] / diagx;
  }
  wg_partials<NV::NV>(loc, 0u, partials);
}

// The reduced (m+1)x(m+1) system for (dlam, dz) and the m-sized directions
// (mma.py:284-314, :322-327); partial pivoting.
template <int M>
__global__ void __launch_bounds__(kMmaBlock) mma_newton_small(MmaSmall* sm, const double* __restrict__ partials,
                                                              int nparts, double epsi) {
  using NV = NewtonNV<M>;
  double tot[NV::NV];
  wg_totals<NV::NV>(partials, nparts, 0u, tot);
  if (threadIdx.x != 0) return;
  constexpr int K = M + 1;
  double A[K][K + 1];
  double alam = 0;
  for (int k = 0; k < M; ++k) alam += sm->a[k] * sm->lam[k];
  double dely[M], diagy[M];
  for (int k = 0; k < M; ++k) {
    dely[k] = sm->c[k] + sm->d[k] * sm->y[k] - sm->lam[k] - epsi / sm->y[k];
    diagy[k] = sm->d[k] + sm->mu[k] / sm->y[k];
    const double dellam = tot[k] - sm->a[k] * sm->z - sm->y[k] - sm->b[k] + epsi / sm->lam[k];
    const double diaglamyi = sm->s[k] / sm->lam[k] + 1.0 / diagy[k];
    A[k][K] = dellam + dely[k] / diagy[k] - tot[M + k];
    A[k][k] = diaglamyi;
    A[k][M] = sm->a[k];
    A[M][k] = sm->a[k];
  }
  int e = 2 * M;
  for (int k = 0; k < M; ++k)
    for (int l = k; l < M; ++l) {
      const double g = tot[e++];
      if (l == k) A[k][k] += g;
      else { A[k][l] = g; A[l][k] = g; }
    }
  A[M][M] = -sm->zet / sm->z;
  A[M][K] = sm->a0 - alam - epsi / sm->z;
  for (int c = 0; c < K; ++c) {        // Gaussian elimination, partial pivoting
    int piv = c;
    for (int r = c + 1; r < K; ++r)
      if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
    if (piv != c)
      for (int q = 0; q <= K; ++q) { const double t = A[c][q]; A[c][q] = A[piv][q]; A[piv][q] = t; }
    for (int r = c + 1; r < K; ++r) {
      const double f = A[r][c] / A[c][c];
      for (int q = c; q <= K; ++q) A[r][q] -= f * A[c][q];
    }
  }
  double sol[K];
  for (int r = K - 1; r >= 0; --r) {
    double t = A[r][K];
    for (int q = r + 1; q < K; ++q) t -= A[r][q] * sol[q];
    sol[r] = t / A[r][r];
  }
  for (int k = 0; k < M; ++k) {
    const double dlam = sol[k];
    const double dy = -dely[k] / diagy[k] + dlam / diagy[k];
    sm->dlam[k] = dlam;
    sm->dy[k] = dy;
    sm->dmu[k] = -sm->mu[k] + epsi / sm->y[k] - sm->mu[k] * dy / sm->y[k];
    sm->ds[k] = -sm->s[k] + epsi / sm->lam[k] - sm->s[k] * dlam / sm->lam[k];
  }
  const double dz = sol[M];
  sm->dz = dz;
  sm->dzet = -sm->zet + epsi / sm->z - sm->zet * dz / sm->z;
}

// Second pass over n: dx, dxsi, deta, the old point, and the largest step ratio
// (mma.py:315-335).  Partial: max ratio.
template <int M>
__global__ void __launch_bounds__(kMmaBlock) mma_newton_b(MmaVecs v, const MmaSmall* __restrict__ sm, double epsi,
                                                          double* __restrict__ partials) {
  int64_t i0, i1;
  chunk(v.n, i0, i1);
  double lam[M], dlam[M];
#pragma unroll
  for (int k = 0; k < M; ++k) { lam[k] = sm->lam[k]; dlam[k] = sm->dlam[k]; }
  double loc[1] = {-INFINITY};
  for (int64_t j = i0 + threadIdx.x; j < i1; j += kMmaBlock) {
    const double x = v.x[j], al = v.alfa[j], be = v.beta[j], xsi = v.xsi[j], eta = v.eta[j];
    const double ux1 = v.upp[j] - x, xl1 = x - v.low[j];
    const double ux2 = ux1 * ux1, xl2 = xl1 * xl1;
    double plam = v.p0[j], qlam = v.q0[j], gtd = 0;
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double Pk = v.P[(int64_t)k * v.n + j], Qk = v.Q[(int64_t)k * v.n + j];
      plam += Pk * lam[k];
      qlam += Qk * lam[k];
      gtd += (Pk / ux2 - Qk / xl2) * dlam[k];
    }
    const double delx = plam / ux2 - qlam / xl2 - epsi / (x - al) + epsi / (be - x);
    const double diagx = 2.0 * (plam / (ux1 * ux2) + qlam / (xl1 * xl2)) + xsi / (x - al) + eta / (be - x);
    const double dx = -delx / diagx - gtd / diagx;
    const double dxsi = -xsi + epsi / (x - al) - xsi * dx / (x - al);
    const double deta = -eta + epsi / (be - x) + eta * dx / (be - x);
    v.dx[j] = dx;
    v.dxsi[j] = dxsi;
    v.deta[j] = deta;
    v.xo[j] = x;
    v.xsio[j] = xsi;
    v.etao[j] = eta;
    const double r = fmax(fmax(-1.01 * dxsi / xsi, -1.01 * deta / eta),
                          fmax(-1.01 * dx / (x - al), 1.01 * dx / (be - x)));
    loc[0] = fmax(loc[0], r);
  }
  wg_partials<1>(loc, 1u, partials);
}

// steg = 1 / max(1, largest ratio) and the old small state (mma.py:328-352).
template <int M>
__global__ void __launch_bounds__(kMmaBlock) mma_step(MmaSmall* sm, const double* __restrict__ partials, int nparts) {
  double tot[1];
  wg_totals<1>(partials, nparts, 1u, tot);
  if (threadIdx.x != 0) return;
  double r = fmax(tot[0], 1.0);
  for (int k = 0; k < M; ++k) {
    r = fmax(r, -1.01 * sm->dy[k] / sm->y[k]);
    r = fmax(r, -1.01 * sm->dlam[k] / sm->lam[k]);
    r = fmax(r, -1.01 * sm->dmu[k] / sm->mu[k]);
    r = fmax(r, -1.01 * sm->ds[k] / sm->s[k]);
    sm->yo[k] = sm->y[k];
    sm->lamo[k] = sm->lam[k];
    sm->muo[k] = sm->mu[k];
    sm->so[k] = sm->s[k];
  }
  r = fmax(r, fmax(-1.01 * sm->dz / sm->z, -1.01 * sm->dzet / sm->zet));
  sm->zo = sm->z;
  sm->zeto = sm->zet;
  sm->steg = 1.0 / r;
}

// Filters (mma.py:58-66).
__global__ void sensitivity_filter_kernel(const int32_t* __restrict__ nbr, const double* __restrict__ w,
                                          const double* __restrict__ hs, int64_t n, int k,
                                          const double* __restrict__ rho, const double* __restrict__ g, int ncols,
                                          double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < ncols; ++c) {
    double acc = 0;
    for (int t = 0; t < k; ++t) {
      const int64_t j = nbr[i * k + t];
      const double r = rho[j];
      acc += w[i * k + t] * (r * g[(int64_t)c * n + j] / fmax(1e-3, r) / hs[j]);
    }
    out[(int64_t)c * n + i] = acc;
  }
}

__global__ void density_filter_kernel(const int32_t* __restrict__ nbr, const double* __restrict__ w,
                                      const double* __restrict__ hs, int64_t n, int k,
                                      const double* __restrict__ rho, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0;
  for (int t = 0; t < k; ++t) acc += w[i * k + t] * rho[nbr[i * k + t]];
  out[i] = acc / hs[i];
}

}  // namespace
}  // namespace hf

using namespace hf;

struct hf_mma {
  int64_t n;
  int m;
  int grid;
  MmaVecs v;
  double* partials;      // device, grid * max NV
  MmaSmall* small;       // device
  MmaResult* result;     // pinned host
  double* block;         // the one device allocation behind v
};

extern "C" int hf_mma_create(int64_t n, int32_t m, hf_mma** out) {
  HF_REQUIRE(out != nullptr, "hf_mma_create: null argument");
  HF_REQUIRE(m >= 1 && m <= kMmaMaxM, "hf_mma_create: m = %d constraints (1..%d)", m, kMmaMaxM);
  HF_REQUIRE(n > m, "hf_mma_create: n = %lld design variables must exceed m = %d", (long long)n, m);
  hf_mma* h = new hf_mma();
  h->n = n;
  h->m = m;
  h->grid = (int)std::min<int64_t>(kMmaMaxGrid, std::max<int64_t>(1, ceil_div(n, 4 * kMmaBlock)));
  const int64_t nvec = 15 + 2 * (int64_t)m;
  double* blk = nullptr;
  if (hipMalloc(&blk, (size_t)nvec * n * sizeof(double)) != hipSuccess) {
    delete h;
    HF_REQUIRE(false, "hf_mma_create: out of device memory for %lld vectors of %lld", (long long)nvec, (long long)n);
  }
  h->block = blk;
  double* p = blk;
  auto take = [&](int64_t cnt) { double* q = p; p += cnt * n; return q; };
  MmaVecs& v = h->v;
  v.n = n;
  v.low = take(1); v.upp = take(1); v.alfa = take(1); v.beta = take(1); v.p0 = take(1); v.q0 = take(1);
  v.P = take(m); v.Q = take(m);
  v.x = take(1); v.xsi = take(1); v.eta = take(1); v.xo = take(1); v.xsio = take(1); v.etao = take(1);
  v.dx = take(1); v.dxsi = take(1); v.deta = take(1);
  constexpr int kMaxNV = NewtonNV<kMmaMaxM>::NV;
  if (hipMalloc(&h->partials, (size_t)h->grid * kMaxNV * sizeof(double)) != hipSuccess ||
      hipMalloc(&h->small, sizeof(MmaSmall)) != hipSuccess ||
      hipHostMalloc(&h->result, sizeof(MmaResult), hipHostMallocDefault) != hipSuccess) {
    hf_mma_destroy(h);
    HF_REQUIRE(false, "hf_mma_create: allocation failed");
  }
  *out = h;
  return 0;
}

extern "C" int hf_mma_destroy(hf_mma* h) {
  if (!h) return 0;
  hipFree(h->block);
  hipFree(h->partials);
  hipFree(h->small);
  hipHostFree(h->result);
  delete h;
  return 0;
}

namespace {

template <int M>
int mma_run(hf_mma* h, int iter, const double* xval, const double* xold1, const double* xold2, const double* xmin,
            const double* xmax, double* low, double* upp, const double* df0dx, const double* fval,
            const double* dfdx, double move, double* xmma, hf_mma_info* info, hipStream_t s) {
  MmaVecs v = h->v;
  const int64_t n = h->n;
  const int G = h->grid;
  HF_CHECK_HIP(hipMemcpyAsync(v.low, low, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HF_CHECK_HIP(hipMemcpyAsync(v.upp, upp, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  mma_setup<M><<<G, kMmaBlock, 0, s>>>(v, iter, xval, xold1, xold2, xmin, xmax, df0dx, dfdx, move, h->partials);
  HF_LAUNCH_CHECK();
  mma_setup_finish<M><<<1, kMmaBlock, 0, s>>>(h->small, h->partials, G, fval);
  HF_LAUNCH_CHECK();
  MmaResult* res = h->result;
  auto residual = [&](double epsi, int trial, double& norm, double& mx) -> int {
    mma_residual<M><<<G, kMmaBlock, 0, s>>>(v, h->small, epsi, trial, h->partials);
    HF_LAUNCH_CHECK();
    mma_residual_finish<M><<<1, kMmaBlock, 0, s>>>(h->small, h->partials, G, epsi, trial, res);
    HF_LAUNCH_CHECK();
    HF_CHECK_HIP(hipStreamSynchronize(s));
    norm = res->norm;
    mx = res->maxabs;
    return 0;
  };
  int newton_steps = 0, trials = 0;
  double rnorm = 0, rmax = 0;
  for (double epsi = 1.0; epsi > 1e-7; epsi *= 0.1) {       // mma.py:224 (epsimin = 1e-7)
    int rc = residual(epsi, 0, rnorm, rmax);
    if (rc) return rc;
    for (int k = 0; rmax > 0.9 * epsi && k < 200; ++k) {   // mma.py:247
      ++newton_steps;
      mma_newton_a<M><<<G, kMmaBlock, 0, s>>>(v, h->small, epsi, h->partials);
      HF_LAUNCH_CHECK();
      mma_newton_small<M><<<1, kMmaBlock, 0, s>>>(h->small, h->partials, G, epsi);
      HF_LAUNCH_CHECK();
      mma_newton_b<M><<<G, kMmaBlock, 0, s>>>(v, h->small, epsi, h->partials);
      HF_LAUNCH_CHECK();
      mma_step<M><<<1, kMmaBlock, 0, s>>>(h->small, h->partials, G);
      HF_LAUNCH_CHECK();
      double nn = 2.0 * rnorm, nm = rmax;
      for (int t = 0; nn > rnorm && t < 50; ++t) {          // mma.py:357
        ++trials;
        rc = residual(epsi, 1, nn, nm);
        if (rc) return rc;
      }
      rnorm = nn;
      rmax = nm;
      HF_REQUIRE(std::isfinite(rnorm), "hf_mma_update_f64: the subproblem's residual is not finite");
    }
  }
  HF_CHECK_HIP(hipMemcpyAsync(xmma, v.x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HF_CHECK_HIP(hipMemcpyAsync(low, v.low, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HF_CHECK_HIP(hipMemcpyAsync(upp, v.upp, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HF_CHECK_HIP(hipStreamSynchronize(s));
  if (info) {
    info->newton_steps = newton_steps;
    info->trials = trials;
    info->res_max = rmax;
    info->res_norm = rnorm;
  }
  return 0;
}

}  // namespace

extern "C" int hf_mma_update_f64(hf_mma* h, int32_t iter, const double* xval, const double* xold1,
                                 const double* xold2, const double* xmin, const double* xmax, double* low,
                                 double* upp, const double* df0dx, const double* fval, const double* dfdx,
                                 double a0, const double* a, const double* c, const double* d, double move,
                                 double* xmma, hf_mma_info* info, hf_stream_t stream) {
  HF_REQUIRE(h && xval && xold1 && xold2 && xmin && xmax && low && upp && df0dx && fval && dfdx && a && c && d &&
                 xmma, "hf_mma_update_f64: null argument");
  hipStream_t s = (hipStream_t)stream;
  MmaSmall host{};
  for (int k = 0; k < h->m; ++k) {
    host.a[k] = a[k];
    host.c[k] = c[k];
    host.d[k] = d[k];
  }
  host.a0 = a0;
  HF_CHECK_HIP(hipMemcpyAsync(h->small, &host, sizeof(MmaSmall), hipMemcpyHostToDevice, s));
  HF_CHECK_HIP(hipStreamSynchronize(s));      // host is a stack object
  switch (h->m) {
    case 1: return mma_run<1>(h, iter, xval, xold1, xold2, xmin, xmax, low, upp, df0dx, fval, dfdx, move, xmma, info, s);
    case 2: return mma_run<2>(h, iter, xval, xold1, xold2, xmin, xmax, low, upp, df0dx, fval, dfdx, move, xmma, info, s);
    case 3: return mma_run<3>(h, iter, xval, xold1, xold2, xmin, xmax, low, upp, df0dx, fval, dfdx, move, xmma, info, s);
    default: return mma_run<4>(h, iter, xval, xold1, xold2, xmin, xmax, low, upp, df0dx, fval, dfdx, move, xmma, info, s);
  }
}

extern "C" int hf_mma_sensitivity_filter_f64(const int32_t* nbr, const double* w, const double* hs, int64_t n,
                                             int32_t k, const double* rho, const double* g, int32_t ncols,
                                             double* out, hf_stream_t stream) {
  HF_REQUIRE(nbr && w && hs && rho && g && out, "hf_mma_sensitivity_filter_f64: null argument");
  HF_REQUIRE(n > 0 && k > 0 && ncols > 0, "hf_mma_sensitivity_filter_f64: empty filter");
  sensitivity_filter_kernel<<<ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(nbr, w, hs, n, k, rho, g, ncols, out);
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_mma_density_filter_f64(const int32_t* nbr, const double* w, const double* hs, int64_t n, int32_t k,
                                         const double* rho, double* out, hf_stream_t stream) {
  HF_REQUIRE(nbr && w && hs && rho && out, "hf_mma_density_filter_f64: null argument");
  HF_REQUIRE(n > 0 && k > 0, "hf_mma_density_filter_f64: empty filter");
  density_filter_kernel<<<ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(nbr, w, hs, n, k, rho, out);
  HF_LAUNCH_CHECK();
  return 0;
}
