// BSR SpMV and device-resident Jacobi-preconditioned PCG / BiCGSTAB.
//
// Replaces jax_solve (jax_fem/solver.py:63-92): Jacobi M = 1/diag(A), Krylov
// with the stop rule ||r||^2 <= max(rtol^2 ||b||^2, atol^2), r the
// unpreconditioned (recursive) residual, maxiter.  BiCGSTAB follows
// jax.scipy.sparse.linalg.bicgstab (JAX, unpinned; van der Vorst's method with
// an early exit when ||s||^2 is below the threshold and breakdown when rho,
// alpha or omega is 0).  PCG is the method for the reference's row-eliminated
// matrix when the initial guess is exact on the Dirichlet rows (then every
// Krylov vector is 0 on them and the iteration runs on the symmetric A_II).
//
// SpMV mapping (wave64): one wavefront per block row.  A block row's values
// are contiguous (blocks of vec*vec); lane l handles entry e = l % BB of block
// rowstart + l / BB + G*it with G = 64 / BB blocks per step (7 for 3x3 blocks,
// lanes 63 idle), so each step streams 63 consecutive values.  Rows are given
// to workgroups in contiguous chunks, and consecutive workgroups are remapped
// so that each XCD (workgroups are dealt round-robin over 8 XCDs) works on one
// contiguous range of rows: neighbouring rows share x entries in that XCD's L2.
//
// Reductions: every workgroup writes FP64 partial sums; the last workgroup to
// finish (atomic ticket) adds them in a fixed order and updates the scalars.
// Results are bitwise reproducible run to run.
#include "common.hpp"
#include "hipfem.h"

namespace hf {
const int32_t* pattern_rowptr(const hf_pattern* p);
const int32_t* pattern_colidx(const hf_pattern* p);
const int32_t* pattern_trans(const hf_pattern* p);
int64_t pattern_nnode(const hf_pattern* p);
int pattern_vec(const hf_pattern* p);

constexpr int kBlock = 256;         // threads per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxGrid = 2048;      // workgroups of the reduction kernels
constexpr int kNumXcd = 8;

struct Scalars {
  double rho, rho_, alpha, omega, rr, bb, thresh, pq, ss, ts, tt;
  int32_t iter, done, status, exit_early;
  uint32_t ticket;
  int32_t maxiter;
};

// Logical workgroup id, remapped so that each XCD owns a contiguous range.
__device__ __forceinline__ int logical_block() {
  const int g = gridDim.x, b = blockIdx.x;
  if (g % kNumXcd != 0) return b;
  return (b % kNumXcd) * (g / kNumXcd) + b / kNumXcd;
}

// Block + grid reduction of NV doubles.  Returns true in the last workgroup,
// where `tot` then holds the grid totals (identical in every thread).
template <int NV>
__device__ bool grid_reduce(double (&v)[NV], double* __restrict__ partials, uint32_t* ticket,
                            double (&tot)[NV]) {
  __shared__ double sm[kWavesPerBlock][NV];
  __shared__ int last;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const double s = wave_sum(v[j]);
    if (lane == 0) sm[w][j] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double s = 0;
      for (int k = 0; k < kWavesPerBlock; ++k) s += sm[k][j];
      partials[(int64_t)blockIdx.x * NV + j] = s;
    }
    __threadfence();
    const uint32_t t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] += partials[(int64_t)b * NV + j];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const double s = wave_sum(acc[j]);
    if (lane == 0) sm[w][j] = s;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double s = 0;
    for (int k = 0; k < kWavesPerBlock; ++k) s += sm[k][j];
    tot[j] = s;
  }
  if (threadIdx.x == 0) *ticket = 0;
  return true;
}

// y_i (all lanes) = sum over the row's blocks of A_blk[i][j] * xv(col*VEC + j).
// TRANS: use the transposed block trans[k], entry [j][i].
template <typename T, int VEC, bool TRANS, typename XF>
__device__ __forceinline__ void bsr_row(int64_t n, const int32_t* __restrict__ rowptr,
                                        const int32_t* __restrict__ colidx, const int32_t* __restrict__ trans,
                                        const T* __restrict__ vals, XF xv, double (&y)[VEC]) {
  constexpr int BB = VEC * VEC;
  constexpr int G = kWave / BB;
  const int lane = threadIdx.x & (kWave - 1);
  const int e = lane % BB, boff = lane / BB;
  const int i = e / VEC, j = e - (e / VEC) * VEC;
  const int32_t k0 = rowptr[n], k1 = rowptr[n + 1];
  double acc = 0;
  if (boff < G) {
    for (int32_t k = k0 + boff; k < k1; k += G) {
      const int32_t c = colidx[k];
      T a;
      if constexpr (TRANS) {
        const int32_t kt = trans[k];
        a = kt >= 0 ? vals[(int64_t)kt * BB + j * VEC + i] : (T)0;
      } else {
        a = __builtin_nontemporal_load(&vals[(int64_t)k * BB + e]);
      }
      acc += (double)(a * xv((int64_t)c * VEC + j));
    }
  }
#pragma unroll
  for (int ii = 0; ii < VEC; ++ii) y[ii] = wave_sum((boff < G && i == ii) ? acc : 0.0);
}

// Contiguous row range of this wave's workgroup.
__device__ __forceinline__ void row_range(int64_t nrows, int64_t& r0, int64_t& r1) {
  const int64_t per = ceil_div(nrows, gridDim.x);
  r0 = (int64_t)logical_block() * per;
  r1 = r0 + per < nrows ? r0 + per : nrows;
}

// ---------------------------------------------------------------------------
// plain SpMV y = A x (or A^T x)
template <typename T, int VEC, bool TRANS>
__global__ void __launch_bounds__(kBlock) spmv_kernel(int64_t nrows, const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ colidx,
                                                      const int32_t* __restrict__ trans,
                                                      const T* __restrict__ vals, const T* __restrict__ x,
                                                      T* __restrict__ y) {
  int64_t r0, r1;
  row_range(nrows, r0, r1);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  auto xv = [&](int64_t c) { return x[c]; };
  for (int64_t n = r0 + w; n < r1; n += kWavesPerBlock) {
    double yy[VEC];
    bsr_row<T, VEC, TRANS>(n, rowptr, colidx, trans, vals, xv, yy);
    if (lane < VEC) {
      double o = yy[0];
#pragma unroll
      for (int ii = 1; ii < VEC; ++ii) o = (lane == ii) ? yy[ii] : o;
      y[n * VEC + lane] = (T)o;
    }
  }
}

// ---------------------------------------------------------------------------
// PCG (Jacobi).  init: r = b - A x, z = dinv r, p_old = 0, beta = 0.
template <typename T, int VEC, bool TRANS>
__global__ void __launch_bounds__(kBlock) pcg_init_kernel(
    int64_t nrows, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const int32_t* __restrict__ trans, const T* __restrict__ vals, const T* __restrict__ dinv,
    const T* __restrict__ b, const T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
    T* __restrict__ p_old, Scalars* sc, double* partials, double rtol2, double atol2) {
  int64_t r0, r1;
  row_range(nrows, r0, r1);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  auto xv = [&](int64_t c) { return x[c]; };
  double loc[3] = {0, 0, 0};  // r.r, r.z, b.b
  for (int64_t n = r0 + w; n < r1; n += kWavesPerBlock) {
    double yy[VEC];
    bsr_row<T, VEC, TRANS>(n, rowptr, colidx, trans, vals, xv, yy);
    if (lane < VEC) {
      double ax = yy[0];
#pragma unroll
      for (int ii = 1; ii < VEC; ++ii) ax = (lane == ii) ? yy[ii] : ax;
      const int64_t d = n * VEC + lane;
      const T rr = (T)((double)b[d] - ax);
      const T zz = dinv[d] * rr;
      r[d] = rr;
      z[d] = zz;
      p_old[d] = (T)0;
      loc[0] += (double)rr * (double)rr;
      loc[1] += (double)rr * (double)zz;
      loc[2] += (double)b[d] * (double)b[d];
    }
  }
  double tot[3];
  if (grid_reduce<3>(loc, partials, &sc->ticket, tot) && threadIdx.x == 0) {
    sc->rr = tot[0];
    sc->rho = tot[1];
    sc->bb = tot[2];
    sc->thresh = fmax(rtol2 * tot[2], atol2);
    sc->alpha = 0;
    sc->iter = 0;
    sc->status = 0;
    sc->done = (tot[0] <= sc->thresh) ? 1 : 0;
    if (!isfinite(tot[0])) { sc->done = 1; sc->status = 3; }
    sc->omega = 0;  // beta for the first direction (p_new = z)
  }
}

// p_new = z + beta p_old (on the fly), q = A p_new, alpha = rho / (p_new . q)
template <typename T, int VEC, bool TRANS>
__global__ void __launch_bounds__(kBlock) pcg_spmv_kernel(
    int64_t nrows, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const int32_t* __restrict__ trans, const T* __restrict__ vals, const T* __restrict__ z,
    const T* __restrict__ p_old, T* __restrict__ p_new, T* __restrict__ q, Scalars* sc,
    double* partials) {
  if (sc->done) return;
  const T beta = (T)sc->omega;
  int64_t r0, r1;
  row_range(nrows, r0, r1);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  auto xv = [&](int64_t c) { return z[c] + beta * p_old[c]; };
  double loc[1] = {0};
  for (int64_t n = r0 + w; n < r1; n += kWavesPerBlock) {
    double yy[VEC];
    bsr_row<T, VEC, TRANS>(n, rowptr, colidx, trans, vals, xv, yy);
    if (lane < VEC) {
      double ap = yy[0];
#pragma unroll
      for (int ii = 1; ii < VEC; ++ii) ap = (lane == ii) ? yy[ii] : ap;
      const int64_t d = n * VEC + lane;
      const T pn = z[d] + beta * p_old[d];
      p_new[d] = pn;
      q[d] = (T)ap;
      loc[0] += (double)pn * ap;
    }
  }
  double tot[1];
  if (grid_reduce<1>(loc, partials, &sc->ticket, tot) && threadIdx.x == 0) {
    sc->pq = tot[0];
    if (!(tot[0] > 0.0) || !isfinite(tot[0])) {
      sc->done = 1;
      sc->status = isfinite(tot[0]) ? 2 : 3;
    } else {
      sc->alpha = sc->rho / tot[0];
    }
  }
}

// x += alpha p, r -= alpha q, z = dinv r, beta = r.z / rho
template <typename T>
__global__ void __launch_bounds__(kBlock) pcg_update_kernel(
    int64_t N, const T* __restrict__ dinv, const T* __restrict__ p, const T* __restrict__ q,
    T* __restrict__ x, T* __restrict__ r, T* __restrict__ z, Scalars* sc, double* partials) {
  if (sc->done) return;
  const T alpha = (T)sc->alpha;
  const int64_t per = ceil_div(N, gridDim.x);
  const int64_t i0 = (int64_t)logical_block() * per;
  const int64_t i1 = i0 + per < N ? i0 + per : N;
  double loc[2] = {0, 0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBlock) {
    x[i] += alpha * p[i];
    const T rr = r[i] - alpha * q[i];
    const T zz = dinv[i] * rr;
    r[i] = rr;
    z[i] = zz;
    loc[0] += (double)rr * (double)zz;
    loc[1] += (double)rr * (double)rr;
  }
  double tot[2];
  if (grid_reduce<2>(loc, partials, &sc->ticket, tot) && threadIdx.x == 0) {
    sc->iter += 1;
    sc->rr = tot[1];
    sc->omega = tot[0] / sc->rho;  // beta
    sc->rho = tot[0];
    if (!isfinite(tot[1])) { sc->done = 1; sc->status = 3; }
    else if (tot[1] <= sc->thresh) { sc->done = 1; sc->status = 0; }
    else if (sc->iter >= sc->maxiter) { sc->done = 1; sc->status = 1; }
  }
}

// ---------------------------------------------------------------------------
// BiCGSTAB (jax semantics).  Vectors: r, rhat, p, q, phat, s, shat, t.
template <typename T, int VEC, bool TRANS>
__global__ void __launch_bounds__(kBlock) bicg_init_kernel(
    int64_t nrows, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const int32_t* __restrict__ trans, const T* __restrict__ vals, const T* __restrict__ b,
    const T* __restrict__ x, T* __restrict__ r, T* __restrict__ rhat, T* __restrict__ p,
    T* __restrict__ q, Scalars* sc, double* partials, double rtol2, double atol2) {
  int64_t r0, r1;
  row_range(nrows, r0, r1);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  auto xv = [&](int64_t c) { return x[c]; };
  double loc[2] = {0, 0};  // r.r (= rhat.r), b.b
  for (int64_t n = r0 + w; n < r1; n += kWavesPerBlock) {
    double yy[VEC];
    bsr_row<T, VEC, TRANS>(n, rowptr, colidx, trans, vals, xv, yy);
    if (lane < VEC) {
      double ax = yy[0];
#pragma unroll
      for (int ii = 1; ii < VEC; ++ii) ax = (lane == ii) ? yy[ii] : ax;
      const int64_t d = n * VEC + lane;
      const T rr = (T)((double)b[d] - ax);
      r[d] = rr; rhat[d] = rr; p[d] = rr; q[d] = rr;
      loc[0] += (double)rr * (double)rr;
      loc[1] += (double)b[d] * (double)b[d];
    }
  }
  double tot[2];
  if (grid_reduce<2>(loc, partials, &sc->ticket, tot) && threadIdx.x == 0) {
    sc->rr = tot[0];
    sc->rho_ = tot[0];
    sc->bb = tot[1];
    sc->rho = 1.0; sc->alpha = 1.0; sc->omega = 1.0;
    sc->thresh = fmax(rtol2 * tot[1], atol2);
    sc->iter = 0; sc->status = 0; sc->exit_early = 0;
    sc->done = (tot[0] <= sc->thresh) ? 1 : 0;
    if (!isfinite(tot[0])) { sc->done = 1; sc->status = 3; }
  }
}

// p = r + beta (p - omega q), phat = dinv p
template <typename T>
__global__ void __launch_bounds__(kBlock) bicg_p_kernel(int64_t N, const T* __restrict__ dinv,
                                                        const T* __restrict__ r, const T* __restrict__ q,
                                                        T* __restrict__ p, T* __restrict__ phat,
                                                        const Scalars* sc) {
  if (sc->done) return;
  const T beta = (T)(sc->rho_ / sc->rho * sc->alpha / sc->omega);
  const T omega = (T)sc->omega;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (int64_t)gridDim.x * kBlock) {
    const T pp = r[i] + beta * (p[i] - omega * q[i]);
    p[i] = pp;
    phat[i] = dinv[i] * pp;
  }
}

// y = A v with a grid-reduced dot.  MODE 0: q = A phat, alpha = rho_ / (rhat . q)
//                                   MODE 1: t = A shat, omega = (t.s) / (t.t)
template <typename T, int VEC, bool TRANS, int MODE>
__global__ void __launch_bounds__(kBlock) bicg_spmv_kernel(
    int64_t nrows, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const int32_t* __restrict__ trans, const T* __restrict__ vals, const T* __restrict__ v,
    const T* __restrict__ w1, T* __restrict__ out, Scalars* sc, double* partials) {
  if (sc->done || (MODE == 1 && sc->exit_early)) return;
  int64_t r0, r1;
  row_range(nrows, r0, r1);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  auto xv = [&](int64_t c) { return v[c]; };
  double loc[2] = {0, 0};
  for (int64_t n = r0 + w; n < r1; n += kWavesPerBlock) {
    double yy[VEC];
    bsr_row<T, VEC, TRANS>(n, rowptr, colidx, trans, vals, xv, yy);
    if (lane < VEC) {
      double o = yy[0];
#pragma unroll
      for (int ii = 1; ii < VEC; ++ii) o = (lane == ii) ? yy[ii] : o;
      const int64_t d = n * VEC + lane;
      const T ot = (T)o;
      out[d] = ot;
      loc[0] += (double)w1[d] * (double)ot;
      loc[1] += (double)ot * (double)ot;
    }
  }
  double tot[2];
  if (grid_reduce<2>(loc, partials, &sc->ticket, tot) && threadIdx.x == 0) {
    if (MODE == 0) {
      sc->alpha = sc->rho_ / tot[0];
      if (!isfinite(sc->alpha)) { sc->done = 1; sc->status = 3; }
    } else {
      sc->ts = tot[0];
      sc->tt = tot[1];
      sc->omega = tot[0] / tot[1];
    }
  }
}

// s = r - alpha q, shat = dinv s, exit_early = s.s < thresh
template <typename T>
__global__ void __launch_bounds__(kBlock) bicg_s_kernel(int64_t N, const T* __restrict__ dinv,
                                                        const T* __restrict__ r, const T* __restrict__ q,
                                                        T* __restrict__ s, T* __restrict__ shat, Scalars* sc,
                                                        double* partials) {
  if (sc->done) return;
  const T alpha = (T)sc->alpha;
  const int64_t per = ceil_div(N, gridDim.x);
  const int64_t i0 = (int64_t)logical_block() * per;
  const int64_t i1 = i0 + per < N ? i0 + per : N;
  double loc[1] = {0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBlock) {
    const T ss = r[i] - alpha * q[i];
    s[i] = ss;
    shat[i] = dinv[i] * ss;
    loc[0] += (double)ss * (double)ss;
  }
  double tot[1];
  if (grid_reduce<1>(loc, partials, &sc->ticket, tot) && threadIdx.x == 0) {
    sc->ss = tot[0];
    sc->exit_early = (tot[0] < sc->thresh) ? 1 : 0;
  }
}

// x += alpha phat (+ omega shat); r = s (- omega t); rho_ = rhat . r; convergence
template <typename T>
__global__ void __launch_bounds__(kBlock) bicg_x_kernel(int64_t N, const T* __restrict__ rhat,
                                                        const T* __restrict__ phat, const T* __restrict__ shat,
                                                        const T* __restrict__ s, const T* __restrict__ t,
                                                        T* __restrict__ x, T* __restrict__ r, Scalars* sc,
                                                        double* partials) {
  if (sc->done) return;
  const bool early = sc->exit_early != 0;
  const T alpha = (T)sc->alpha, omega = (T)sc->omega;
  const int64_t per = ceil_div(N, gridDim.x);
  const int64_t i0 = (int64_t)logical_block() * per;
  const int64_t i1 = i0 + per < N ? i0 + per : N;
  double loc[2] = {0, 0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBlock) {
    T rr;
    if (early) {
      x[i] += alpha * phat[i];
      rr = s[i];
    } else {
      x[i] += alpha * phat[i] + omega * shat[i];
      rr = s[i] - omega * t[i];
    }
    r[i] = rr;
    loc[0] += (double)rhat[i] * (double)rr;
    loc[1] += (double)rr * (double)rr;
  }
  double tot[2];
  if (grid_reduce<2>(loc, partials, &sc->ticket, tot) && threadIdx.x == 0) {
    sc->iter += 1;
    const double rho_used = sc->rho_;
    sc->rho = rho_used;
    sc->rho_ = tot[0];
    sc->rr = tot[1];
    if (!isfinite(tot[1])) { sc->done = 1; sc->status = 3; }
    else if (tot[1] <= sc->thresh) { sc->done = 1; sc->status = 0; }
    else if ((!early && sc->omega == 0.0) || sc->alpha == 0.0 || rho_used == 0.0) { sc->done = 1; sc->status = 2; }
    else if (sc->iter >= sc->maxiter) { sc->done = 1; sc->status = 1; }
    if (early && sc->status == 0) sc->done = 1;
  }
}

// ---------------------------------------------------------------------------
__global__ void set_maxiter_kernel(Scalars* sc, int32_t maxiter) {
  sc->maxiter = maxiter;
  sc->ticket = 0;
  sc->done = 0;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) dot_kernel(int64_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                     double* partials, uint32_t* ticket, double* out) {
  const int64_t per = ceil_div(n, gridDim.x);
  const int64_t i0 = (int64_t)blockIdx.x * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  double loc[1] = {0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBlock) loc[0] += (double)x[i] * (double)y[i];
  double tot[1];
  if (grid_reduce<1>(loc, partials, ticket, tot) && threadIdx.x == 0) *out = tot[0];
}

// ---------------------------------------------------------------------------
struct Workspace {
  Scalars* sc;
  double* partials;
  char* vec;  // vectors follow
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
static constexpr size_t kHeader = 4096 + (size_t)kMaxGrid * 4 * sizeof(double);

// Optional HIP-event timing of the Krylov SpMV kernel (first SpMV of every poll
// batch), for roofline reporting; off unless hf_profile_enable(1) was called.
static bool g_prof = false;
static double g_prof_ms = 0.0;
static int64_t g_prof_n = 0;
static hipEvent_t g_ev[2] = {nullptr, nullptr};

static int prof_begin(hipStream_t s) {
  if (!g_prof) return 0;
  if (!g_ev[0]) {
    HF_CHECK_HIP(hipEventCreate(&g_ev[0]));
    HF_CHECK_HIP(hipEventCreate(&g_ev[1]));
  }
  HF_CHECK_HIP(hipEventRecord(g_ev[0], s));
  return 0;
}
static int prof_end(hipStream_t s) {
  if (!g_prof) return 0;
  HF_CHECK_HIP(hipEventRecord(g_ev[1], s));
  return 0;
}
static int prof_collect() {
  if (!g_prof) return 0;
  float ms = 0.f;
  HF_CHECK_HIP(hipEventElapsedTime(&ms, g_ev[0], g_ev[1]));
  g_prof_ms += ms;
  g_prof_n += 1;
  return 0;
}

static inline int rows_grid(int64_t nrows) {
  int64_t g = ceil_div(nrows, kWavesPerBlock * 2);
  if (g > kMaxGrid) g = kMaxGrid;
  if (g >= 64) g = (g / kNumXcd) * kNumXcd;
  return (int)(g < 1 ? 1 : g);
}
static inline int elem_grid(int64_t n) {
  int64_t g = ceil_div(n, kBlock * 4);
  if (g > kMaxGrid) g = kMaxGrid;
  if (g >= 64) g = (g / kNumXcd) * kNumXcd;
  return (int)(g < 1 ? 1 : g);
}

template <typename T, int VEC, bool TRANS>
static int spmv_launch(const hf_pattern* p, const T* vals, const T* x, T* y, hipStream_t s) {
  const int64_t nrows = pattern_nnode(p);
  spmv_kernel<T, VEC, TRANS><<<rows_grid(nrows), kBlock, 0, s>>>(nrows, pattern_rowptr(p), pattern_colidx(p),
                                                                  pattern_trans(p), vals, x, y);
  HF_LAUNCH_CHECK();
  return 0;
}

template <typename T, bool TRANS>
static int spmv_dispatch(const hf_pattern* p, const T* vals, const T* x, T* y, hipStream_t s) {
  HF_REQUIRE(p && vals && x && y, "hf_spmv: null argument");
  switch (pattern_vec(p)) {
    case 1: return spmv_launch<T, 1, TRANS>(p, vals, x, y, s);
    case 3: return spmv_launch<T, 3, TRANS>(p, vals, x, y, s);
    default: HF_REQUIRE(false, "hf_spmv: vec=%d not supported", pattern_vec(p));
  }
}

template <typename T, int VEC, bool TRANS>
static int pcg_run(const hf_pattern* p, const T* vals, const T* dinv, const T* b, T* x, int maxiter, double rtol,
                   double atol, int poll, void* ws, hf_krylov_info* info, hipStream_t s) {
  const int64_t nrows = pattern_nnode(p), N = nrows * VEC;
  Scalars* sc = (Scalars*)ws;
  double* partials = (double*)((char*)ws + 4096);
  T* v = (T*)((char*)ws + kHeader);
  const size_t stride = align256(N * sizeof(T)) / sizeof(T);
  T *r = v, *z = v + stride, *p0 = v + 2 * stride, *p1 = v + 3 * stride, *q = v + 4 * stride;
  const int32_t *rp = pattern_rowptr(p), *ci = pattern_colidx(p), *tr = pattern_trans(p);
  const int gr = rows_grid(nrows), ge = elem_grid(N);
  set_maxiter_kernel<<<1, 1, 0, s>>>(sc, maxiter);
  pcg_init_kernel<T, VEC, TRANS><<<gr, kBlock, 0, s>>>(nrows, rp, ci, tr, vals, dinv, b, x, r, z, p0, sc,
                                                       partials, rtol * rtol, atol * atol);
  HF_LAUNCH_CHECK();
  Scalars host{};
  int it = 0;
  for (;;) {
    for (int k = 0; k < poll; ++k, ++it) {
      T* pold = (it & 1) ? p1 : p0;
      T* pnew = (it & 1) ? p0 : p1;
      if (k == 0 && prof_begin(s)) return -2;
      pcg_spmv_kernel<T, VEC, TRANS><<<gr, kBlock, 0, s>>>(nrows, rp, ci, tr, vals, z, pold, pnew, q, sc, partials);
      if (k == 0 && prof_end(s)) return -2;
      pcg_update_kernel<T><<<ge, kBlock, 0, s>>>(N, dinv, pnew, q, x, r, z, sc, partials);
    }
    HF_LAUNCH_CHECK();
    HF_CHECK_HIP(hipMemcpyAsync(&host, sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
    HF_CHECK_HIP(hipStreamSynchronize(s));
    if (!host.done || host.iter > it - poll) {
      if (prof_collect()) return -2;  // the timed SpMV did real work
    }
    if (host.done || it >= maxiter + poll) break;
  }
  if (info) {
    info->iterations = host.iter;
    info->status = host.done ? host.status : 1;
    info->res_norm = sqrt(host.rr);
    info->b_norm = sqrt(host.bb);
  }
  return 0;
}

template <typename T, int VEC, bool TRANS>
static int bicg_run(const hf_pattern* p, const T* vals, const T* dinv, const T* b, T* x, int maxiter, double rtol,
                    double atol, int poll, void* ws, hf_krylov_info* info, hipStream_t s) {
  const int64_t nrows = pattern_nnode(p), N = nrows * VEC;
  Scalars* sc = (Scalars*)ws;
  double* partials = (double*)((char*)ws + 4096);
  T* v = (T*)((char*)ws + kHeader);
  const size_t stride = align256(N * sizeof(T)) / sizeof(T);
  T *r = v, *rhat = v + stride, *pp = v + 2 * stride, *q = v + 3 * stride, *phat = v + 4 * stride,
    *sv = v + 5 * stride, *shat = v + 6 * stride, *t = v + 7 * stride;
  const int32_t *rp = pattern_rowptr(p), *ci = pattern_colidx(p), *tr = pattern_trans(p);
  const int gr = rows_grid(nrows), ge = elem_grid(N);
  set_maxiter_kernel<<<1, 1, 0, s>>>(sc, maxiter);
  bicg_init_kernel<T, VEC, TRANS><<<gr, kBlock, 0, s>>>(nrows, rp, ci, tr, vals, b, x, r, rhat, pp, q, sc, partials,
                                                        rtol * rtol, atol * atol);
  HF_LAUNCH_CHECK();
  Scalars host{};
  int it = 0;
  for (;;) {
    for (int k = 0; k < poll; ++k, ++it) {
      bicg_p_kernel<T><<<ge, kBlock, 0, s>>>(N, dinv, r, q, pp, phat, sc);
      if (k == 0 && prof_begin(s)) return -2;
      bicg_spmv_kernel<T, VEC, TRANS, 0><<<gr, kBlock, 0, s>>>(nrows, rp, ci, tr, vals, phat, rhat, q, sc, partials);
      if (k == 0 && prof_end(s)) return -2;
      bicg_s_kernel<T><<<ge, kBlock, 0, s>>>(N, dinv, r, q, sv, shat, sc, partials);
      bicg_spmv_kernel<T, VEC, TRANS, 1><<<gr, kBlock, 0, s>>>(nrows, rp, ci, tr, vals, shat, sv, t, sc, partials);
      bicg_x_kernel<T><<<ge, kBlock, 0, s>>>(N, rhat, phat, shat, sv, t, x, r, sc, partials);
    }
    HF_LAUNCH_CHECK();
    HF_CHECK_HIP(hipMemcpyAsync(&host, sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
    HF_CHECK_HIP(hipStreamSynchronize(s));
    if (!host.done || host.iter > it - poll) {
      if (prof_collect()) return -2;
    }
    if (host.done || it >= maxiter + poll) break;
  }
  if (info) {
    info->iterations = host.iter;
    info->status = host.done ? host.status : 1;
    info->res_norm = sqrt(host.rr);
    info->b_norm = sqrt(host.bb);
  }
  return 0;
}

template <typename T>
static int krylov_dispatch(int method, const hf_pattern* p, const T* vals, const T* dinv, const T* b, T* x,
                           int transpose, int maxiter, double rtol, double atol, int poll, void* ws,
                           hf_krylov_info* info, hipStream_t s) {
  HF_REQUIRE(p && vals && dinv && b && x && ws, "hf_krylov_solve: null argument");
  HF_REQUIRE(maxiter >= 0, "hf_krylov_solve: maxiter < 0");
  if (poll < 1) poll = 1;
  const int vec = pattern_vec(p);
  HF_REQUIRE(vec == 1 || vec == 3, "hf_krylov_solve: vec=%d not supported", vec);
  HF_REQUIRE(method == HF_KRYLOV_PCG || method == HF_KRYLOV_BICGSTAB, "hf_krylov_solve: unknown method %d", method);
#define KRY(M, V, TR) M<T, V, TR>(p, vals, dinv, b, x, maxiter, rtol, atol, poll, ws, info, s)
  if (method == HF_KRYLOV_PCG) {
    if (vec == 3) return transpose ? KRY(pcg_run, 3, true) : KRY(pcg_run, 3, false);
    return transpose ? KRY(pcg_run, 1, true) : KRY(pcg_run, 1, false);
  }
  if (vec == 3) return transpose ? KRY(bicg_run, 3, true) : KRY(bicg_run, 3, false);
  return transpose ? KRY(bicg_run, 1, true) : KRY(bicg_run, 1, false);
#undef KRY
}

}  // namespace hf

using namespace hf;

extern "C" size_t hf_krylov_workspace_size(int32_t method, int64_t N, int32_t elem_bytes) {
  const int nvec = (method == HF_KRYLOV_BICGSTAB) ? 8 : 5;
  return kHeader + (size_t)nvec * align256((size_t)N * elem_bytes);
}

extern "C" int hf_krylov_solve_f64(int32_t method, const hf_pattern* p, const double* bsr, const double* dinv,
                                   const double* b, double* x, int32_t transpose, int32_t maxiter, double rtol,
                                   double atol, int32_t poll, void* ws, hf_krylov_info* info, hf_stream_t s) {
  return krylov_dispatch<double>(method, p, bsr, dinv, b, x, transpose, maxiter, rtol, atol, poll, ws, info,
                                 (hipStream_t)s);
}
extern "C" int hf_krylov_solve_f32(int32_t method, const hf_pattern* p, const float* bsr, const float* dinv,
                                   const float* b, float* x, int32_t transpose, int32_t maxiter, double rtol,
                                   double atol, int32_t poll, void* ws, hf_krylov_info* info, hf_stream_t s) {
  return krylov_dispatch<float>(method, p, bsr, dinv, b, x, transpose, maxiter, rtol, atol, poll, ws, info,
                                (hipStream_t)s);
}

extern "C" int hf_spmv_bsr_f64(const hf_pattern* p, const double* a, const double* x, double* y, hf_stream_t s) {
  return spmv_dispatch<double, false>(p, a, x, y, (hipStream_t)s);
}
extern "C" int hf_spmv_bsr_f32(const hf_pattern* p, const float* a, const float* x, float* y, hf_stream_t s) {
  return spmv_dispatch<float, false>(p, a, x, y, (hipStream_t)s);
}
extern "C" int hf_spmv_bsr_t_f64(const hf_pattern* p, const double* a, const double* x, double* y, hf_stream_t s) {
  return spmv_dispatch<double, true>(p, a, x, y, (hipStream_t)s);
}
extern "C" int hf_spmv_bsr_t_f32(const hf_pattern* p, const float* a, const float* x, float* y, hf_stream_t s) {
  return spmv_dispatch<float, true>(p, a, x, y, (hipStream_t)s);
}

extern "C" int hf_profile_enable(int32_t on) {
  g_prof = on != 0;
  g_prof_ms = 0.0;
  g_prof_n = 0;
  return 0;
}
extern "C" int hf_profile_read(int64_t* launches, double* total_ms) {
  HF_REQUIRE(launches && total_ms, "hf_profile_read: null argument");
  *launches = g_prof_n;
  *total_ms = g_prof_ms;
  return 0;
}

extern "C" size_t hf_dot_workspace_size(void) { return 4096 + (size_t)kMaxGrid * sizeof(double) + 256; }

template <typename T>
static int dot_impl(const T* x, const T* y, int64_t n, void* ws, double* result, hipStream_t s) {
  HF_REQUIRE(x && y && ws && result, "hf_dot: null argument");
  uint32_t* ticket = (uint32_t*)ws;
  double* out = (double*)((char*)ws + 256);
  double* partials = (double*)((char*)ws + 4096);
  HF_CHECK_HIP(hipMemsetAsync(ticket, 0, sizeof(uint32_t), s));
  const int g = elem_grid(n);
  dot_kernel<T><<<g, kBlock, 0, s>>>(n, x, y, partials, ticket, out);
  HF_LAUNCH_CHECK();
  HF_CHECK_HIP(hipMemcpyAsync(result, out, sizeof(double), hipMemcpyDeviceToHost, s));
  HF_CHECK_HIP(hipStreamSynchronize(s));
  return 0;
}
extern "C" int hf_dot_f64(const double* x, const double* y, int64_t n, void* ws, double* r, hf_stream_t s) {
  return dot_impl<double>(x, y, n, ws, r, (hipStream_t)s);
}
extern "C" int hf_dot_f32(const float* x, const float* y, int64_t n, void* ws, double* r, hf_stream_t s) {
  return dot_impl<float>(x, y, n, ws, r, (hipStream_t)s);
}
