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
// SpMV: sliced ELL over node blocks, column-major blocks, one wavefront per
// slice of 21 node rows (krylov_common.hpp, sweep).  Slices are given to
// workgroups in contiguous chunks, and consecutive workgroups are remapped so
// that each XCD (workgroups are dealt round-robin over 8 XCDs) works on one
// contiguous range of rows: neighbouring rows share x entries in that XCD's L2.
//
// PCG is the Chronopoulos-Gear form: the SpMV multiplies the preconditioned
// residual u (one gathered vector), the direction p and s = A p follow by
// recurrences in the elementwise kernel, and all dot products of an iteration
// are reduced in one phase (one scalar kernel per iteration).
//
// Reductions: every workgroup of a vector kernel writes its FP64 partial sums
// with plain stores; a one-workgroup scalar kernel launched next adds them in
// a fixed order and updates the scalars (alpha, beta, flags).  No fences or
// atomics between workgroups: on a multi-XCD GPU a device-scope release per
// workgroup costs an L2 write-back, while the kernel boundary already makes
// the partials visible.  Results are bitwise reproducible run to run.
#include <hip/hip_ext.h>

#include "common.hpp"
#include "hipfem.h"

#include "krylov_common.hpp"

namespace hf {
// ---------------------------------------------------------------------------
// plain SpMV y = A x (or A^T x)
template <typename T, int VEC, bool TRANS>
__global__ void __launch_bounds__(kBlock) spmv_kernel(MatArgs m, const T* __restrict__ vals,
                                                      const T* __restrict__ x, T* __restrict__ y) {
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return x[c]; };
  auto epi = [&](int64_t n, int i, T v) { y[n * VEC + i] = v; };
  sweep<T, VEC, TRANS>(m, r0, r1, vals, xv, epi);
}

// ---------------------------------------------------------------------------
// PCG (Jacobi), Chronopoulos-Gear form:
//   u = M r, w = A u, gamma = r.u, delta = w.u,
//   beta = gamma / gamma_prev, alpha = gamma / (delta - beta gamma / alpha_prev),
//   p = u + beta p, s = w + beta s (= A p), x += alpha p, r -= alpha s.
// delta - beta gamma / alpha_prev equals p.Ap: <= 0 is a breakdown (status 2).
// Partials: region A (the elementwise / init kernels) and region B (the SpMV).
constexpr int kPartB = 3 * kMaxGrid;

// init: r = b - A x, u = dinv r, p = s = 0; partials (A) r.u, r.r, b.b
template <typename T, int VEC, bool TRANS>
__global__ void __launch_bounds__(kBlock) cg_init_kernel(MatArgs m, const T* __restrict__ vals,
                                                         const T* __restrict__ dinv, const T* __restrict__ b,
                                                         const T* __restrict__ x, T* __restrict__ r,
                                                         T* __restrict__ u, T* __restrict__ p, T* __restrict__ sv,
                                                         double* partials) {
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return x[c]; };
  double loc[3] = {0, 0, 0};
  auto epi = [&](int64_t n, int i, T ax) {
    const int64_t d = n * VEC + i;
    const T rr = b[d] - ax;
    const T uu = dinv[d] * rr;
    r[d] = rr;
    u[d] = uu;
    p[d] = (T)0;
    sv[d] = (T)0;
    loc[0] += (double)rr * (double)uu;
    loc[1] += (double)rr * (double)rr;
    loc[2] += (double)b[d] * (double)b[d];
  };
  sweep<T, VEC, TRANS>(m, r0, r1, vals, xv, epi);
  block_partials<3>(loc, partials);
}

// w = A u; partial (B) u.w.  sc = nullptr at set-up (the flags are not yet reset).
template <typename T, int VEC, bool TRANS>
__global__ void __launch_bounds__(kBlock) cg_spmv_kernel(MatArgs m, const T* __restrict__ vals,
                                                         const T* __restrict__ u, T* __restrict__ w,
                                                         const Scalars* sc, double* partials) {
  if (sc && sc->done) return;
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return u[c]; };
  double loc[1] = {0};
  auto epi = [&](int64_t n, int i, T v) {
    const int64_t d = n * VEC + i;
    w[d] = v;
    loc[0] += (double)u[d] * (double)v;
  };
  sweep<T, VEC, TRANS>(m, r0, r1, vals, xv, epi);
  block_partials<1>(loc, partials);
}

// Stop threshold: max(rtol^2 * ref, atol^2), ref = ||b||^2 or ||r0||^2 (rel_r0).
__global__ void __launch_bounds__(kScalarBlock) cg_init_scalars(Scalars* sc, const double* partA, int nA,
                                                                const double* partB, int nB, double rtol2,
                                                                double atol2, int rel_r0, int maxiter) {
  double a[3], d[1];
  sum_partials<3>(partA, nA, a);
  sum_partials<1>(partB, nB, d);
  if (threadIdx.x != 0) return;
  sc->rho = a[0];   // gamma
  sc->rr = a[1];
  sc->bb = a[2];
  sc->pq = d[0];    // delta
  sc->thresh = fmax(rtol2 * (rel_r0 ? a[1] : a[2]), atol2);
  sc->omega = 0;    // beta of the first direction (p = u)
  sc->iter = 0;
  sc->status = 0;
  sc->maxiter = maxiter;
  sc->done = (a[1] <= sc->thresh) ? 1 : 0;
  if (!isfinite(a[1])) { sc->done = 1; sc->status = 3; }
  else if (!sc->done && maxiter <= 0) { sc->done = 1; sc->status = 1; }
  else if (!sc->done && !(d[0] > 0.0)) { sc->done = 1; sc->status = isfinite(d[0]) ? 2 : 3; }
  sc->alpha = sc->done ? 0.0 : a[0] / d[0];
}

// p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = dinv r; partials (A) r.u, r.r
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_update_kernel(int64_t N, const T* __restrict__ dinv,
                                                           T* __restrict__ u, const T* __restrict__ w,
                                                           T* __restrict__ p, T* __restrict__ sv, T* __restrict__ x,
                                                           T* __restrict__ r, const Scalars* sc, double* partials) {
  if (sc->done) return;
  const T alpha = (T)sc->alpha, beta = (T)sc->omega;
  const int64_t per = ceil_div(N, gridDim.x);
  const int64_t i0 = (int64_t)logical_block() * per;
  const int64_t i1 = i0 + per < N ? i0 + per : N;
  double loc[2] = {0, 0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBlock) {
    const T pp = u[i] + beta * p[i];
    const T ss = w[i] + beta * sv[i];
    p[i] = pp;
    sv[i] = ss;
    x[i] += alpha * pp;
    const T rr = r[i] - alpha * ss;
    const T uu = dinv[i] * rr;
    r[i] = rr;
    u[i] = uu;
    loc[0] += (double)rr * (double)uu;
    loc[1] += (double)rr * (double)rr;
  }
  block_partials<2>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) cg_scalars(Scalars* sc, const double* partA, int nA,
                                                           const double* partB, int nB) {
  if (sc->done) return;
  double a[2], d[1];
  sum_partials<2>(partA, nA, a);
  sum_partials<1>(partB, nB, d);
  if (threadIdx.x != 0) return;
  sc->iter += 1;
  sc->rr = a[1];
  if (!isfinite(a[1])) { sc->done = 1; sc->status = 3; return; }
  if (a[1] <= sc->thresh) { sc->done = 1; sc->status = 0; return; }
  if (sc->iter >= sc->maxiter) { sc->done = 1; sc->status = 1; return; }
  const double gamma = a[0], beta = gamma / sc->rho;
  const double denom = d[0] - beta * gamma / sc->alpha;  // p.Ap of the next direction
  if (!(denom > 0.0) || !isfinite(denom)) {
    sc->done = 1;
    sc->status = isfinite(denom) ? 2 : 3;
    return;
  }
  sc->omega = beta;
  sc->alpha = gamma / denom;
  sc->rho = gamma;
  sc->pq = denom;
}

// ---------------------------------------------------------------------------
// BiCGSTAB (jax semantics).  Vectors: r, rhat, p, q, phat, s, shat, t.
template <typename T, int VEC, bool TRANS>
__global__ void __launch_bounds__(kBlock) bicg_init_kernel(MatArgs m, const T* __restrict__ vals,
                                                           const T* __restrict__ b, const T* __restrict__ x,
                                                           T* __restrict__ r, T* __restrict__ rhat, T* __restrict__ p,
                                                           T* __restrict__ q, double* partials) {
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return x[c]; };
  double loc[2] = {0, 0};  // r.r (= rhat.r), b.b
  auto epi = [&](int64_t n, int i, T ax) {
    const int64_t d = n * VEC + i;
    const T rr = b[d] - ax;
    r[d] = rr; rhat[d] = rr; p[d] = rr; q[d] = rr;
    loc[0] += (double)rr * (double)rr;
    loc[1] += (double)b[d] * (double)b[d];
  };
  sweep<T, VEC, TRANS>(m, r0, r1, vals, xv, epi);
  block_partials<2>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) bicg_init_scalars(Scalars* sc, const double* partials, int nparts,
                                                            double rtol2, double atol2, int rel_r0, int maxiter) {
  double tot[2];
  sum_partials<2>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  sc->rr = tot[0];
  sc->rho_ = tot[0];
  sc->bb = tot[1];
  sc->rho = 1.0; sc->alpha = 1.0; sc->omega = 1.0;
  sc->thresh = fmax(rtol2 * (rel_r0 ? tot[0] : tot[1]), atol2);
  sc->iter = 0; sc->status = 0; sc->exit_early = 0;
  sc->maxiter = maxiter;
  sc->done = (tot[0] <= sc->thresh) ? 1 : 0;
  if (!isfinite(tot[0])) { sc->done = 1; sc->status = 3; }
  else if (maxiter <= 0 && !sc->done) { sc->done = 1; sc->status = 1; }
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

// out = A v; partials w1.out, out.out.  MODE 0: q = A phat (alpha); MODE 1: t = A shat (omega)
template <typename T, int VEC, bool TRANS, int MODE>
__global__ void __launch_bounds__(kBlock) bicg_spmv_kernel(MatArgs m, const T* __restrict__ vals,
                                                           const T* __restrict__ v, const T* __restrict__ w1,
                                                           T* __restrict__ out, const Scalars* sc, double* partials) {
  if (sc->done || (MODE == 1 && sc->exit_early)) return;
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return v[c]; };
  double loc[2] = {0, 0};
  auto epi = [&](int64_t n, int i, T o) {
    const int64_t d = n * VEC + i;
    out[d] = o;
    loc[0] += (double)w1[d] * (double)o;
    loc[1] += (double)o * (double)o;
  };
  sweep<T, VEC, TRANS>(m, r0, r1, vals, xv, epi);
  block_partials<2>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) bicg_alpha_scalars(Scalars* sc, const double* partials, int nparts) {
  if (sc->done) return;
  double tot[2];
  sum_partials<2>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  sc->alpha = sc->rho_ / tot[0];
  if (!isfinite(sc->alpha)) { sc->done = 1; sc->status = 3; }
}

__global__ void __launch_bounds__(kScalarBlock) bicg_omega_scalars(Scalars* sc, const double* partials, int nparts) {
  if (sc->done || sc->exit_early) return;
  double tot[2];
  sum_partials<2>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  sc->ts = tot[0];
  sc->tt = tot[1];
  sc->omega = tot[0] / tot[1];
}

// s = r - alpha q, shat = dinv s; partial s.s
template <typename T>
__global__ void __launch_bounds__(kBlock) bicg_s_kernel(int64_t N, const T* __restrict__ dinv,
                                                        const T* __restrict__ r, const T* __restrict__ q,
                                                        T* __restrict__ s, T* __restrict__ shat, const Scalars* sc,
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
  block_partials<1>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) bicg_s_scalars(Scalars* sc, const double* partials, int nparts) {
  if (sc->done) return;
  double tot[1];
  sum_partials<1>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  sc->ss = tot[0];
  sc->exit_early = (tot[0] < sc->thresh) ? 1 : 0;
}

// x += alpha phat (+ omega shat); r = s (- omega t); partials rhat.r, r.r
template <typename T>
__global__ void __launch_bounds__(kBlock) bicg_x_kernel(int64_t N, const T* __restrict__ rhat,
                                                        const T* __restrict__ phat, const T* __restrict__ shat,
                                                        const T* __restrict__ s, const T* __restrict__ t,
                                                        T* __restrict__ x, T* __restrict__ r, const Scalars* sc,
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
  block_partials<2>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) bicg_end_scalars(Scalars* sc, const double* partials, int nparts) {
  if (sc->done) return;
  double tot[2];
  sum_partials<2>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  const bool early = sc->exit_early != 0;
  sc->iter += 1;
  const double rho_used = sc->rho_;
  sc->rho = rho_used;
  sc->rho_ = tot[0];
  sc->rr = tot[1];
  if (!isfinite(tot[1])) { sc->done = 1; sc->status = 3; }
  else if (tot[1] <= sc->thresh || early) { sc->done = 1; sc->status = 0; }
  else if (sc->omega == 0.0 || sc->alpha == 0.0 || rho_used == 0.0) { sc->done = 1; sc->status = 2; }
  else if (sc->iter >= sc->maxiter) { sc->done = 1; sc->status = 1; }
}

// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) dot_kernel(int64_t n, const T* __restrict__ x, const T* __restrict__ y,
                                                     double* partials) {
  const int64_t per = ceil_div(n, gridDim.x);
  const int64_t i0 = (int64_t)blockIdx.x * per;
  const int64_t i1 = i0 + per < n ? i0 + per : n;
  double loc[1] = {0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBlock) loc[0] += (double)x[i] * (double)y[i];
  block_partials<1>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) dot_final(const double* partials, int nparts, double* out) {
  double tot[1];
  sum_partials<1>(partials, nparts, tot);
  if (threadIdx.x == 0) *out = tot[0];
}

// ---------------------------------------------------------------------------
// Optional HIP-event timing of the Krylov SpMV kernel (first SpMV of every poll
// batch), for roofline reporting; off unless hf_profile_enable(1) was called.
bool g_prof = false;
static double g_prof_ms = 0.0;
static int64_t g_prof_n = 0;
static hipEvent_t g_ev[2] = {nullptr, nullptr};

int prof_begin(hipStream_t s) {
  if (!g_prof) return 0;
  if (!g_ev[0]) {
    HF_CHECK_HIP(hipEventCreate(&g_ev[0]));
    HF_CHECK_HIP(hipEventCreate(&g_ev[1]));
  }
  HF_CHECK_HIP(hipEventRecord(g_ev[0], s));
  return 0;
}
// The pair of timing events, created on first use; a launch given to
// hipExtLaunchKernelGGL with them records at the kernel's own start and end.
int prof_events(hipEvent_t* e0, hipEvent_t* e1) {
  if (!g_ev[0]) {
    HF_CHECK_HIP(hipEventCreate(&g_ev[0]));
    HF_CHECK_HIP(hipEventCreate(&g_ev[1]));
  }
  *e0 = g_ev[0];
  *e1 = g_ev[1];
  return 0;
}
int prof_end(hipStream_t s) {
  if (!g_prof) return 0;
  HF_CHECK_HIP(hipEventRecord(g_ev[1], s));
  return 0;
}
int prof_collect() {
  if (!g_prof) return 0;
  float ms = 0.f;
  HF_CHECK_HIP(hipEventElapsedTime(&ms, g_ev[0], g_ev[1]));
  g_prof_ms += ms;
  g_prof_n += 1;
  return 0;
}

// Host side of the convergence poll: after each batch of iterations the
// scalars are copied to pinned host memory behind an event; the host launches
// the next batch before it waits for the previous batch's copy, so the GPU
// always has a batch queued.  One pinned pair + events per host thread.
static thread_local PollState g_poll;

int poll_init() {
  if (!g_poll.host) {
    HF_CHECK_HIP(hipHostMalloc((void**)&g_poll.host, 2 * sizeof(Scalars), hipHostMallocDefault));
    HF_CHECK_HIP(hipEventCreateWithFlags(&g_poll.ev[0], hipEventDisableTiming));
    HF_CHECK_HIP(hipEventCreateWithFlags(&g_poll.ev[1], hipEventDisableTiming));
  }
  return 0;
}
int poll_post(const Scalars* sc, int slot, hipStream_t s, PollState* ps) {
  PollState& P = ps ? *ps : g_poll;
  HF_CHECK_HIP(hipMemcpyAsync(&P.host[slot], sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
  HF_CHECK_HIP(hipEventRecord(P.ev[slot], s));
  return 0;
}
int poll_wait(int slot, Scalars& out, PollState* ps) {
  PollState& P = ps ? *ps : g_poll;
  HF_CHECK_HIP(hipEventSynchronize(P.ev[slot]));
  out = P.host[slot];
  return 0;
}


template <typename T, int VEC, bool TRANS>
static int spmv_launch(const hf_pattern* p, const T* vals, const T* x, T* y, hipStream_t s) {
  const MatArgs m = mat_args(p);
  spmv_kernel<T, VEC, TRANS><<<rows_grid(m.nslices), kBlock, 0, s>>>(m, vals, x, y);
  HF_LAUNCH_CHECK();
  return 0;
}

template <typename T, bool TRANS>
static int spmv_dispatch(const hf_pattern* p, const T* vals, const T* x, T* y, hipStream_t s) {
  HF_REQUIRE(p && vals && x && y, "hf_spmv: null argument");
  switch (pattern_vec(p)) {
    case 1: return spmv_launch<T, 1, TRANS>(p, vals, x, y, s);
    case 2: return spmv_launch<T, 2, TRANS>(p, vals, x, y, s);
    case 3: return spmv_launch<T, 3, TRANS>(p, vals, x, y, s);
    default: HF_REQUIRE(false, "hf_spmv: vec=%d not supported", pattern_vec(p));
  }
}

template <typename T, int VEC, bool TRANS>
static int pcg_run(const hf_pattern* p, const T* vals, const T* dinv, const T* b, T* x, int maxiter, double rtol,
                   double atol, int poll, int rel_r0, void* ws, hf_krylov_info* info, hipStream_t s,
                   PollState* ps) {
  const MatArgs m = mat_args(p);
  const int64_t N = m.nrows * VEC;
  Scalars* sc = (Scalars*)ws;
  double* partA = (double*)((char*)ws + 4096);
  double* partB = partA + kPartB;
  T* v = (T*)((char*)ws + kHeader);
  const size_t stride = align256(N * sizeof(T)) / sizeof(T);
  T *r = v, *u = v + stride, *pv = v + 2 * stride, *sv = v + 3 * stride, *w = v + 4 * stride;
  const int gr = rows_grid(m.nslices), ge = elem_grid(N);
  static_assert(kPartB + kMaxGrid <= 4 * kMaxGrid, "partials exceed the workspace header");
  cg_init_kernel<T, VEC, TRANS><<<gr, kBlock, 0, s>>>(m, vals, dinv, b, x, r, u, pv, sv, partA);
  cg_spmv_kernel<T, VEC, TRANS><<<gr, kBlock, 0, s>>>(m, vals, u, w, nullptr, partB);
  cg_init_scalars<<<1, kScalarBlock, 0, s>>>(sc, partA, gr, partB, gr, rtol * rtol, atol * atol, rel_r0, maxiter);
  HF_LAUNCH_CHECK();
  // ps != nullptr: the caller's poll state (no timing: it would create events)
  const bool prof_ok = ps == nullptr;
  if (!ps && poll_init()) return -2;
  Scalars host{};
  int it = 0, batch = 0;
  bool timed_pending = false;
  int timed_batch = -1;
  for (;;) {
    for (int k = 0; k < poll; ++k, ++it) {
      const bool timed = prof_ok && k == poll / 2 && !timed_pending;  // a steady-state launch, mid-batch
      cg_update_kernel<T><<<ge, kBlock, 0, s>>>(N, dinv, u, w, pv, sv, x, r, sc, partA);
      if (timed && g_prof) {
        hipEvent_t e0, e1;
        if (prof_events(&e0, &e1)) return -2;
        hipExtLaunchKernelGGL((cg_spmv_kernel<T, VEC, TRANS>), dim3(gr), dim3(kBlock), 0, s, e0, e1, 0, m, vals, u,
                              w, (const Scalars*)sc, partB);
        timed_pending = true;
        timed_batch = batch;
      } else {
        cg_spmv_kernel<T, VEC, TRANS><<<gr, kBlock, 0, s>>>(m, vals, u, w, sc, partB);
      }
      cg_scalars<<<1, kScalarBlock, 0, s>>>(sc, partA, ge, partB, gr);
    }
    HF_LAUNCH_CHECK();
    if (poll_post(sc, batch & 1, s, ps)) return -2;
    if (batch > 0) {
      // wait for the previous batch while this one runs
      if (poll_wait((batch - 1) & 1, host, ps)) return -2;
      if (timed_pending && timed_batch == batch - 1) {
        // the timed SpMV ran in the batch just waited for; it did work unless
        // the solve had already finished before it
        if (!host.done || host.iter > (batch - 1) * poll + poll / 2) {
          if (prof_collect()) return -2;
        }
        timed_pending = false;
      }
      if (host.done) break;
    }
    ++batch;
    if (it >= maxiter + 2 * poll) {
      if (poll_wait((batch - 1) & 1, host, ps)) return -2;
      break;
    }
  }
  HF_CHECK_HIP(hipStreamSynchronize(s));
  HF_CHECK_HIP(hipMemcpy(&host, sc, sizeof(Scalars), hipMemcpyDeviceToHost));
  timed_pending = false;
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
                    double atol, int poll, int rel_r0, void* ws, hf_krylov_info* info, hipStream_t s,
                    PollState* ps) {
  const MatArgs m = mat_args(p);
  const int64_t N = m.nrows * VEC;
  Scalars* sc = (Scalars*)ws;
  double* partials = (double*)((char*)ws + 4096);
  T* v = (T*)((char*)ws + kHeader);
  const size_t stride = align256(N * sizeof(T)) / sizeof(T);
  T *r = v, *rhat = v + stride, *pp = v + 2 * stride, *q = v + 3 * stride, *phat = v + 4 * stride,
    *sv = v + 5 * stride, *shat = v + 6 * stride, *t = v + 7 * stride;
  const int gr = rows_grid(m.nslices), ge = elem_grid(N);
  bicg_init_kernel<T, VEC, TRANS><<<gr, kBlock, 0, s>>>(m, vals, b, x, r, rhat, pp, q, partials);
  bicg_init_scalars<<<1, kScalarBlock, 0, s>>>(sc, partials, gr, rtol * rtol, atol * atol, rel_r0, maxiter);
  HF_LAUNCH_CHECK();
  // ps != nullptr: the caller's poll state (no timing: it would create events)
  const bool prof_ok = ps == nullptr;
  if (!ps && poll_init()) return -2;
  Scalars host{};
  int it = 0, batch = 0;
  bool timed_pending = false;
  int timed_batch = -1;
  for (;;) {
    for (int k = 0; k < poll; ++k, ++it) {
      const bool timed = prof_ok && k == poll / 2 && !timed_pending;  // a steady-state launch, mid-batch
      bicg_p_kernel<T><<<ge, kBlock, 0, s>>>(N, dinv, r, q, pp, phat, sc);
      if (timed && prof_begin(s)) return -2;
      bicg_spmv_kernel<T, VEC, TRANS, 0><<<gr, kBlock, 0, s>>>(m, vals, phat, rhat, q, sc, partials);
      if (timed && prof_end(s)) return -2;
      if (timed && g_prof) {
        timed_pending = true;
        timed_batch = batch;
      }
      bicg_alpha_scalars<<<1, kScalarBlock, 0, s>>>(sc, partials, gr);
      bicg_s_kernel<T><<<ge, kBlock, 0, s>>>(N, dinv, r, q, sv, shat, sc, partials);
      bicg_s_scalars<<<1, kScalarBlock, 0, s>>>(sc, partials, ge);
      bicg_spmv_kernel<T, VEC, TRANS, 1><<<gr, kBlock, 0, s>>>(m, vals, shat, sv, t, sc, partials);
      bicg_omega_scalars<<<1, kScalarBlock, 0, s>>>(sc, partials, gr);
      bicg_x_kernel<T><<<ge, kBlock, 0, s>>>(N, rhat, phat, shat, sv, t, x, r, sc, partials);
      bicg_end_scalars<<<1, kScalarBlock, 0, s>>>(sc, partials, ge);
    }
    HF_LAUNCH_CHECK();
    if (poll_post(sc, batch & 1, s, ps)) return -2;
    if (batch > 0) {
      if (poll_wait((batch - 1) & 1, host, ps)) return -2;
      if (timed_pending && timed_batch == batch - 1) {
        // the timed SpMV ran in the batch just waited for; it did work unless
        // the solve had already finished before it
        if (!host.done || host.iter > (batch - 1) * poll + poll / 2) {
          if (prof_collect()) return -2;
        }
        timed_pending = false;
      }
      if (host.done) break;
    }
    ++batch;
    if (it >= maxiter + 2 * poll) {
      if (poll_wait((batch - 1) & 1, host, ps)) return -2;
      break;
    }
  }
  HF_CHECK_HIP(hipStreamSynchronize(s));
  HF_CHECK_HIP(hipMemcpy(&host, sc, sizeof(Scalars), hipMemcpyDeviceToHost));
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
                           int flags, int maxiter, double rtol, double atol, int poll, void* ws,
                           hf_krylov_info* info, hipStream_t s, PollState* ps = nullptr) {
  const int transpose = flags & HF_KRYLOV_TRANSPOSE;
  const int rel_r0 = (flags & HF_KRYLOV_REL_R0) ? 1 : 0;
  HF_REQUIRE(p && vals && dinv && b && x && ws, "hf_krylov_solve: null argument");
  HF_REQUIRE(maxiter >= 0, "hf_krylov_solve: maxiter < 0");
  if (poll < 1) poll = 1;
  const int vec = pattern_vec(p);
  HF_REQUIRE(vec >= 1 && vec <= 3, "hf_krylov_solve: vec=%d not supported", vec);
  HF_REQUIRE(method == HF_KRYLOV_PCG || method == HF_KRYLOV_BICGSTAB, "hf_krylov_solve: unknown method %d", method);
#define KRY(M, V, TR) M<T, V, TR>(p, vals, dinv, b, x, maxiter, rtol, atol, poll, rel_r0, ws, info, s, ps)
  if (method == HF_KRYLOV_PCG) {
    if (vec == 3) return transpose ? KRY(pcg_run, 3, true) : KRY(pcg_run, 3, false);
    if (vec == 2) return transpose ? KRY(pcg_run, 2, true) : KRY(pcg_run, 2, false);
    return transpose ? KRY(pcg_run, 1, true) : KRY(pcg_run, 1, false);
  }
  if (vec == 3) return transpose ? KRY(bicg_run, 3, true) : KRY(bicg_run, 3, false);
  if (vec == 2) return transpose ? KRY(bicg_run, 2, true) : KRY(bicg_run, 2, false);
  return transpose ? KRY(bicg_run, 1, true) : KRY(bicg_run, 1, false);
#undef KRY
}

}  // namespace hf

using namespace hf;

extern "C" size_t hf_krylov_workspace_size(int32_t method, int64_t N, int32_t elem_bytes) {
  const int nvec = (method == HF_KRYLOV_BICGSTAB) ? 8 : 5;
  return kHeader + (size_t)nvec * align256((size_t)N * elem_bytes);
}

extern "C" int hf_krylov_solve_f64(int32_t method, const hf_pattern* p, const double* vals, const double* dinv,
                                   const double* b, double* x, int32_t flags, int32_t maxiter, double rtol,
                                   double atol, int32_t poll, void* ws, hf_krylov_info* info, hf_stream_t s) {
  return krylov_dispatch<double>(method, p, vals, dinv, b, x, flags, maxiter, rtol, atol, poll, ws, info,
                                 (hipStream_t)s);
}
extern "C" int hf_krylov_solve_f32(int32_t method, const hf_pattern* p, const float* vals, const float* dinv,
                                   const float* b, float* x, int32_t flags, int32_t maxiter, double rtol,
                                   double atol, int32_t poll, void* ws, hf_krylov_info* info, hf_stream_t s) {
  return krylov_dispatch<float>(method, p, vals, dinv, b, x, flags, maxiter, rtol, atol, poll, ws, info,
                                (hipStream_t)s);
}

// ---------------------------------------------------------------------------
// Allocation-free Krylov solve.  The workspace is a file-scope device array
// (every GPU the library runs on holds its own copy, at its own address, until
// the process ends).  hf_init() resolves that address once per device
// with hipGetSymbolAddress and makes the device's poll state (two pinned
// scalar slots, two events); after that the entry point itself calls no
// allocating routine.  hf_init is the library's set-up call for the current
// device; the other entry points do not need it.  In host code `scratch` names
// a host-side stand-in, never the device memory: only the resolved pointer is
// handed to kernels.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
struct NoAllocDevice {
  void* ptr = nullptr;  // device address of this GPU's `scratch`
  PollState poll;
  bool ready = false;
};
NoAllocDevice g_noalloc[kMaxDevices];

// the 4 KB scalar header starts from zeros (a kernel, not a memset node, so a
// caller may capture it)
__global__ void clear_header_kernel(uint32_t* p, int nwords) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) p[i] = 0u;
}
}  // namespace

template <typename T>
static int krylov_noalloc(int method, const hf_pattern* p, const T* vals, const T* dinv, const T* b, T* x,
                          int flags, int maxiter, double rtol, double atol, int poll, hf_krylov_info* info,
                          hipStream_t s) {
  HF_REQUIRE(p, "hf_krylov_solve_noalloc_f64: null pattern");
  int dev = -1;
  HF_CHECK_HIP(hipGetDevice(&dev));
  HF_REQUIRE(dev >= 0 && dev < kMaxDevices && g_noalloc[dev].ready,
             "hf_krylov_solve_noalloc_f64: call hf_init() with device %d current first", dev);
  HF_REQUIRE(method == HF_KRYLOV_PCG || method == HF_KRYLOV_BICGSTAB,
             "hf_krylov_solve_noalloc_f64: unknown method %d", method);
  const int64_t N = pattern_nnode(p) * pattern_vec(p);
  const size_t need = hf_krylov_workspace_size(method, N, (int32_t)sizeof(T));
  HF_REQUIRE(need <= kScratchBytes,
             "hf_krylov_solve_noalloc_f64: N = %lld needs %zu bytes of workspace; the static scratch holds %zu "
             "(use hf_krylov_solve_f64 with a caller workspace)", (long long)N, need, kScratchBytes);
  NoAllocDevice& d = g_noalloc[dev];
  clear_header_kernel<<<(1024 + 255) / 256, 256, 0, s>>>((uint32_t*)d.ptr, 1024);
  HF_LAUNCH_CHECK();
  return krylov_dispatch<T>(method, p, vals, dinv, b, x, flags, maxiter, rtol, atol, poll, d.ptr, info, s, &d.poll);
}

extern "C" int hf_init(void) {
  int dev = -1;
  HF_CHECK_HIP(hipGetDevice(&dev));
  HF_REQUIRE(dev >= 0 && dev < kMaxDevices, "hf_init: device %d out of range", dev);
  NoAllocDevice& d = g_noalloc[dev];
  if (d.ready) return 0;
  void* ptr = nullptr;
  HF_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  HF_REQUIRE(ptr != nullptr, "hf_init: no device address for the scratch array");
  if (!d.poll.host) HF_CHECK_HIP(hipHostMalloc((void**)&d.poll.host, 2 * sizeof(Scalars), hipHostMallocDefault));
  if (!d.poll.ev[0]) HF_CHECK_HIP(hipEventCreateWithFlags(&d.poll.ev[0], hipEventDisableTiming));
  if (!d.poll.ev[1]) HF_CHECK_HIP(hipEventCreateWithFlags(&d.poll.ev[1], hipEventDisableTiming));
  d.ptr = ptr;
  d.ready = true;
  return 0;
}

extern "C" int hf_krylov_solve_noalloc_f64(int32_t method, const hf_pattern* p, const double* vals,
                                           const double* dinv, const double* b, double* x, int32_t flags,
                                           int32_t maxiter, double rtol, double atol, int32_t poll,
                                           hf_krylov_info* info, hf_stream_t s) {
  return krylov_noalloc<double>(method, p, vals, dinv, b, x, flags, maxiter, rtol, atol, poll, info, (hipStream_t)s);
}
extern "C" int hf_spmv_f64(const hf_pattern* p, const double* a, const double* x, double* y, hf_stream_t s) {
  return spmv_dispatch<double, false>(p, a, x, y, (hipStream_t)s);
}
extern "C" int hf_spmv_f32(const hf_pattern* p, const float* a, const float* x, float* y, hf_stream_t s) {
  return spmv_dispatch<float, false>(p, a, x, y, (hipStream_t)s);
}
extern "C" int hf_spmv_t_f64(const hf_pattern* p, const double* a, const double* x, double* y, hf_stream_t s) {
  return spmv_dispatch<double, true>(p, a, x, y, (hipStream_t)s);
}
extern "C" int hf_spmv_t_f32(const hf_pattern* p, const float* a, const float* x, float* y, hf_stream_t s) {
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

extern "C" size_t hf_dot_workspace_size(void) { return 4096 + (size_t)kMaxGrid * sizeof(double); }

template <typename T>
static int dot_impl(const T* x, const T* y, int64_t n, void* ws, double* result, hipStream_t s) {
  HF_REQUIRE(x && y && ws && result, "hf_dot: null argument");
  double* out = (double*)ws;
  double* partials = (double*)((char*)ws + 4096);
  const int g = elem_grid(n);
  dot_kernel<T><<<g, kBlock, 0, s>>>(n, x, y, partials);
  dot_final<<<1, kScalarBlock, 0, s>>>(partials, g, out);
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
