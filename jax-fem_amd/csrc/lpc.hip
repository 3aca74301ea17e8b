// Learned preconditioner: a 3-layer 3x3x3 convolutional network on the
// structured node grid, its training kernels, and a flexible-CG solve that
// applies it (no reference counterpart: the reference preconditions with
// Jacobi only, solver.py:68-69).
//
// Node grid: nodes (i, j, k) of a structured box in box_mesh order
// n = (i*ny + j)*nz + k (jax_fem/generate_mesh.py:170-187); tensors are
// channels-last (batch, n, channel), so the displacement DOF vector (3 per
// node, interleaved) is already a 3-channel grid.  Zero padding outside.
// Weights W[tap][ci][co], tap = ((di+1)*3 + (dj+1))*3 + (dk+1).
//
// Preconditioner (D = diag A, mask = 0 on Dirichlet DOFs, s = rms(D^-1/2 r)):
//   f  = [D^-1/2 r / s (3 ch), g (1 ch, normalised log-modulus)]
//   h1 = tanh(conv1(f)), h2 = tanh(conv2(h1)), y = conv3(h2)   (4->16->16->3)
//   M(r) = D^-1 r + s * D^-1/2 * mask * y
// so M is Jacobi plus a learned correction, exactly zero on Dirichlet rows and
// positively homogeneous in r.  M is nonlinear, so the Krylov method is
// flexible CG (Polak-Ribiere beta); a breakdown, a non-finite value or
// stagnation ends the flexible solve and Jacobi-PCG continues from its iterate.
//
// Kernels are FP32 on the vector ALU: a thread per node holds COUT
// accumulators; the layer's weights are staged in LDS and read as broadcasts.
#include "krylov_common.hpp"

namespace hf {

struct CnnGrid {
  int32_t nx, ny, nz, batch;
};

__device__ __forceinline__ bool grid_nbr(const CnnGrid& g, int64_t n, int tap, int64_t& m) {
  const int64_t plane = (int64_t)g.ny * g.nz;
  const int i = (int)(n / plane), j = (int)((n / g.nz) % g.ny), k = (int)(n % g.nz);
  const int di = tap / 9 - 1, dj = (tap / 3) % 3 - 1, dk = tap % 3 - 1;
  const int ii = i + di, jj = j + dj, kk = k + dk;
  if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) return false;
  m = ((int64_t)ii * g.ny + jj) * g.nz + kk;
  return true;
}

template <int C>
__device__ __forceinline__ void load_ch(const float* __restrict__ p, float (&x)[C]) {
  if constexpr (C % 4 == 0) {
#pragma unroll
    for (int c = 0; c < C; c += 4) {
      const float4 v = *reinterpret_cast<const float4*>(p + c);
      x[c] = v.x; x[c + 1] = v.y; x[c + 2] = v.z; x[c + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) x[c] = p[c];
  }
}

// out[b, n, co] = act(bias[co] + sum_{tap, ci} W[tap][ci][co] in[b, n + off(tap), ci])
template <int CIN, int COUT, int ACT>
__global__ void __launch_bounds__(256) conv3d_fwd(const float* __restrict__ in, const float* __restrict__ w,
                                                  const float* __restrict__ bias, CnnGrid g,
                                                  float* __restrict__ out, const int32_t* skip) {
  __shared__ float sw[27 * CIN * COUT];
  if (skip && *skip) return;
  for (int e = threadIdx.x; e < 27 * CIN * COUT; e += blockDim.x) sw[e] = w[e];
  __syncthreads();
  const int64_t nn = (int64_t)g.nx * g.ny * g.nz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn * g.batch) return;
  const int64_t b = t / nn, n = t - b * nn;
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; ++co) acc[co] = bias[co];
  for (int tap = 0; tap < 27; ++tap) {
    int64_t m;
    if (!grid_nbr(g, n, tap, m)) continue;
    float x[CIN];
    load_ch<CIN>(in + (b * nn + m) * CIN, x);
    const float* wt = sw + tap * CIN * COUT;
#pragma unroll
    for (int ci = 0; ci < CIN; ++ci)
#pragma unroll
      for (int co = 0; co < COUT; ++co) acc[co] += wt[ci * COUT + co] * x[ci];
  }
  float* o = out + t * COUT;
#pragma unroll
  for (int co = 0; co < COUT; ++co) o[co] = ACT ? tanhf(acc[co]) : acc[co];
}

// gpre = gout * (1 - y^2) (tanh backward), elementwise
__global__ void tanh_bwd(const float* __restrict__ y, const float* __restrict__ gout, int64_t n,
                         float* __restrict__ gpre) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  gpre[t] = gout[t] * (1.0f - y[t] * y[t]);
}

// gin[b, m, ci] = sum_{tap, co} W[tap][ci][co] gpre[b, m - off(tap), co]
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) conv3d_bwd_data(const float* __restrict__ gpre, const float* __restrict__ w,
                                                       CnnGrid g, float* __restrict__ gin) {
  __shared__ float sw[27 * CIN * COUT];
  for (int e = threadIdx.x; e < 27 * CIN * COUT; e += blockDim.x) sw[e] = w[e];
  __syncthreads();
  const int64_t nn = (int64_t)g.nx * g.ny * g.nz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn * g.batch) return;
  const int64_t b = t / nn, m = t - b * nn;
  float acc[CIN];
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) acc[ci] = 0.f;
  for (int tap = 0; tap < 27; ++tap) {
    int64_t n;
    if (!grid_nbr(g, m, 26 - tap, n)) continue;  // n = m - off(tap)
    float gp[COUT];
    load_ch<COUT>(gpre + (b * nn + n) * COUT, gp);
    const float* wt = sw + tap * CIN * COUT;
#pragma unroll
    for (int ci = 0; ci < CIN; ++ci)
#pragma unroll
      for (int co = 0; co < COUT; ++co) acc[ci] += wt[ci * COUT + co] * gp[co];
  }
  float* o = gin + t * CIN;
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) o[ci] = acc[ci];
}

// Weight/bias gradient partials: workgroup c covers nodes [c*chunk, (c+1)*chunk)
// of the flattened (batch, node) range; thread (ci, co) accumulates the 27 taps.
// partial layout: part[c][(tap*CIN + ci)*COUT + co], bias part[c][27*CIN*COUT + co].
template <int CIN, int COUT>
__global__ void __launch_bounds__(CIN * COUT) conv3d_bwd_weight_part(const float* __restrict__ in,
                                                                     const float* __restrict__ gpre, CnnGrid g,
                                                                     int64_t chunk, float* __restrict__ part) {
  constexpr int NB = 8;  // nodes staged per step
  __shared__ float sin_[NB][27][CIN];
  __shared__ float sg[NB][COUT];
  const int ci = threadIdx.x / COUT, co = threadIdx.x - (threadIdx.x / COUT) * COUT;
  const int64_t nn = (int64_t)g.nx * g.ny * g.nz, total = nn * g.batch;
  const int64_t t0 = (int64_t)blockIdx.x * chunk;
  const int64_t t1 = t0 + chunk < total ? t0 + chunk : total;
  float acc[27];
#pragma unroll
  for (int tap = 0; tap < 27; ++tap) acc[tap] = 0.f;
  float accb = 0.f;
  for (int64_t base = t0; base < t1; base += NB) {
    // stage NB nodes' gradients and their 27 neighbours' inputs
    for (int e = threadIdx.x; e < NB * 27 * CIN; e += CIN * COUT) {
      const int q = e / (27 * CIN), rem = e - q * 27 * CIN, tap = rem / CIN, c = rem - (rem / CIN) * CIN;
      const int64_t t = base + q;
      float v = 0.f;
      if (t < t1) {
        const int64_t b = t / nn, n = t - b * nn;
        int64_t m;
        if (grid_nbr(g, n, tap, m)) v = in[(b * nn + m) * CIN + c];
      }
      sin_[q][tap][c] = v;
    }
    for (int e = threadIdx.x; e < NB * COUT; e += CIN * COUT) {
      const int q = e / COUT, c = e - (e / COUT) * COUT;
      const int64_t t = base + q;
      sg[q][c] = t < t1 ? gpre[t * COUT + c] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const float gv = sg[q][co];
      if (ci == 0) accb += gv;
#pragma unroll
      for (int tap = 0; tap < 27; ++tap) acc[tap] += sin_[q][tap][ci] * gv;
    }
    __syncthreads();
  }
  float* p = part + (int64_t)blockIdx.x * (27 * CIN * COUT + COUT);
#pragma unroll
  for (int tap = 0; tap < 27; ++tap) p[(tap * CIN + ci) * COUT + co] = acc[tap];
  if (ci == 0) p[27 * CIN * COUT + co] = accb;
}

// sum partials over workgroups in a fixed order
__global__ void sum_parts(const float* __restrict__ part, int nparts, int width, float* __restrict__ gw,
                          float* __restrict__ gb, int nw) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= width) return;
  float s = 0.f;
  for (int c = 0; c < nparts; ++c) s += part[(int64_t)c * width + e];
  if (e < nw) gw[e] = s;
  else gb[e - nw] = s;
}

// ---- preconditioner input / output ------------------------------------------
// f[b, n] = (r[3n+i] sqrt(dinv[3n+i]) / s, i<3 ; gfield[n])
__global__ void lpc_features(const float* __restrict__ r, const float* __restrict__ dinv,
                             const float* __restrict__ gfield, int64_t nnode, int batch,
                             const float* __restrict__ scale, float* __restrict__ f, const int32_t* skip) {
  if (skip && *skip) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode * batch) return;
  const int64_t b = t / nnode, n = t - b * nnode;
  const float inv_s = 1.0f / scale[b];
  float4 v;
  v.x = r[t * 3 + 0] * sqrtf(dinv[n * 3 + 0]) * inv_s;
  v.y = r[t * 3 + 1] * sqrtf(dinv[n * 3 + 1]) * inv_s;
  v.z = r[t * 3 + 2] * sqrtf(dinv[n * 3 + 2]) * inv_s;
  v.w = gfield[n];
  *reinterpret_cast<float4*>(f + t * 4) = v;
}

// z = dinv r + s sqrt(dinv) mask y  (y = last conv output, 3 channels)
__global__ void lpc_output(const float* __restrict__ r, const float* __restrict__ dinv,
                           const float* __restrict__ mask, const float* __restrict__ y, int64_t nnode, int batch,
                           const float* __restrict__ scale, float* __restrict__ z) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode * 3 * batch) return;
  const int64_t b = t / (nnode * 3), d = t - b * nnode * 3;
  z[t] = dinv[d] * r[t] + scale[b] * sqrtf(dinv[d]) * mask[d] * y[t];
}

// ---- flexible CG with the learned preconditioner ----------------------------
struct LpcArgs {
  const float *w1, *b1, *w2, *b2, *w3, *b3;
  const float* gfield;  // nnode
  const float* mask;    // N (1 interior, 0 Dirichlet)
  CnnGrid grid;         // batch = 1
};

struct FcgScalars {
  float s;       // rms of D^-1/2 r
  int32_t stall;
  double best_rr;
};

// init: r = b - A x; partials r.r, b.b, r.D^-1.r
template <int VEC>
__global__ void __launch_bounds__(kBlock) fcg_init_kernel(MatArgs m, const float* __restrict__ vals,
                                                          const float* __restrict__ dinv, const float* __restrict__ b,
                                                          const float* __restrict__ x, float* __restrict__ r,
                                                          float* __restrict__ z, float* __restrict__ p_old,
                                                          double* partials) {
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return x[c]; };
  double loc[3] = {0, 0, 0};
  auto epi = [&](int64_t n, int i, float ax) {
    const int64_t d = n * VEC + i;
    const float rr = b[d] - ax;
    r[d] = rr;
    z[d] = 0.f;
    p_old[d] = 0.f;
    loc[0] += (double)rr * (double)rr;
    loc[1] += (double)b[d] * (double)b[d];
    loc[2] += (double)rr * (double)rr * (double)dinv[d];
  };
  sweep<float, VEC, false>(m, r0, r1, vals, xv, epi);
  block_partials<3>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) fcg_init_scalars(Scalars* sc, FcgScalars* fs, float* scale,
                                                                 const double* partials, int nparts, int64_t N,
                                                                 double rtol2, double atol2, int rel_r0,
                                                                 int maxiter) {
  double tot[3];
  sum_partials<3>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  sc->rr = tot[0];
  sc->bb = tot[1];
  sc->thresh = fmax(rtol2 * (rel_r0 ? tot[0] : tot[1]), atol2);
  sc->rho = 0;     // r.z of the previous step (0: first step, beta = 0)
  sc->omega = 0;   // beta
  sc->alpha = 0;
  sc->iter = 0;
  sc->status = 0;
  sc->maxiter = maxiter;
  sc->done = (tot[0] <= sc->thresh) ? 1 : 0;
  if (!isfinite(tot[0])) { sc->done = 1; sc->status = 3; }
  else if (maxiter <= 0 && !sc->done) { sc->done = 1; sc->status = 1; }
  fs->s = (float)sqrt(fmax(tot[2], 1e-300) / (double)N);
  fs->stall = 0;
  fs->best_rr = tot[0];
  *scale = fs->s;
}

// conv3 fused with the output: z_new = dinv r + s sqrt(dinv) mask y; partials r.z_new, r.z_old
template <int CIN>
__global__ void __launch_bounds__(256) fcg_conv3_out(const float* __restrict__ in, const float* __restrict__ w,
                                                     const float* __restrict__ bias, CnnGrid g,
                                                     const float* __restrict__ r, const float* __restrict__ dinv,
                                                     const float* __restrict__ mask, const FcgScalars* fs,
                                                     const Scalars* sc, float* __restrict__ z, double* partials) {
  __shared__ float sw[27 * CIN * 3];
  if (sc->done) return;
  for (int e = threadIdx.x; e < 27 * CIN * 3; e += blockDim.x) sw[e] = w[e];
  __syncthreads();
  const int64_t nn = (int64_t)g.nx * g.ny * g.nz;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double loc[2] = {0, 0};
  if (n < nn) {
    float acc[3] = {bias[0], bias[1], bias[2]};
    for (int tap = 0; tap < 27; ++tap) {
      int64_t m;
      if (!grid_nbr(g, n, tap, m)) continue;
      float x[CIN];
      load_ch<CIN>(in + m * CIN, x);
      const float* wt = sw + tap * CIN * 3;
#pragma unroll
      for (int ci = 0; ci < CIN; ++ci)
#pragma unroll
        for (int co = 0; co < 3; ++co) acc[co] += wt[ci * 3 + co] * x[ci];
    }
    const float s = fs->s;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int64_t d = n * 3 + i;
      const float zn = dinv[d] * r[d] + s * sqrtf(dinv[d]) * mask[d] * acc[i];
      loc[0] += (double)r[d] * (double)zn;
      loc[1] += (double)r[d] * (double)z[d];
      z[d] = zn;
    }
  }
  block_partials<2>(loc, partials);
}

// beta (Polak-Ribiere) = (r.z_new - r.z_old) / rz_prev; rho = r.z_new
__global__ void __launch_bounds__(kScalarBlock) fcg_beta_scalars(Scalars* sc, const double* partials, int nparts) {
  if (sc->done) return;
  double tot[2];
  sum_partials<2>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  const double rz = tot[0];
  if (!(rz > 0.0) || !isfinite(rz)) {  // M not positive on r: hand over to Jacobi-PCG
    sc->done = 1;
    sc->status = isfinite(rz) ? 2 : 3;
    return;
  }
  sc->omega = (sc->iter == 0 || sc->rho == 0.0) ? 0.0 : fmax((rz - tot[1]) / sc->rho, 0.0);
  sc->rho = rz;
}

// x += alpha p, r -= alpha q; partials r.r, r.D^-1.r
__global__ void __launch_bounds__(kBlock) fcg_update_kernel(int64_t N, const float* __restrict__ dinv,
                                                            const float* __restrict__ p, const float* __restrict__ q,
                                                            float* __restrict__ x, float* __restrict__ r,
                                                            const Scalars* sc, double* partials) {
  if (sc->done) return;
  const float alpha = (float)sc->alpha;
  const int64_t per = ceil_div(N, gridDim.x);
  const int64_t i0 = (int64_t)logical_block() * per;
  const int64_t i1 = i0 + per < N ? i0 + per : N;
  double loc[2] = {0, 0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBlock) {
    x[i] += alpha * p[i];
    const float rr = r[i] - alpha * q[i];
    r[i] = rr;
    loc[0] += (double)rr * (double)rr;
    loc[1] += (double)rr * (double)rr * (double)dinv[i];
  }
  block_partials<2>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) fcg_end_scalars(Scalars* sc, FcgScalars* fs, float* scale,
                                                                const double* partials, int nparts, int64_t N,
                                                                int stall_window) {
  if (sc->done) return;
  double tot[2];
  sum_partials<2>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  sc->iter += 1;
  sc->rr = tot[0];
  fs->s = (float)sqrt(fmax(tot[1], 1e-300) / (double)N);
  *scale = fs->s;
  if (tot[0] < 0.81 * fs->best_rr) {  // progress: ||r|| down by 10 %
    fs->best_rr = tot[0];
    fs->stall = 0;
  } else {
    fs->stall += 1;
  }
  if (!isfinite(tot[0])) { sc->done = 1; sc->status = 3; }
  else if (tot[0] <= sc->thresh) { sc->done = 1; sc->status = 0; }
  else if (fs->stall >= stall_window) { sc->done = 1; sc->status = 4; }
  else if (sc->iter >= sc->maxiter) { sc->done = 1; sc->status = 1; }
}

}  // namespace hf

using namespace hf;

static inline CnnGrid mkgrid(int32_t nx, int32_t ny, int32_t nz, int32_t batch) {
  CnnGrid g;
  g.nx = nx; g.ny = ny; g.nz = nz; g.batch = batch;
  return g;
}

template <int CIN, int COUT>
static int conv_fwd_dispatch2(const float* in, const float* w, const float* b, int act, CnnGrid g, float* out,
                              hipStream_t s) {
  const int64_t n = (int64_t)g.nx * g.ny * g.nz * g.batch;
  const unsigned grid = (unsigned)ceil_div(n, 256);
  if (act) conv3d_fwd<CIN, COUT, 1><<<grid, 256, 0, s>>>(in, w, b, g, out, nullptr);
  else conv3d_fwd<CIN, COUT, 0><<<grid, 256, 0, s>>>(in, w, b, g, out, nullptr);
  HF_LAUNCH_CHECK();
  return 0;
}

#define HF_CNN_SHAPES(X) X(4, 16) X(16, 16) X(16, 3)

extern "C" int hf_cnn_conv3d_fwd_f32(const float* in, int32_t cin, const float* w, const float* b, int32_t cout,
                                     int32_t act, int32_t nx, int32_t ny, int32_t nz, int32_t batch, float* out,
                                     hf_stream_t stream) {
  HF_REQUIRE(in && w && b && out, "hf_cnn_conv3d_fwd_f32: null argument");
  const CnnGrid g = mkgrid(nx, ny, nz, batch);
  hipStream_t s = (hipStream_t)stream;
#define X(CI, CO) if (cin == CI && cout == CO) return conv_fwd_dispatch2<CI, CO>(in, w, b, act, g, out, s);
  HF_CNN_SHAPES(X)
#undef X
  HF_REQUIRE(false, "hf_cnn_conv3d_fwd_f32: layer %dx%d not built", cin, cout);
}

extern "C" int hf_cnn_conv3d_bwd_data_f32(const float* gpre, int32_t cout, const float* w, int32_t cin, int32_t nx,
                                          int32_t ny, int32_t nz, int32_t batch, float* gin, hf_stream_t stream) {
  HF_REQUIRE(gpre && w && gin, "hf_cnn_conv3d_bwd_data_f32: null argument");
  const CnnGrid g = mkgrid(nx, ny, nz, batch);
  const int64_t n = (int64_t)nx * ny * nz * batch;
  hipStream_t s = (hipStream_t)stream;
#define X(CI, CO)                                                                                         \
  if (cin == CI && cout == CO) {                                                                          \
    conv3d_bwd_data<CI, CO><<<(unsigned)ceil_div(n, 256), 256, 0, s>>>(gpre, w, g, gin);                   \
    HF_LAUNCH_CHECK();                                                                                    \
    return 0;                                                                                             \
  }
  HF_CNN_SHAPES(X)
#undef X
  HF_REQUIRE(false, "hf_cnn_conv3d_bwd_data_f32: layer %dx%d not built", cin, cout);
}

extern "C" size_t hf_cnn_bwd_weight_workspace(int32_t cin, int32_t cout, int32_t nx, int32_t ny, int32_t nz,
                                              int32_t batch) {
  const int64_t n = (int64_t)nx * ny * nz * batch;
  const int64_t nparts = ceil_div(n, 1024);
  return (size_t)nparts * (27 * cin * cout + cout) * sizeof(float);
}

extern "C" int hf_cnn_conv3d_bwd_weight_f32(const float* in, int32_t cin, const float* gpre, int32_t cout,
                                            int32_t nx, int32_t ny, int32_t nz, int32_t batch, void* workspace,
                                            float* gw, float* gb, hf_stream_t stream) {
  HF_REQUIRE(in && gpre && workspace && gw && gb, "hf_cnn_conv3d_bwd_weight_f32: null argument");
  const CnnGrid g = mkgrid(nx, ny, nz, batch);
  const int64_t n = (int64_t)nx * ny * nz * batch;
  const int64_t chunk = 1024;
  const int nparts = (int)ceil_div(n, chunk);
  const int width = 27 * cin * cout + cout;
  float* part = (float*)workspace;
  hipStream_t s = (hipStream_t)stream;
#define X(CI, CO)                                                                                          \
  if (cin == CI && cout == CO) {                                                                           \
    conv3d_bwd_weight_part<CI, CO><<<nparts, CI * CO, 0, s>>>(in, gpre, g, chunk, part);                    \
    sum_parts<<<(unsigned)ceil_div(width, 256), 256, 0, s>>>(part, nparts, width, gw, gb, 27 * CI * CO);   \
    HF_LAUNCH_CHECK();                                                                                     \
    return 0;                                                                                              \
  }
  HF_CNN_SHAPES(X)
#undef X
  HF_REQUIRE(false, "hf_cnn_conv3d_bwd_weight_f32: layer %dx%d not built", cin, cout);
}

extern "C" int hf_cnn_tanh_bwd_f32(const float* y, const float* gout, int64_t n, float* gpre, hf_stream_t stream) {
  HF_REQUIRE(y && gout && gpre, "hf_cnn_tanh_bwd_f32: null argument");
  tanh_bwd<<<(unsigned)ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(y, gout, n, gpre);
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_lpc_features_f32(const float* r, const float* dinv, const float* gfield, int64_t nnode,
                                   int32_t batch, const float* scale, float* f, hf_stream_t stream) {
  HF_REQUIRE(r && dinv && gfield && scale && f, "hf_lpc_features_f32: null argument");
  lpc_features<<<(unsigned)ceil_div(nnode * batch, 256), 256, 0, (hipStream_t)stream>>>(r, dinv, gfield, nnode,
                                                                                         batch, scale, f, nullptr);
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_lpc_output_f32(const float* r, const float* dinv, const float* mask, const float* y, int64_t nnode,
                                 int32_t batch, const float* scale, float* z, hf_stream_t stream) {
  HF_REQUIRE(r && dinv && mask && y && scale && z, "hf_lpc_output_f32: null argument");
  lpc_output<<<(unsigned)ceil_div(nnode * 3 * batch, 256), 256, 0, (hipStream_t)stream>>>(r, dinv, mask, y, nnode,
                                                                                            batch, scale, z);
  HF_LAUNCH_CHECK();
  return 0;
}

// ---- the solve ---------------------------------------------------------------
extern "C" size_t hf_lpc_workspace_size(int64_t nnode) {
  const int64_t N = nnode * 3;
  // scalars/partials header, FcgScalars + scale, vectors r z p0 p1 q, features, h1, h2
  return kHeader + 256 + 5 * align256(N * 4) + align256(nnode * 4 * 4) + 2 * align256(nnode * 16 * 4);
}

extern "C" int hf_lpc_solve_f32(const hf_pattern* p, const float* vals, const float* dinv, const float* b, float* x,
                                const float* const* weights, const float* gfield, const float* mask, int32_t nx,
                                int32_t ny, int32_t nz, int32_t maxiter, double rtol, double atol, int32_t poll,
                                int32_t flags, int32_t stall_window, void* ws, hf_krylov_info* info,
                                hf_stream_t stream) {
  HF_REQUIRE(p && vals && dinv && b && x && weights && gfield && mask && ws, "hf_lpc_solve_f32: null argument");
  HF_REQUIRE(pattern_vec(p) == 3, "hf_lpc_solve_f32: vec must be 3");
  const MatArgs m = mat_args(p);
  HF_REQUIRE((int64_t)nx * ny * nz == m.nrows, "hf_lpc_solve_f32: grid %dx%dx%d != %lld nodes", nx, ny, nz,
             (long long)m.nrows);
  hipStream_t s = (hipStream_t)stream;
  const int64_t nnode = m.nrows, N = nnode * 3;
  if (poll < 1) poll = 1;
  const int rel_r0 = (flags & HF_KRYLOV_REL_R0) ? 1 : 0;
  Scalars* sc = (Scalars*)ws;
  double* partials = (double*)((char*)ws + 4096);
  char* cur = (char*)ws + kHeader;
  FcgScalars* fs = (FcgScalars*)cur;
  float* scale = (float*)(cur + 128);
  cur += 256;
  auto take = [&](size_t bytes) { char* r_ = cur; cur += align256(bytes); return (float*)r_; };
  float *r = take(N * 4), *z = take(N * 4), *p0 = take(N * 4), *p1 = take(N * 4), *q = take(N * 4);
  float *f = take(nnode * 16), *h1 = take(nnode * 64), *h2 = take(nnode * 64);
  const CnnGrid g = mkgrid(nx, ny, nz, 1);
  const int gr = rows_grid(m.nslices), ge = elem_grid(N);
  const unsigned gn = (unsigned)ceil_div(nnode, 256);
  fcg_init_kernel<3><<<gr, kBlock, 0, s>>>(m, vals, dinv, b, x, r, z, p0, partials);
  fcg_init_scalars<<<1, kScalarBlock, 0, s>>>(sc, fs, scale, partials, gr, N, rtol * rtol, atol * atol, rel_r0,
                                             maxiter);
  HF_LAUNCH_CHECK();
  if (poll_init()) return -2;
  Scalars host{};
  int it = 0, batch = 0;
  for (;;) {
    for (int k = 0; k < poll; ++k, ++it) {
      float* pold = (it & 1) ? p1 : p0;
      float* pnew = (it & 1) ? p0 : p1;
      lpc_features<<<gn, 256, 0, s>>>(r, dinv, gfield, nnode, 1, scale, f, &sc->done);
      conv3d_fwd<4, 16, 1><<<gn, 256, 0, s>>>(f, weights[0], weights[1], g, h1, &sc->done);
      conv3d_fwd<16, 16, 1><<<gn, 256, 0, s>>>(h1, weights[2], weights[3], g, h2, &sc->done);
      fcg_conv3_out<16><<<gn, 256, 0, s>>>(h2, weights[4], weights[5], g, r, dinv, mask, fs, sc, z, partials);
      fcg_beta_scalars<<<1, kScalarBlock, 0, s>>>(sc, partials, (int)gn);
      pcg_spmv_kernel<float, 3, false><<<gr, kBlock, 0, s>>>(m, vals, z, pold, pnew, q, sc, partials);
      pcg_alpha_scalars<<<1, kScalarBlock, 0, s>>>(sc, partials, gr);
      fcg_update_kernel<<<ge, kBlock, 0, s>>>(N, dinv, pnew, q, x, r, sc, partials);
      fcg_end_scalars<<<1, kScalarBlock, 0, s>>>(sc, fs, scale, partials, ge, N, stall_window);
    }
    HF_LAUNCH_CHECK();
    if (poll_post(sc, batch & 1, s)) return -2;
    if (batch > 0) {
      if (poll_wait((batch - 1) & 1, host)) return -2;
      if (host.done) break;
    }
    ++batch;
    if (it >= maxiter + 2 * poll) break;
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
