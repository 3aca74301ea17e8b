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
// truncated flexible CG (Notay): each direction is A-orthogonalised against
// the last 1..4; the correction is scaled down where needed so that
// r.z >= 0.1 r.D^-1.r; when the recursive residual meets the tolerance the true
// residual is recomputed (up to 3 replacements).  Breakdown, a non-finite value
// or stagnation ends the flexible solve and the caller continues with
// Jacobi-PCG.
//
// Two implementations of the network: FP32 vector-ALU kernels (a thread per
// node, weights in LDS) used for training with their backward kernels, and
// the inference path of the solve on MFMA (f16 activations, f32
// accumulation), tested against the FP32 path.
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

// ---- MFMA inference path: f16 activations, f32 accumulation ----------------
// v_mfma_f32_16x16x16f16 computes D(16x16) += A(16x16) B(16x16); lane l holds
// A[l%16][4(l/16)+j], B[4(l/16)+j][l%16] and D[4(l/16)+j][l%16], j = 0..3.
// A conv layer is run as D[co][node] = sum_tap W_tap^T[co][ci] in_tap[ci][node]
// over tiles of 16 consecutive nodes, so a lane loads 4 channels of one
// neighbour node (8 bytes) and stores 4 output channels of one node: the
// channels-last f16 layout a layer writes is the one the next layer reads.
// Layer 1 (4 channels) packs 4 taps into one K = 16 step (7 steps, the 28th tap
// is zero); layer 3 pads its 3 output channels to 16.
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kPackW1 = 7 * 64, kPackW2 = 27 * 64, kPackW3 = 27 * 64;  // h4 entries

struct PackedNet {
  const h4 *w1, *w2, *w3;          // [step][lane] MFMA A operands
  const float *b1, *b2, *b3;
};

// Pack W1 (27x4x16), W2 (27x16x16), W3 (27x16x3) fp32 into per-lane f16 A operands.
__global__ void pack_weights(const float* __restrict__ w1, const float* __restrict__ w2,
                             const float* __restrict__ w3, h4* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kPackW1 + kPackW2 + kPackW3) return;
  const int lane = e & 63, co = lane & 15, grp = lane >> 4;
  h4 v;
  if (e < kPackW1) {
    const int tap = 4 * (e >> 6) + grp;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (_Float16)(tap < 27 ? w1[(tap * 4 + j) * 16 + co] : 0.f);
  } else if (e < kPackW1 + kPackW2) {
    const int tap = (e - kPackW1) >> 6;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (_Float16)w2[(tap * 16 + 4 * grp + j) * 16 + co];
  } else {
    const int tap = (e - kPackW1 - kPackW2) >> 6;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = (_Float16)(co < 3 ? w3[(tap * 16 + 4 * grp + j) * 3 + co] : 0.f);
  }
  out[e] = v;
}

// Tiles: a workgroup computes TI x TJ x TK output nodes (8 k-lines of 16, two per
// wavefront) from its (TI+2) x (TJ+2) x (TK+2) input halo staged in LDS, zero
// outside the grid, so the 27 taps of a lane are LDS reads without bounds checks.
// Each input node is read from global memory about 1.7 times instead of 27.
constexpr int TI = 2, TJ = 4, TK = 16;
constexpr int HI = TI + 2, HJ = TJ + 2, HK = TK + 2, HN = HI * HJ * HK;  // 432 halo nodes
constexpr int kTileGridMax = 2048;

struct TileGeom {
  int nbj, nbk;
  int64_t ntiles;
};

__host__ __device__ inline TileGeom tile_geom(const CnnGrid& g) {
  TileGeom t;
  const int nbi = (g.nx + TI - 1) / TI;
  t.nbj = (g.ny + TJ - 1) / TJ;
  t.nbk = (g.nz + TK - 1) / TK;
  t.ntiles = (int64_t)nbi * t.nbj * t.nbk;
  return t;
}

__device__ __forceinline__ void tile_origin(const TileGeom& tg, int64_t t, int& i0, int& j0, int& k0) {
  k0 = (int)(t % tg.nbk) * TK;
  j0 = (int)((t / tg.nbk) % tg.nbj) * TJ;
  i0 = (int)(t / ((int64_t)tg.nbk * tg.nbj)) * TI;
}

// halo node h -> global node index, or -1 outside the grid
__device__ __forceinline__ int64_t halo_node(const CnnGrid& g, int i0, int j0, int k0, int h) {
  const int hk = h % HK, hj = (h / HK) % HJ, hi = h / (HK * HJ);
  const int ii = i0 - 1 + hi, jj = j0 - 1 + hj, kk = k0 - 1 + hk;
  if (ii < 0 || ii >= g.nx || jj < 0 || jj >= g.ny || kk < 0 || kk >= g.nz) return -1;
  return ((int64_t)ii * g.ny + jj) * g.nz + kk;
}

// halo index of the tap neighbour of output node (li, lj, kk) of the tile
__device__ __forceinline__ int halo_tap(int li, int lj, int kk, int tap) {
  return ((li + 1 + tap / 9 - 1) * HJ + (lj + 1 + (tap / 3) % 3 - 1)) * HK + (kk + 1 + tap % 3 - 1);
}

__device__ __forceinline__ f4 bias4(const float* __restrict__ b, int grp, int nvalid) {
  f4 a;
#pragma unroll
  for (int j = 0; j < 4; ++j) a[j] = (4 * grp + j < nvalid) ? b[4 * grp + j] : 0.f;
  return a;
}

// Layer 1 (4 -> 16, tanh) with the input features formed while staging the halo:
// f = (r_i sqrt(dinv_i) / s, i < 3; gfield).  Four taps share one K = 16 step.
__global__ void __launch_bounds__(256) conv1_tile(const float* __restrict__ r, const float* __restrict__ dinv,
                                                  const float* __restrict__ gfield, const float* __restrict__ scale,
                                                  PackedNet net, CnnGrid g, h4* __restrict__ h1,
                                                  const int32_t* skip) {
  __shared__ h4 sf[HN];
  if (skip && *skip) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, kk = lane & 15, grp = lane >> 4;
  h4 wa[7];
#pragma unroll
  for (int t = 0; t < 7; ++t) wa[t] = net.w1[t * 64 + lane];
  const f4 b0 = bias4(net.b1, grp, 16);
  const float inv_s = 1.0f / scale[0];
  const TileGeom tg = tile_geom(g);
  const h4 zero = {0, 0, 0, 0};
  for (int64_t tile = blockIdx.x; tile < tg.ntiles; tile += gridDim.x) {
    int i0, j0, k0;
    tile_origin(tg, tile, i0, j0, k0);
    for (int h = threadIdx.x; h < HN; h += blockDim.x) {
      const int64_t m = halo_node(g, i0, j0, k0, h);
      h4 v = zero;
      if (m >= 0) {
#pragma unroll
        for (int i = 0; i < 3; ++i) v[i] = (_Float16)(r[m * 3 + i] * sqrtf(dinv[m * 3 + i]) * inv_s);
        v[3] = (_Float16)gfield[m];
      }
      sf[h] = v;
    }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int line = wv * 2 + x, li = line / TJ, lj = line % TJ;
      f4 acc = b0;
#pragma unroll
      for (int t = 0; t < 7; ++t) {
        const int tap = 4 * t + grp;
        const h4 bv = tap < 27 ? sf[halo_tap(li, lj, kk, tap)] : zero;
        acc = __builtin_amdgcn_mfma_f32_16x16x16f16(wa[t], bv, acc, 0, 0, 0);
      }
      const int ii = i0 + li, jj = j0 + lj, kn = k0 + kk;
      if (ii < g.nx && jj < g.ny && kn < g.nz) {
        const int64_t n = ((int64_t)ii * g.ny + jj) * g.nz + kn;
        h4 o;
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = (_Float16)tanhf(acc[j]);
        h1[n * 4 + grp] = o;
      }
    }
    __syncthreads();  // the halo is restaged for the next tile
  }
}

// Stage the 16-channel halo of a tile (4 h4 per node).
__device__ __forceinline__ void stage16(const h4* __restrict__ in, const CnnGrid& g, int i0, int j0, int k0,
                                        h4* __restrict__ sh) {
  const h4 zero = {0, 0, 0, 0};
  for (int e = threadIdx.x; e < HN * 4; e += blockDim.x) {
    const int64_t m = halo_node(g, i0, j0, k0, e >> 2);
    sh[e] = m >= 0 ? in[m * 4 + (e & 3)] : zero;
  }
}

// 27-tap 16-channel MFMA chain for output node (li, lj, kk) of the staged tile.
__device__ __forceinline__ f4 conv16_lds(const h4 (&wa)[27], const h4* __restrict__ sh, int li, int lj, int kk,
                                         int grp, f4 acc) {
  h4 bv[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) bv[t] = sh[halo_tap(li, lj, kk, t) * 4 + grp];
#pragma unroll
  for (int t = 0; t < 27; ++t) acc = __builtin_amdgcn_mfma_f32_16x16x16f16(wa[t], bv[t], acc, 0, 0, 0);
  return acc;
}

// Layer 2 (16 -> 16, tanh)
__global__ void __launch_bounds__(256) conv2_tile(const h4* __restrict__ h1, PackedNet net, CnnGrid g,
                                                  h4* __restrict__ h2, const int32_t* skip) {
  __shared__ h4 sh[HN * 4];
  if (skip && *skip) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, kk = lane & 15, grp = lane >> 4;
  h4 wa[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) wa[t] = net.w2[t * 64 + lane];
  const f4 b0 = bias4(net.b2, grp, 16);
  const TileGeom tg = tile_geom(g);
  for (int64_t tile = blockIdx.x; tile < tg.ntiles; tile += gridDim.x) {
    int i0, j0, k0;
    tile_origin(tg, tile, i0, j0, k0);
    stage16(h1, g, i0, j0, k0, sh);
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int line = wv * 2 + x, li = line / TJ, lj = line % TJ;
      const f4 acc = conv16_lds(wa, sh, li, lj, kk, grp, b0);
      const int ii = i0 + li, jj = j0 + lj, kn = k0 + kk;
      if (ii < g.nx && jj < g.ny && kn < g.nz) {
        const int64_t n = ((int64_t)ii * g.ny + jj) * g.nz + kn;
        h4 o;
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = (_Float16)tanhf(acc[j]);
        h2[n * 4 + grp] = o;
      }
    }
    __syncthreads();
  }
}

// Layer 3 (16 -> 3) and z = D^-1 r + s D^-1/2 mask y (the preconditioner applied once)
__global__ void __launch_bounds__(256) conv3_apply_tile(const h4* __restrict__ h2, PackedNet net, CnnGrid g,
                                                        const float* __restrict__ r, const float* __restrict__ dinv,
                                                        const float* __restrict__ mask,
                                                        const float* __restrict__ scale, float* __restrict__ z) {
  __shared__ h4 sh[HN * 4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, kk = lane & 15, grp = lane >> 4;
  h4 wa[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) wa[t] = net.w3[t * 64 + lane];
  const f4 b0 = bias4(net.b3, grp, 3);
  const float s = scale[0];
  const TileGeom tg = tile_geom(g);
  for (int64_t tile = blockIdx.x; tile < tg.ntiles; tile += gridDim.x) {
    int i0, j0, k0;
    tile_origin(tg, tile, i0, j0, k0);
    stage16(h2, g, i0, j0, k0, sh);
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int line = wv * 2 + x, li = line / TJ, lj = line % TJ;
      const f4 acc = conv16_lds(wa, sh, li, lj, kk, grp, b0);
      const int ii = i0 + li, jj = j0 + lj, kn = k0 + kk;
      if (grp == 0 && ii < g.nx && jj < g.ny && kn < g.nz) {
        const int64_t n = ((int64_t)ii * g.ny + jj) * g.nz + kn;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int64_t d = n * 3 + i;
          z[d] = dinv[d] * r[d] + s * sqrtf(dinv[d]) * mask[d] * acc[i];
        }
      }
    }
    __syncthreads();
  }
}

// Layer 3 (16 -> 3) as a smoothing step: e += D^-1 r + s D^-1/2 mask y
__global__ void __launch_bounds__(256) conv3_accum_tile(const h4* __restrict__ h2, PackedNet net, CnnGrid g,
                                                        const float* __restrict__ r, const float* __restrict__ dinv,
                                                        const float* __restrict__ mask,
                                                        const float* __restrict__ scale, float* __restrict__ e,
                                                        const int32_t* skip) {
  __shared__ h4 sh[HN * 4];
  if (skip && *skip) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, kk = lane & 15, grp = lane >> 4;
  h4 wa[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) wa[t] = net.w3[t * 64 + lane];
  const f4 b0 = bias4(net.b3, grp, 3);
  const float s = scale[0];
  const TileGeom tg = tile_geom(g);
  for (int64_t tile = blockIdx.x; tile < tg.ntiles; tile += gridDim.x) {
    int i0, j0, k0;
    tile_origin(tg, tile, i0, j0, k0);
    stage16(h2, g, i0, j0, k0, sh);
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int line = wv * 2 + x, li = line / TJ, lj = line % TJ;
      const f4 acc = conv16_lds(wa, sh, li, lj, kk, grp, b0);
      const int ii = i0 + li, jj = j0 + lj, kn = k0 + kk;
      if (grp == 0 && ii < g.nx && jj < g.ny && kn < g.nz) {
        const int64_t n = ((int64_t)ii * g.ny + jj) * g.nz + kn;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int64_t d = n * 3 + i;
          e[d] += dinv[d] * r[d] + s * sqrtf(dinv[d]) * mask[d] * acc[i];
        }
      }
    }
    __syncthreads();
  }
}

// scale = rms(D^-1/2 r): partials then one workgroup
__global__ void __launch_bounds__(kBlock) rdr_partials(const float* __restrict__ r, const float* __restrict__ dinv,
                                                       int64_t N, double* partials, const int32_t* skip) {
  if (skip && *skip) return;
  double loc[1] = {0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x)
    loc[0] += (double)r[i] * (double)r[i] * (double)dinv[i];
  block_partials<1>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) rms_scalar(const double* partials, int nparts, int64_t N,
                                                           float* scale, const int32_t* skip) {
  if (skip && *skip) return;
  double tot[1];
  sum_partials<1>(partials, nparts, tot);
  if (threadIdx.x == 0) scale[0] = (float)sqrt(fmax(tot[0], 1e-300) / (double)N);
}

static inline unsigned tile_grid(const CnnGrid& g) {
  const int64_t nt = tile_geom(g).ntiles;
  return (unsigned)(nt < kTileGridMax ? nt : kTileGridMax);
}

// ---- flexible CG with the learned preconditioner ----------------------------
struct LpcArgs {
  const float *w1, *b1, *w2, *b2, *w3, *b3;
  const float* gfield;  // nnode
  const float* mask;    // N (1 interior, 0 Dirichlet)
  CnnGrid grid;         // batch = 1
};

constexpr int kMaxTrunc = 4;  // directions kept by the truncated flexible CG

struct FcgScalars {
  float s;       // rms of D^-1/2 r
  int32_t stall;
  double best_rr;
  double rdr;    // r.D^-1.r of the current residual
  double theta;  // weight of the network correction in this step's z
  double pq[kMaxTrunc];    // p_j.A p_j of the kept directions (ring slots)
  double beta[kMaxTrunc];  // this step's A-orthogonalisation coefficients
  int32_t base;            // iteration of the last (re)start: directions before it are dropped
  int32_t improved;        // ||r|| has dropped below 0.9 ||r_start|| since the last (re)start
};

// z = D^-1 r + theta c keeps r.z >= kPosFloor r.D^-1.r (theta in [0, 1]).
constexpr double kPosFloor = 0.1;

struct Ring {
  float* v[kMaxTrunc];
};

// init: r = b - A x; partials r.r, b.b, r.D^-1.r
__global__ void __launch_bounds__(kBlock) fcg_init_kernel(MatArgs m, const float* __restrict__ vals,
                                                          const float* __restrict__ dinv, const float* __restrict__ b,
                                                          const float* __restrict__ x, float* __restrict__ r,
                                                          double* partials) {
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return x[c]; };
  double loc[3] = {0, 0, 0};
  auto epi = [&](int64_t n, int i, float ax) {
    const int64_t d = n * 3 + i;
    const float rr = b[d] - ax;
    r[d] = rr;
    loc[0] += (double)rr * (double)rr;
    loc[1] += (double)b[d] * (double)b[d];
    loc[2] += (double)rr * (double)rr * (double)dinv[d];
  };
  sweep<float, 3, false>(m, r0, r1, vals, xv, epi);
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
  sc->rho = 0;
  sc->alpha = 0;
  sc->iter = 0;
  sc->status = 0;
  sc->maxiter = maxiter;
  sc->done = (tot[0] <= sc->thresh) ? 1 : 0;
  if (!isfinite(tot[0])) { sc->done = 1; sc->status = 3; }
  else if (maxiter <= 0 && !sc->done) { sc->done = 1; sc->status = 1; }
  fs->s = (float)sqrt(fmax(tot[2], 1e-300) / (double)N);
  fs->stall = 0;
  fs->improved = 0;
  fs->best_rr = tot[0];
  fs->rdr = tot[2];
  fs->theta = 1.0;
  for (int j = 0; j < kMaxTrunc; ++j) { fs->pq[j] = 1.0; fs->beta[j] = 0.0; }
  fs->base = 0;
  *scale = fs->s;
}

// Residual replacement after the recursive residual met the tolerance:
// partials of r = b - A x (recomputed by fcg_init_kernel).  Accept, or restart
// the flexible CG from the true residual with an empty direction ring.
__global__ void __launch_bounds__(kScalarBlock) fcg_verify_scalars(Scalars* sc, FcgScalars* fs, float* scale,
                                                                   const double* partials, int nparts, int64_t N) {
  double tot[3];
  sum_partials<3>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  sc->rr = tot[0];
  if (!isfinite(tot[0])) { sc->done = 1; sc->status = 3; return; }
  if (tot[0] <= sc->thresh) { sc->done = 1; sc->status = 0; return; }
  if (sc->iter >= sc->maxiter) { sc->done = 1; sc->status = 1; return; }
  sc->done = 0;
  sc->status = 0;
  fs->base = sc->iter;
  fs->rdr = tot[2];
  fs->s = (float)sqrt(fmax(tot[2], 1e-300) / (double)N);
  fs->stall = 0;
  fs->improved = 0;
  fs->best_rr = tot[0];
  *scale = fs->s;
}

// conv3 (MFMA, LDS-tiled) fused with the output scaling: c = s sqrt(dinv) mask y
// (the network's correction to the Jacobi step).  Partials: r.c, then per kept
// direction j (D^-1 r).q_j and c.q_j.  The grid loops over tiles (at most
// kTileGridMax workgroups), so the partials stay bounded.
template <int M>
__global__ void __launch_bounds__(256) fcg_conv3_out(const h4* __restrict__ h2, PackedNet net, CnnGrid g,
                                                     const float* __restrict__ r, const float* __restrict__ dinv,
                                                     const float* __restrict__ mask, const FcgScalars* fs,
                                                     const Scalars* sc, Ring Q, float* __restrict__ c,
                                                     double* partials) {
  __shared__ h4 sh[HN * 4];
  if (sc->done) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, kk = lane & 15, grp = lane >> 4;
  h4 wa[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) wa[t] = net.w3[t * 64 + lane];
  const f4 b0 = bias4(net.b3, grp, 3);
  const int k = sc->iter - fs->base;
  const int nq = k < M ? k : M;
  const float s = fs->s;
  double loc[1 + 2 * M];
#pragma unroll
  for (int j = 0; j < 1 + 2 * M; ++j) loc[j] = 0.0;
  const TileGeom tg = tile_geom(g);
  for (int64_t tile = blockIdx.x; tile < tg.ntiles; tile += gridDim.x) {
    int i0, j0, k0;
    tile_origin(tg, tile, i0, j0, k0);
    stage16(h2, g, i0, j0, k0, sh);
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int line = wv * 2 + x, li = line / TJ, lj = line % TJ;
      const f4 acc = conv16_lds(wa, sh, li, lj, kk, grp, b0);
      const int ii = i0 + li, jj = j0 + lj, kn = k0 + kk;
      if (grp == 0 && ii < g.nx && jj < g.ny && kn < g.nz) {
        const int64_t n = ((int64_t)ii * g.ny + jj) * g.nz + kn;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int64_t d = n * 3 + i;
          const float cn = s * sqrtf(dinv[d]) * mask[d] * acc[i];
          const float jr = dinv[d] * r[d];
          c[d] = cn;
          loc[0] += (double)r[d] * (double)cn;
#pragma unroll
          for (int j = 0; j < M; ++j) {
            if (j < nq) {
              const double qj = Q.v[j][d];
              loc[1 + 2 * j] += (double)jr * qj;
              loc[2 + 2 * j] += (double)cn * qj;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  block_partials<1 + 2 * M>(loc, partials);
}

// theta (positivity safeguard), rho = r.z, beta_j = -z.q_j / p_j.q_j
template <int M>
__global__ void __launch_bounds__(kScalarBlock) fcg_beta_scalars(Scalars* sc, FcgScalars* fs, const double* partials,
                                                                 int nparts) {
  if (sc->done) return;
  double tot[1 + 2 * M];
  sum_partials<1 + 2 * M>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  const double a = fs->rdr, rc = tot[0];
  bool finite = isfinite(rc) && isfinite(a);
  for (int j = 0; j < 2 * M; ++j) finite = finite && isfinite(tot[1 + j]);
  if (!finite) {
    sc->done = 1;
    sc->status = 3;
    return;
  }
  double theta = 1.0;
  if (a + rc < kPosFloor * a) theta = (kPosFloor - 1.0) * a / rc;  // rc < 0 here, theta in [0, 1)
  const double rz = a + theta * rc;
  if (!(rz > 0.0)) {  // r = 0 on the free rows: nothing left to do
    sc->done = 1;
    sc->status = 2;
    return;
  }
  fs->theta = theta;
  const int k = sc->iter - fs->base;
  const int nq = k < M ? k : M;
  for (int j = 0; j < M; ++j)
    fs->beta[j] = j < nq ? -(tot[1 + 2 * j] + theta * tot[2 + 2 * j]) / fs->pq[j] : 0.0;
  sc->rho = rz;
}

// p = D^-1 r + theta c + sum_j beta_j p_j into the ring slot iter % M
template <int M>
__global__ void __launch_bounds__(kBlock) fcg_form_p(int64_t N, const float* __restrict__ r,
                                                     const float* __restrict__ dinv, const float* __restrict__ c,
                                                     Ring P, const Scalars* sc, const FcgScalars* fs) {
  if (sc->done) return;
  const float theta = (float)fs->theta;
  float beta[M];
#pragma unroll
  for (int j = 0; j < M; ++j) beta[j] = (float)fs->beta[j];
  const int k = sc->iter - fs->base;
  const int nq = k < M ? k : M;
  float* __restrict__ dst = P.v[k % M];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
    float p = dinv[i] * r[i] + theta * c[i];
#pragma unroll
    for (int j = 0; j < M; ++j)
      if (j < nq) p += beta[j] * P.v[j][i];
    dst[i] = p;
  }
}

// q = A p (slot iter % M); partials p.q, r.p
template <int M>
__global__ void __launch_bounds__(kBlock) fcg_spmv_kernel(MatArgs m, const float* __restrict__ vals,
                                                          const float* __restrict__ r, Ring P, Ring Q,
                                                          const Scalars* sc, const FcgScalars* fs,
                                                          double* partials) {
  if (sc->done) return;
  const int slot = (sc->iter - fs->base) % M;
  const float* __restrict__ p = P.v[slot];
  float* __restrict__ q = Q.v[slot];
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t j) { return p[j]; };
  double loc[2] = {0, 0};
  auto epi = [&](int64_t n, int i, float ap) {
    const int64_t d = n * 3 + i;
    q[d] = ap;
    loc[0] += (double)p[d] * (double)ap;
    loc[1] += (double)r[d] * (double)p[d];
  };
  sweep<float, 3, false>(m, r0, r1, vals, xv, epi);
  block_partials<2>(loc, partials);
}

template <int M>
__global__ void __launch_bounds__(kScalarBlock) fcg_alpha_scalars(Scalars* sc, FcgScalars* fs, const double* partials,
                                                                  int nparts) {
  if (sc->done) return;
  double tot[2];
  sum_partials<2>(partials, nparts, tot);
  if (threadIdx.x != 0) return;
  sc->pq = tot[0];
  if (!(tot[0] > 0.0) || !isfinite(tot[0]) || !isfinite(tot[1])) {
    sc->done = 1;
    sc->status = isfinite(tot[0]) && isfinite(tot[1]) ? 2 : 3;
  } else {
    sc->alpha = tot[1] / tot[0];
    fs->pq[(sc->iter - fs->base) % M] = tot[0];
  }
}

// x += alpha p, r -= alpha q; partials r.r, r.D^-1.r
template <int M>
__global__ void __launch_bounds__(kBlock) fcg_update_kernel(int64_t N, const float* __restrict__ dinv, Ring P,
                                                            Ring Q, float* __restrict__ x, float* __restrict__ r,
                                                            const Scalars* sc, const FcgScalars* fs,
                                                            double* partials) {
  if (sc->done) return;
  const float alpha = (float)sc->alpha;
  const int slot = (sc->iter - fs->base) % M;
  const float* __restrict__ p = P.v[slot];
  const float* __restrict__ q = Q.v[slot];
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
  fs->rdr = tot[1];
  fs->s = (float)sqrt(fmax(tot[1], 1e-300) / (double)N);
  *scale = fs->s;
  if (tot[0] < 0.81 * fs->best_rr) {  // progress: ||r|| down by 10 %
    fs->best_rr = tot[0];
    fs->stall = 0;
    fs->improved = 1;
  } else {
    fs->stall += 1;
  }
  if (!isfinite(tot[0])) { sc->done = 1; sc->status = 3; }
  else if (tot[0] <= sc->thresh) { sc->done = 1; sc->status = 0; }
  // stagnation: once ||r|| has dropped 10 % below its start, no new best for
  // max(stall_window, iter / 2) iterations; before that, 10 stall windows.  The
  // residual of CG-type methods can grow for a long stretch while a localised load
  // spreads through the body.
  else if (fs->improved ? fs->stall >= max(stall_window, (sc->iter - fs->base) / 2)
                        : (int64_t)(sc->iter - fs->base) >= (int64_t)10 * stall_window) {
    sc->done = 1;
    sc->status = 4;
  }
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
template <int M>
static void fcg_batch(int poll, const MatArgs& m, const float* vals, const float* dinv, float* x, float* r, float* c,
                      const Ring& P, const Ring& Q, h4* f, h4* h1, h4* h2, const PackedNet& net,
                      const float* gfield, const float* mask, const CnnGrid& g, Scalars* sc, FcgScalars* fs,
                      float* scale, double* partials, int stall_window, hipStream_t s) {
  const int64_t nnode = m.nrows, N = nnode * 3;
  const int gr = rows_grid(m.nslices), ge = elem_grid(N);
  const unsigned gt = tile_grid(g);
  (void)f;
  for (int k = 0; k < poll; ++k) {
    conv1_tile<<<gt, 256, 0, s>>>(r, dinv, gfield, scale, net, g, h1, &sc->done);
    conv2_tile<<<gt, 256, 0, s>>>(h1, net, g, h2, &sc->done);
    fcg_conv3_out<M><<<gt, 256, 0, s>>>(h2, net, g, r, dinv, mask, fs, sc, Q, c, partials);
    fcg_beta_scalars<M><<<1, kScalarBlock, 0, s>>>(sc, fs, partials, (int)gt);
    fcg_form_p<M><<<ge, kBlock, 0, s>>>(N, r, dinv, c, P, sc, fs);
    fcg_spmv_kernel<M><<<gr, kBlock, 0, s>>>(m, vals, r, P, Q, sc, fs, partials);
    fcg_alpha_scalars<M><<<1, kScalarBlock, 0, s>>>(sc, fs, partials, gr);
    fcg_update_kernel<M><<<ge, kBlock, 0, s>>>(N, dinv, P, Q, x, r, sc, fs, partials);
    fcg_end_scalars<<<1, kScalarBlock, 0, s>>>(sc, fs, scale, partials, ge, N, stall_window);
  }
}

static PackedNet pack_net(const float* const* weights, h4* packed, hipStream_t s) {
  const int tot = kPackW1 + kPackW2 + kPackW3;
  pack_weights<<<(tot + 255) / 256, 256, 0, s>>>(weights[0], weights[2], weights[4], packed);
  PackedNet net;
  net.w1 = packed;
  net.w2 = packed + kPackW1;
  net.w3 = packed + kPackW1 + kPackW2;
  net.b1 = weights[1];
  net.b2 = weights[3];
  net.b3 = weights[5];
  return net;
}

static inline size_t net_bytes(int64_t nnode) {
  // packed weights, features (h4 per node), h1 and h2 (4 h4 per node)
  return align256((size_t)(kPackW1 + kPackW2 + kPackW3) * sizeof(h4)) + align256(nnode * 8) +
         2 * align256(nnode * 32);
}

namespace hf {
// ---- the network as a multigrid fine-level smoother (mg.hip) -----------------
// buffer: packed weights | biases (48 floats) | scale | partials | h1 | h2
namespace {
constexpr size_t kSmPacked = (size_t)(kPackW1 + kPackW2 + kPackW3) * sizeof(h4);
inline size_t sm_off_bias() { return align256(kSmPacked); }
inline size_t sm_off_scale() { return sm_off_bias() + 256; }
inline size_t sm_off_part() { return sm_off_scale() + 256; }
inline size_t sm_off_h1() { return sm_off_part() + align256((size_t)kMaxGrid * sizeof(double)); }
__global__ void copy_bias(const float* __restrict__ b1, const float* __restrict__ b2, const float* __restrict__ b3,
                          float* __restrict__ out) {
  const int i = threadIdx.x;
  if (i < 16) out[i] = b1[i];
  else if (i < 32) out[i] = b2[i - 16];
  else if (i < 35) out[i] = b3[i - 32];
  else if (i < 48) out[i] = 0.f;
}
}  // namespace

size_t lpc_smoother_bytes(int64_t nnode) { return sm_off_h1() + 2 * align256((size_t)nnode * 32); }

int lpc_smoother_load(const LpcSmoother& L, const float* const* weights, hipStream_t s) {
  for (int i = 0; i < 6; ++i) HF_REQUIRE(weights[i], "learned smoother: null weight %d", i);
  char* buf = (char*)L.buf;
  h4* packed = (h4*)buf;
  const int tot = kPackW1 + kPackW2 + kPackW3;
  pack_weights<<<(tot + 255) / 256, 256, 0, s>>>(weights[0], weights[2], weights[4], packed);
  copy_bias<<<1, 64, 0, s>>>(weights[1], weights[3], weights[5], (float*)(buf + sm_off_bias()));
  HF_LAUNCH_CHECK();
  return 0;
}

int lpc_smoother_apply(const LpcSmoother& L, const float* res, const float* dinv, const float* mask, float* e,
                       const int32_t* skip, hipStream_t s) {
  char* buf = (char*)L.buf;
  const h4* packed = (const h4*)buf;
  const float* bias = (const float*)(buf + sm_off_bias());
  PackedNet net;
  net.w1 = packed;
  net.w2 = packed + kPackW1;
  net.w3 = packed + kPackW1 + kPackW2;
  net.b1 = bias;
  net.b2 = bias + 16;
  net.b3 = bias + 32;
  float* scale = (float*)(buf + sm_off_scale());
  double* part = (double*)(buf + sm_off_part());
  const int64_t nnode = (int64_t)L.nx * L.ny * L.nz, N = 3 * nnode;
  h4* h1 = (h4*)(buf + sm_off_h1());
  h4* h2 = (h4*)(buf + sm_off_h1() + align256((size_t)nnode * 32));
  const CnnGrid g = mkgrid(L.nx, L.ny, L.nz, 1);
  const int ge = elem_grid(N);
  rdr_partials<<<ge, kBlock, 0, s>>>(res, dinv, N, part, skip);
  rms_scalar<<<1, kScalarBlock, 0, s>>>(part, ge, N, scale, skip);
  const unsigned gt = tile_grid(g);
  conv1_tile<<<gt, 256, 0, s>>>(res, dinv, L.gfield, scale, net, g, h1, skip);
  conv2_tile<<<gt, 256, 0, s>>>(h1, net, g, h2, skip);
  conv3_accum_tile<<<gt, 256, 0, s>>>(h2, net, g, res, dinv, mask, scale, e, skip);
  HF_LAUNCH_CHECK();
  return 0;
}
}  // namespace hf

extern "C" size_t hf_lpc_apply_workspace_size(int64_t nnode) { return kHeader + 256 + net_bytes(nnode); }

extern "C" int hf_lpc_apply_f32(const float* r, const float* dinv, const float* gfield, const float* mask,
                                const float* const* weights, int32_t nx, int32_t ny, int32_t nz, void* ws, float* z,
                                hf_stream_t stream) {
  HF_REQUIRE(r && dinv && gfield && mask && weights && ws && z, "hf_lpc_apply_f32: null argument");
  HF_REQUIRE(nx > 0 && ny > 0 && nz > 0, "hf_lpc_apply_f32: empty grid");
  for (int i = 0; i < 6; ++i) HF_REQUIRE(weights[i], "hf_lpc_apply_f32: null weight %d", i);
  hipStream_t s = (hipStream_t)stream;
  const int64_t nnode = (int64_t)nx * ny * nz, N = nnode * 3;
  double* partials = (double*)((char*)ws + 4096);
  char* cur = (char*)ws + kHeader;
  float* scale = (float*)cur;
  cur += 256;
  h4* packed = (h4*)cur;
  cur += align256((size_t)(kPackW1 + kPackW2 + kPackW3) * sizeof(h4));
  h4* f = (h4*)cur;
  cur += align256(nnode * 8);
  h4* h1 = (h4*)cur;
  cur += align256(nnode * 32);
  h4* h2 = (h4*)cur;
  const CnnGrid g = mkgrid(nx, ny, nz, 1);
  const PackedNet net = pack_net(weights, packed, s);
  const int ge = elem_grid(N);
  rdr_partials<<<ge, kBlock, 0, s>>>(r, dinv, N, partials, nullptr);
  rms_scalar<<<1, kScalarBlock, 0, s>>>(partials, ge, N, scale, nullptr);
  const unsigned gt = tile_grid(g);
  (void)f;
  conv1_tile<<<gt, 256, 0, s>>>(r, dinv, gfield, scale, net, g, h1, nullptr);
  conv2_tile<<<gt, 256, 0, s>>>(h1, net, g, h2, nullptr);
  conv3_apply_tile<<<gt, 256, 0, s>>>(h2, net, g, r, dinv, mask, scale, z);
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" size_t hf_lpc_workspace_size(int64_t nnode) {
  const int64_t N = nnode * 3;
  // scalars/partials header, FcgScalars + scale, vectors r c P[4] Q[4], network buffers
  return kHeader + 256 + (2 + 2 * kMaxTrunc) * align256(N * 4) + net_bytes(nnode);
}

extern "C" int hf_lpc_solve_f32(const hf_pattern* p, const float* vals, const float* dinv, const float* b, float* x,
                                const float* const* weights, const float* gfield, const float* mask, int32_t nx,
                                int32_t ny, int32_t nz, int32_t maxiter, double rtol, double atol, int32_t poll,
                                int32_t flags, int32_t stall_window, int32_t trunc, void* ws, hf_krylov_info* info,
                                hf_stream_t stream) {
  HF_REQUIRE(p && vals && dinv && b && x && weights && gfield && mask && ws, "hf_lpc_solve_f32: null argument");
  HF_REQUIRE(pattern_vec(p) == 3, "hf_lpc_solve_f32: vec must be 3");
  HF_REQUIRE(trunc >= 1 && trunc <= kMaxTrunc, "hf_lpc_solve_f32: trunc %d outside 1..%d", trunc, kMaxTrunc);
  const MatArgs m = mat_args(p);
  HF_REQUIRE((int64_t)nx * ny * nz == m.nrows, "hf_lpc_solve_f32: grid %dx%dx%d != %lld nodes", nx, ny, nz,
             (long long)m.nrows);
  static_assert(kTileGridMax * (1 + 2 * kMaxTrunc) <= kMaxGrid * 4, "fcg_conv3_out partials exceed the header");
  hipStream_t s = (hipStream_t)stream;
  const int64_t nnode = m.nrows, N = nnode * 3;
  if (poll < 1) poll = 1;
  const int rel_r0 = (flags & HF_KRYLOV_REL_R0) ? 1 : 0;
  Scalars* sc = (Scalars*)ws;
  double* partials = (double*)((char*)ws + 4096);
  char* cur = (char*)ws + kHeader;
  FcgScalars* fs = (FcgScalars*)cur;
  static_assert(sizeof(FcgScalars) <= 128, "FcgScalars overlaps the scale slot");
  float* scale = (float*)(cur + 128);
  cur += 256;
  auto take = [&](size_t bytes) { char* r_ = cur; cur += align256(bytes); return (float*)r_; };
  float *r = take(N * 4), *c = take(N * 4);
  Ring P, Q;
  for (int j = 0; j < kMaxTrunc; ++j) P.v[j] = take(N * 4);
  for (int j = 0; j < kMaxTrunc; ++j) Q.v[j] = take(N * 4);
  h4* packed = (h4*)take((size_t)(kPackW1 + kPackW2 + kPackW3) * sizeof(h4));
  h4* f = (h4*)take(nnode * 8);
  h4* h1 = (h4*)take(nnode * 32);
  h4* h2 = (h4*)take(nnode * 32);
  const CnnGrid g = mkgrid(nx, ny, nz, 1);
  const PackedNet net = pack_net(weights, packed, s);
  const int gr = rows_grid(m.nslices);
  fcg_init_kernel<<<gr, kBlock, 0, s>>>(m, vals, dinv, b, x, r, partials);
  fcg_init_scalars<<<1, kScalarBlock, 0, s>>>(sc, fs, scale, partials, gr, N, rtol * rtol, atol * atol, rel_r0,
                                             maxiter);
  HF_LAUNCH_CHECK();
  if (poll_init()) return -2;
  Scalars host{};
  constexpr int kMaxReplace = 3;
  for (int replaced = 0;; ++replaced) {
    int it = 0, batch = 0;
    for (;;) {
      switch (trunc) {
#define X(MM) \
  case MM: fcg_batch<MM>(poll, m, vals, dinv, x, r, c, P, Q, f, h1, h2, net, gfield, mask, g, sc, fs, scale, \
                         partials, stall_window, s); break;
        X(1) X(2) X(3) X(4)
#undef X
      }
      it += poll;
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
    if (!host.done || host.status != 0 || replaced == kMaxReplace) break;
    // the recursive residual met the tolerance: check the true one
    fcg_init_kernel<<<gr, kBlock, 0, s>>>(m, vals, dinv, b, x, r, partials);
    fcg_verify_scalars<<<1, kScalarBlock, 0, s>>>(sc, fs, scale, partials, gr, N);
    HF_LAUNCH_CHECK();
    HF_CHECK_HIP(hipMemcpyAsync(&host, sc, sizeof(Scalars), hipMemcpyDeviceToHost, s));
    HF_CHECK_HIP(hipStreamSynchronize(s));
    if (host.done) break;
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
