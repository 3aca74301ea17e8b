// Device-side pieces shared by the Krylov drivers (krylov.hip, lpc.hip):
// scalars, fence-free partial reductions, XCD-aware ranges, the sliced-ELL
// slice sweep, workspace layout and the host-side poll / profiling hooks.
#pragma once
#include "common.hpp"
#include "hipfem.h"

namespace hf {
const int32_t* pattern_rowptr(const hf_pattern* p);
const int32_t* pattern_colidx(const hf_pattern* p);
const uint8_t* pattern_cslot(const hf_pattern* p);
const int32_t* pattern_trans(const hf_pattern* p);
int64_t pattern_nnode(const hf_pattern* p);
int64_t pattern_nblocks(const hf_pattern* p);
int pattern_max_row_blocks(const hf_pattern* p);
int64_t pattern_nslices(const hf_pattern* p);
const int32_t* pattern_slice_ptr(const hf_pattern* p);
const int32_t* pattern_sell_col(const hf_pattern* p);
const int32_t* pattern_sell_tpos(const hf_pattern* p);
const int32_t* pattern_sell_row(const hf_pattern* p);
const int32_t* pattern_seg(const hf_pattern* p);
const int32_t* pattern_contrib(const hf_pattern* p);
const int32_t* pattern_diag(const hf_pattern* p);
int pattern_R(const hf_pattern* p);
int pattern_nn(const hf_pattern* p);
int pattern_vec(const hf_pattern* p);

constexpr int kBlock = 256;         // threads per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxGrid = 8192;      // workgroups of the reduction kernels
constexpr int kNumXcd = 8;

struct Scalars {
  double rho, rho_, alpha, omega, rr, bb, thresh, pq, ss, ts, tt;
  int32_t iter, done, status, exit_early;
  int32_t maxiter, pad;
  double pqs[2];  // the multigrid FCG's p.q, by iteration parity
};

// Logical workgroup id, remapped so that each XCD owns a contiguous range.
__device__ __forceinline__ int logical_block() {
  const int g = gridDim.x, b = blockIdx.x;
  if (g % kNumXcd != 0) return b;
  return (b % kNumXcd) * (g / kNumXcd) + b / kNumXcd;
}

// Block reduction of NV doubles; thread 0 writes this workgroup's partials.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* __restrict__ partials) {
  __shared__ double sm[kWavesPerBlock][NV];
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
  }
}

// Sum of nparts partials in a fixed order, for a one-workgroup kernel; every
// thread receives the totals.
constexpr int kScalarBlock = 1024;   // threads of the one-workgroup scalar kernels
constexpr int kScalarWaves = kScalarBlock / kWave;

template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ partials, int nparts, double (&tot)[NV]) {
  __shared__ double sm[kScalarWaves][NV];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  double acc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = 0;
  for (int b = threadIdx.x; b < nparts; b += kScalarBlock)
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
    for (int k = 0; k < kScalarWaves; ++k) s += sm[k][j];
    tot[j] = s;
  }
}

// ---------------------------------------------------------------------------
// Slice sweeps over the sliced-ELL matrix (layout in assembly.hip): one
// wavefront per slice of R node rows, lane = r*VEC + i owns scalar row
// VEC*(slice*R + r) + i.  For every slot t of the slice the lane reads the
// block column c = sell_col[t*R + r] and the VEC values of its row of the
// block (one VEC-word load; the wavefront reads L*VEC consecutive values),
// gathers x at c, and accumulates.  Slots are independent, so the unrolled
// loop keeps several slots' loads in flight; no cross-lane reduction.
// TRANS (A^T x): the block of A^T in slot t is the transpose of the block
// whose value base is sell_tpos[t*R + r]; entry (i, jj) = vals[base + jj*VEC + i].
struct MatArgs {
  int64_t nrows, nslices;
  const int32_t* slice_ptr;
  const int32_t* sell_col;
  const int32_t* sell_tpos;
  const int32_t* sell_row;  // node of each slice row (-1: padding)
};

template <int VEC>
struct SellShape {
  static constexpr int R = kWave / VEC;   // node rows per slice: 64, 32, 21
  static constexpr int L = R * VEC;
  static constexpr int BB = VEC * VEC;
};

// LDS written by the lanes of one wavefront and read back by other lanes of it.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// y = A x (or A^T x) over the slices [s0, s1), one wavefront per slice; epi(n, i, y)
// receives scalar row (node n, component i).  xv(k) returns entry k of x.
// VEC = 3 (column form): lane (r, i), lane = 3r + i < 63, streams column i of
// its row's block in each slot (3 contiguous values), gathers the single entry
// x[col*3 + i] and keeps 3 partial outputs; the 3 lanes of a row are summed
// through LDS once per slice.  Each entry of x is gathered once per block,
// not once per lane of the row.
// TRANS: the block of A^T at (n, c) is the transpose of the stored block at
// sell_tpos (block (c, n), column-major), so lane (r, i) reads row i of it.
template <typename T, int VEC, bool TRANS, typename XF, typename EPI>
__device__ __forceinline__ void sweep(const MatArgs& m, int64_t s0, int64_t s1, const T* __restrict__ vals, XF xv,
                                      EPI epi) {
  using S = SellShape<VEC>;
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int r = lane / VEC, i = lane - (lane / VEC) * VEC;
  const bool in_slice = lane < S::L;
  if constexpr (VEC == 1) {
    for (int64_t s = s0 + w; s < s1; s += kWavesPerBlock) {
      const int32_t t0 = m.slice_ptr[s], t1 = m.slice_ptr[s + 1];
      T y = 0;
      if (in_slice) {
#pragma unroll 8
        for (int32_t t = t0; t < t1; ++t) {
          const int64_t cr = (int64_t)t * S::R + r;
          const T x = xv((int64_t)m.sell_col[cr]);
          if constexpr (TRANS) {
            const int32_t tp = m.sell_tpos[cr];
            y += (tp >= 0 ? vals[tp] : (T)0) * x;
          } else {
            y += __builtin_nontemporal_load(vals + (int64_t)t * S::L + lane) * x;
          }
        }
      }
      const int32_t n = in_slice ? m.sell_row[s * S::R + r] : -1;
      if (n >= 0) epi((int64_t)n, 0, y);
    }
  } else {
    __shared__ T red[kWavesPerBlock][kWave * VEC];
    for (int64_t s = s0 + w; s < s1; s += kWavesPerBlock) {
      const int32_t t0 = m.slice_ptr[s], t1 = m.slice_ptr[s + 1];
      T y[VEC];
#pragma unroll
      for (int k = 0; k < VEC; ++k) y[k] = 0;
      if (in_slice) {
#pragma unroll 8
        for (int32_t t = t0; t < t1; ++t) {
          const int64_t cr = (int64_t)t * S::R + r;
          const T x = xv((int64_t)m.sell_col[cr] * VEC + i);
          if constexpr (TRANS) {
            const int32_t tp = m.sell_tpos[cr];
#pragma unroll
            for (int k = 0; k < VEC; ++k) y[k] += (tp >= 0 ? vals[(int64_t)tp + k * VEC + i] : (T)0) * x;
          } else {
            const T* vc = vals + ((int64_t)t * S::L + lane) * VEC;
#pragma unroll
            for (int k = 0; k < VEC; ++k) y[k] += __builtin_nontemporal_load(vc + k) * x;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < VEC; ++k) red[w][lane * VEC + k] = y[k];
      wave_lds_sync();
      if (in_slice) {
        T yi = 0;
#pragma unroll
        for (int j = 0; j < VEC; ++j) yi += red[w][(r * VEC + j) * VEC + i];
        const int32_t n = m.sell_row[s * S::R + r];
        if (n >= 0) epi((int64_t)n, i, yi);
      }
      wave_lds_sync();  // the next slice rewrites red
    }
  }
}

// Contiguous slice range of this workgroup (XCD-remapped).
__device__ __forceinline__ void row_range(int64_t nslices, int64_t& r0, int64_t& r1) {
  const int64_t per = ceil_div(nslices, gridDim.x);
  r0 = (int64_t)logical_block() * per;
  r1 = r0 + per < nslices ? r0 + per : nslices;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kHeader = 4096 + (size_t)kMaxGrid * 4 * sizeof(double);

inline int rows_grid(int64_t nslices) {
  int64_t g = ceil_div(nslices, kWavesPerBlock);
  if (g >= 64) g = ceil_div(g, kNumXcd) * kNumXcd;  // whole XCDs, rounded up: no workgroup gets a second round
  if (g > kMaxGrid) g = kMaxGrid;
  return (int)(g < 1 ? 1 : g);
}
inline int elem_grid(int64_t n) {
  int64_t g = ceil_div(n, kBlock * 4);
  if (g >= 64) g = ceil_div(g, kNumXcd) * kNumXcd;  // whole XCDs, rounded up: no workgroup gets a second round
  if (g > kMaxGrid) g = kMaxGrid;
  return (int)(g < 1 ? 1 : g);
}

inline MatArgs mat_args(const hf_pattern* p) {
  MatArgs m;
  m.nrows = pattern_nnode(p);
  m.nslices = pattern_nslices(p);
  m.slice_ptr = pattern_slice_ptr(p);
  m.sell_col = pattern_sell_col(p);
  m.sell_tpos = pattern_sell_tpos(p);
  m.sell_row = pattern_sell_row(p);
  return m;
}

// host-side hooks implemented in krylov.hip
extern bool g_prof;
int prof_begin(hipStream_t s);
int prof_end(hipStream_t s);
int prof_collect();
int prof_events(hipEvent_t* e0, hipEvent_t* e1);
// Convergence poll: two pinned host slots for the scalars and an event per slot.
// ps == nullptr: the calling thread's own state, created by poll_init() on
// first use; the allocation-free Krylov entry point passes a per-device state
// made by hf_noalloc_init().
struct PollState {
  Scalars* host = nullptr;  // 2 slots, pinned
  hipEvent_t ev[2] = {nullptr, nullptr};
};
int poll_init();
int poll_post(const Scalars* sc, int slot, hipStream_t s, PollState* ps = nullptr);
// the learned network as a multigrid fine-level smoother (lpc.hip, used by mg.hip):
// e += D^-1 res + s D^-1/2 mask N([D^-1/2 res / s, g]), s = rms(D^-1/2 res), on the
// nx*ny*nz node grid; its kernels exit at once while *skip != 0
struct LpcSmoother {
  void* buf = nullptr;            // lpc_smoother_bytes(nnode), packed by lpc_smoother_load
  const float* gfield = nullptr;  // nnode input feature (caller's memory)
  int nx = 0, ny = 0, nz = 0;
};
size_t lpc_smoother_bytes(int64_t nnode);
int lpc_smoother_load(const LpcSmoother& L, const float* const* weights, hipStream_t s);
int lpc_smoother_apply(const LpcSmoother& L, const float* res, const float* dinv, const float* mask, float* e,
                       const int32_t* skip, hipStream_t s);
int poll_wait(int slot, Scalars& out, PollState* ps = nullptr);

}  // namespace hf
