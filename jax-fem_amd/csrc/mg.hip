// Geometric multigrid preconditioner on the structured node lattice of a box
// mesh (node n = (x*ny + y)*nz + z, generate_mesh.py:170-187), vec = 3, FP32,
// and the flexible CG it preconditions.  No reference counterpart: the
// reference's Krylov solves use Jacobi only (jax_fem/solver.py:68-69).
//
// Levels.  Level 0 is the tangent's 27-point node-block stencil with the
// Dirichlet DOFs masked out of its rows and columns (S0 = M A M).  Level l+1
// halves the intervals of every dimension (n -> ceil(n/2); coarse node c sits
// at fine position min(2c, n), so an odd count ends in one short interval) and
// carries the Galerkin operator A_{l+1} = P^T A_l P of the trilinear
// prolongation P (per dimension: weight 1 at the coarse node's own position,
// 1/2 at a fine-only neighbour).  A_l P is formed first (fine rows, 27 slots of
// coarse columns around x/2), then P^T (A_l P); both stay 27-point.  The
// coarsest level (<= kDenseMax DOF) is inverted densely in FP64 (Gauss-Jordan
// in LDS on the symmetrised matrix).
//
// Stencils are stored slot-major: S[(o*9 + 3i + j)*nnode + n] is entry (i, j)
// of the block that couples node n to its neighbour at offset
// o = (dx+1)*9 + (dy+1)*3 + (dz+1); slots that leave the lattice hold zeros.
// A thread per node reads its 243 values with every wavefront load coalesced,
// and neighbour indices are clamped to the lattice (the zero slots cancel them).
//
// V-cycle at level l (input r, output e), nu sweeps per side:
//   e = w Dinv r;  t = r - A e;  r_{l+1} = P^T t;  e_{l+1} = V(r_{l+1});
//   e += M P e_{l+1};  e = e + w Dinv (r - A e)
// with Dinv the inverse of each node's 3x3 diagonal block and M the Dirichlet
// mask (level 0 only).  The same smoother on both sides makes the cycle a
// symmetric operator.
//
// Outer method: flexible CG with one kept direction (Notay's FCG(1)) in the
// Chronopoulos-Gear arrangement of krylov.hip: z = V(r), w = A z and the dot
// products r.z, z.w, z.q in one phase, then
//   beta = -z.q / p.q,  p = z + beta p,  q = w + beta q  (q = A p),
//   p.q <- z.w + 2 beta z.q + beta^2 p.q,  alpha = r.z / p.q,
//   x += alpha p,  r -= alpha q.
// The flexible beta keeps the method valid when the preconditioner is not
// exactly symmetric (FP32 rounding, or a learned fine-level smoother).
#include "common.hpp"
#include "hipfem.h"

#include <hip/hip_ext.h>
#include <hip/hip_fp16.h>

#include "krylov_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

namespace hf {
const int32_t* pattern_blkpos(const hf_pattern* p);
const int32_t* pattern_sell_pos(const hf_pattern* p);
const int32_t* pattern_dpos(const hf_pattern* p);

namespace {
constexpr int kMgMaxLevels = 16;
constexpr int kDenseMax = 88;   // FP64 dense coarsest matrix in 62 KB of LDS
constexpr int kMgBlock = 256;
constexpr int kSt = 243;        // 27 slots x 9 values
constexpr int kMaxDegree = 8;   // smoothing steps per side
// A level's smoothing coefficients in device memory: s0, c1[kMaxDegree], c2[kMaxDegree]
// (written on the set-up stream, read by the cycle's kernels, so a captured
// cycle stays valid across set-ups).
constexpr int kCoefStride = 1 + 2 * kMaxDegree;

struct Lat {
  int nx, ny, nz;
  int64_t n;
};

inline Lat make_lat(int nx, int ny, int nz) {
  Lat g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.n = (int64_t)nx * ny * nz;
  return g;
}

__device__ __forceinline__ void lat_xyz(const Lat& g, int64_t n, int& x, int& y, int& z) {
  z = (int)(n % g.nz);
  const int64_t q = n / g.nz;
  y = (int)(q % g.ny);
  x = (int)(q / g.ny);
}
// the same for node ids below 2^31, with 32-bit divisions (the 64-bit ones are
// long software sequences; kernels that decode a node per matrix block use this)
__device__ __forceinline__ void lat_xyz32(const Lat& g, uint32_t n, int& x, int& y, int& z) {
  const uint32_t q = n / (uint32_t)g.nz;
  z = (int)(n - q * (uint32_t)g.nz);
  y = (int)(q % (uint32_t)g.ny);
  x = (int)(q / (uint32_t)g.ny);
}
__device__ __forceinline__ int64_t lat_id(const Lat& g, int x, int y, int z) {
  return ((int64_t)x * g.ny + y) * g.nz + z;
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The fine level of a symmetric operator stores only the diagonal and the 13
// upper offsets (o >= 13) of its stencil: block (n, o) of a lower offset is the
// transpose of block (n + off(o), 26 - o), the neighbour's upper slot 13 - o.
// Half the bytes of the cycle's largest stream; stencil_row reads a neighbour's
// slot at the step its owner reads it, so most of those second reads hit L2.
constexpr int kHalfSlots = 14;
__host__ __device__ constexpr int diag_slot(bool half) { return half ? 0 : 13; }

// Block (n, o) of a stencil with N nodes into B (row-major); m is the (clamped)
// neighbour index and valid whether the neighbour is inside the lattice.  Full
// stencils hold zeros in slots that leave the lattice.
template <bool HALF>
__device__ __forceinline__ void load_block(const float* __restrict__ S, int64_t N, int64_t n, int o, int64_t m,
                                           bool valid, float (&B)[9]) {
  if constexpr (!HALF) {
#pragma unroll
    for (int e = 0; e < 9; ++e) B[e] = S[(int64_t)(o * 9 + e) * N + n];
  } else {
    if (o >= 13) {
#pragma unroll
      for (int e = 0; e < 9; ++e) B[e] = S[(int64_t)((o - 13) * 9 + e) * N + n];
    } else {
      const float vf = valid ? 1.f : 0.f;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) B[3 * i + j] = vf * S[(int64_t)((13 - o) * 9 + 3 * j + i) * N + m];
    }
  }
}

// The fine level's stencil in FP16 for the smoothing passes (HF_MG_FINE_F16):
// the half-stored slots with each block's 9 entries packed in 5 half2 words
// (entries 2k, 2k + 1; the fifth word's upper half is zero), word-major over
// the nodes: W[(s*5 + k)*nnode + n].  The entries are scaled by a power of two
// that puts the largest magnitude in [2^14, 2^15) (exact: FP16's 11-bit
// significand is the only rounding), and the products are scaled back by its
// inverse.  280 bytes a node instead of 504.
constexpr int kF16Words = 5;
template <bool HALF>
__device__ __forceinline__ void load_block(const __half2* __restrict__ W, int64_t N, int64_t n, int o, int64_t m,
                                           bool valid, float (&B)[9]) {
  static_assert(HALF, "FP16 stencils are half-stored");
  const bool own = o >= 13;
  const __half2* p = W + (int64_t)((own ? o - 13 : 13 - o) * kF16Words) * N + (own ? n : m);
  float T[2 * kF16Words];
#pragma unroll
  for (int k = 0; k < kF16Words; ++k) {
    const float2 f = __half22float2(p[(int64_t)k * N]);
    T[2 * k] = f.x;
    T[2 * k + 1] = f.y;
  }
  if (own) {
#pragma unroll
    for (int e = 0; e < 9; ++e) B[e] = T[e];
  } else {
    const float vf = valid ? 1.f : 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) B[3 * i + j] = vf * T[3 * j + i];
  }
}
template <typename ST>
constexpr bool is_f16() { return std::is_same<ST, __half2>::value; }

// 1-D transfer between nf fine intervals and their coarse nodes.
__device__ __forceinline__ int cpos(int c, int nf) { return 2 * c < nf ? 2 * c : nf; }
__device__ __forceinline__ bool is_cpos(int f, int nf) { return (f % 2 == 0) || f == nf; }
__device__ __forceinline__ float p1(int f, int c, int nf) {
  const int p = cpos(c, nf);
  if (f == p) return 1.f;
  if ((f == p - 1 || f == p + 1) && !is_cpos(f, nf)) return 0.5f;
  return 0.f;
}

// Contiguous node range of this workgroup (XCD-remapped, as the Krylov kernels).
__device__ __forceinline__ void node_range(int64_t n, int64_t& i0, int64_t& i1) {
  const int64_t per = ceil_div(n, gridDim.x);
  i0 = (int64_t)logical_block() * per;
  i1 = i0 + per < n ? i0 + per : n;
}
inline int node_grid(int64_t n) {
  int64_t g = ceil_div(n, kMgBlock);
  if (g >= 64) g = ceil_div(g, kNumXcd) * kNumXcd;  // whole XCDs, rounded up: no workgroup gets a second round
  if (g > kMaxGrid) g = kMaxGrid;
  return (int)(g < 1 ? 1 : g);
}

// The 27 slots in mirrored-pair order: 13 (the diagonal), then p, 26 - p for
// p = 0..12 (t = 2p + 1, 2p + 2); see stencil_row.
__device__ __forceinline__ int pair_slot(int t) {
  return t == 0 ? 13 : ((t & 1) ? (t - 1) / 2 : 26 - (t - 2) / 2);
}

// out = sum over the 27 slots of S[n] v[neighbour]  (clamped neighbours).
// Offsets go in mirrored pairs (p, 26 - p).  With half storage both blocks of a
// pair come from stored slot 13 - p (node n's own, and its neighbour
// n + off(p)'s transposed), and every node reads that slot at the same step,
// so the neighbour's copy is read while its owner's wave fetches it as well;
// in offset order the second read came half a kernel later and mostly missed
// L2 (fine-pass fetch 290 -> 210 MB at 65^3 nodes).  The pair loop is not
// unrolled: unrolled, the compiler hoists all 27 block loads (256 VGPRs).
template <bool HALF, typename ST>
__device__ __forceinline__ void stencil_row_stored(const ST* __restrict__ S, const Lat& g, int64_t n, int x, int y,
                                                   int z, const float* __restrict__ v, float (&out)[3],
                                                   const float* __restrict__ scl) {
  {
    float B[9];
    load_block<HALF>(S, g.n, n, 13, n, true, B);
    const float v0 = v[3 * n], v1 = v[3 * n + 1], v2 = v[3 * n + 2];
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = B[3 * i] * v0 + B[3 * i + 1] * v1 + B[3 * i + 2] * v2;
  }
#pragma unroll 1
  for (int p = 0; p < 13; ++p) {
    const int ox = p / 9 - 1, oy = (p / 3) % 3 - 1, oz = p % 3 - 1;
    const int xl = x + ox, yl = y + oy, zl = z + oz;
    const int xu = x - ox, yu = y - oy, zu = z - oz;
    const int64_t ml = lat_id(g, clampi(xl, 0, g.nx - 1), clampi(yl, 0, g.ny - 1), clampi(zl, 0, g.nz - 1));
    const int64_t mu = lat_id(g, clampi(xu, 0, g.nx - 1), clampi(yu, 0, g.ny - 1), clampi(zu, 0, g.nz - 1));
    const bool valid_l = xl >= 0 && xl < g.nx && yl >= 0 && yl < g.ny && zl >= 0 && zl < g.nz;
    float Bl[9], Bu[9];
    load_block<HALF>(S, g.n, n, p, ml, valid_l, Bl);
    load_block<HALF>(S, g.n, n, 26 - p, mu, true, Bu);
    const float l0 = v[3 * ml], l1 = v[3 * ml + 1], l2 = v[3 * ml + 2];
    const float u0 = v[3 * mu], u1 = v[3 * mu + 1], u2 = v[3 * mu + 2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      out[i] += (Bl[3 * i] * l0 + Bl[3 * i + 1] * l1 + Bl[3 * i + 2] * l2) +
                (Bu[3 * i] * u0 + Bu[3 * i + 1] * u1 + Bu[3 * i + 2] * u2);
  }
  if constexpr (is_f16<ST>()) {
    const float sinv = scl[1];
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] *= sinv;
  }
}

// Matrix-free fine level (hf_mg_set_cell_operator): A = sum_c scale[c] K over the
// lattice's cells, K the 24x24 cell matrix at unit scale in 3x3 blocks (q, q') of
// corners q = 4 dx + 2 dy + dz (dx, dy, dz in {0, 1}).  A box cell and its law are
// symmetric under the reflections through the cell's mid-planes, so block (a, b) =
// R_a B_{a^b} R_a with B_q = block (0, q) and R_a the reflection of the axes where
// corner a has a 1, and B_q^T = R_q B_q R_q: the 8 blocks' upper triangles (48
// values, held in scalar registers) give all 576 entries.  The host checks this
// before it accepts a matrix.
struct MfOp {
  float b[8 * 6];  // B_q entries (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
};
__host__ __device__ constexpr int mf_tri(int r, int c) {
  const int lo = r < c ? r : c, hi = r < c ? c : r;
  return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
}
__host__ __device__ constexpr float mf_refl(int q, int axis) { return ((q >> (2 - axis)) & 1) ? -1.f : 1.f; }
// sign of entry (r, c) of block (a, b) relative to B_{a^b}'s stored upper-triangle value
__host__ __device__ constexpr float mf_sign(int a, int b, int r, int c) {
  return mf_refl(a, r) * mf_refl(a, c) * (r > c ? mf_refl(a ^ b, r) * mf_refl(a ^ b, c) : 1.f);
}

// Row of node (x, y, z): over its cells c = (x - 1 + i, y - 1 + j, z - 1 + k), in
// which it is corner a = 7 - (4i + 2j + k), scale[c] sum_{q' != a} K(a, q')
// (v[corner q'] - v[node]).  K's block rows sum to zero (a rigid translation
// carries no force), so this is the row of A, with the diagonal block implied:
// the differences of neighbouring values keep their digits where a smooth v
// would cancel them in sum_q' K(a, q') v[q'] (the FP32 error then grows like
// 1/h), and the operator annihilates translations exactly.  The 26 neighbours'
// differences are formed once; a clamped neighbour meets only cells outside the
// lattice, whose scale is taken as 0.
__device__ __forceinline__ void mf_row(const MfOp* __restrict__ op, const float* __restrict__ cs, const Lat& g,
                                       int x, int y, int z, const float* __restrict__ v, float (&out)[3]) {
  // 32-bit offsets (3 n < 2^32: hf_mg_create keeps n below 2^30): per axis, the
  // clamped neighbour coordinates' share of 3 * (lattice index) and of the cell index
  const uint32_t nz = (uint32_t)g.nz, nyz = (uint32_t)g.ny * nz;
  const uint32_t ncz = nz - 1, ncyz = (uint32_t)(g.ny - 1) * ncz;
  uint32_t X[3], Y[3], Z[3], CX[2], CY[2], CZ[2];
  bool inx[2], iny[2], inz[2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    X[d] = 3u * (uint32_t)clampi(x + d - 1, 0, g.nx - 1) * nyz;
    Y[d] = 3u * (uint32_t)clampi(y + d - 1, 0, g.ny - 1) * nz;
    Z[d] = 3u * (uint32_t)clampi(z + d - 1, 0, g.nz - 1);
  }
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    inx[d] = x - 1 + d >= 0 && x - 1 + d < g.nx - 1;
    iny[d] = y - 1 + d >= 0 && y - 1 + d < g.ny - 1;
    inz[d] = z - 1 + d >= 0 && z - 1 + d < g.nz - 1;
    CX[d] = (uint32_t)clampi(x - 1 + d, 0, g.nx - 2) * ncyz;
    CY[d] = (uint32_t)clampi(y - 1 + d, 0, g.ny - 2) * ncz;
    CZ[d] = (uint32_t)clampi(z - 1 + d, 0, g.nz - 2);
  }
  const float* pn = v + (X[1] + Y[1] + Z[1]);
  const float v0 = pn[0], v1 = pn[1], v2 = pn[2];
  float u[27][3];
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    if (o == 13) {
      u[o][0] = u[o][1] = u[o][2] = 0.f;
      continue;
    }
    const float* pm = v + (X[o / 9] + Y[(o / 3) % 3] + Z[o % 3]);   // one 64-bit address, 3 loads
    u[o][0] = pm[0] - v0;
    u[o][1] = pm[1] - v1;
    u[o][2] = pm[2] - v2;
  }
  out[0] = out[1] = out[2] = 0.f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int i = c >> 2, j = (c >> 1) & 1, k = c & 1;
    // the clamped index is always a valid cell: load, then zero outside the lattice
    const float scl = cs[CX[i] + CY[j] + CZ[k]];
    const float sc = (inx[i] && iny[j] && inz[k]) ? scl : 0.f;
    const int qa = 7 - c;
    float t[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int qb = 0; qb < 8; ++qb) {
      if (qb == qa) continue;   // the node itself: difference 0
      const int o = (i + (qb >> 2)) * 9 + (j + ((qb >> 1) & 1)) * 3 + (k + (qb & 1));
      const float* B = op->b + (qa ^ qb) * 6;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
          // the sign is a compile-time constant here: it folds into the FMA
          const float bv = B[mf_tri(r, cc)];
          if (mf_sign(qa, qb, r, cc) > 0.f)
            t[r] += bv * u[o][cc];
          else
            t[r] -= bv * u[o][cc];
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) out[r] += sc * t[r];
  }
}

// out = row n of the level's operator times v: a stored stencil (FP32, or FP16
// words with scl[1] the inverse scale) or the matrix-free cell operator (scl the
// per-cell scales)
template <bool HALF, typename ST = float>
__device__ __forceinline__ void stencil_row(const ST* __restrict__ S, const Lat& g, int64_t n, int x, int y, int z,
                                            const float* __restrict__ v, float (&out)[3],
                                            const float* __restrict__ scl = nullptr) {
  if constexpr (std::is_same<ST, MfOp>::value) {
    mf_row(S, scl, g, x, y, z, v, out);
  } else {
    stencil_row_stored<HALF>(S, g, n, x, y, z, v, out, scl);
  }
}

__device__ __forceinline__ void block_apply(const float* __restrict__ D, int64_t n, int64_t node, const float (&v)[3],
                                            float (&out)[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
    out[i] = D[(3 * i) * n + node] * v[0] + D[(3 * i + 1) * n + node] * v[1] + D[(3 * i + 2) * n + node] * v[2];
}

// ---------------------------------------------------------------------------
// set-up kernels

// Level 0: S0 = M A M from the tangent's SELL values, one thread per (node,
// slot = blockIdx.y).  On the lattice a row's columns are its in-lattice
// neighbours in sorted order, which is the lexicographic order of the offsets,
// so the block of offset o is the row's k-th, k the number of in-lattice
// offsets before o, and sits in slot k of the node's slice row.
// mg_check_lattice verifies that structure once per pattern (colidx and
// blkpos against the rule, and the row lengths); mg_fine_stencil then reads
// only the values and the mask.
struct LatSlot {
  bool inside;
  int k;          // index of the offset's block within the row
  int64_t c;      // neighbour node
  int count;      // in-lattice neighbours of the node (row length)
};
__device__ __forceinline__ LatSlot lat_slot(const Lat& g, int64_t n, int o) {
  int p[3];
  lat_xyz(g, n, p[0], p[1], p[2]);
  const int np[3] = {g.nx, g.ny, g.nz};
  const int off[3] = {o / 9 - 1, (o / 3) % 3 - 1, o % 3 - 1};
  LatSlot r;
  r.inside = true;
  int na[3], rank[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int lo = p[d] > 0, hi = p[d] < np[d] - 1;
    na[d] = 1 + lo + hi;
    rank[d] = off[d] < 0 ? 0 : (off[d] == 0 ? lo : lo + 1);
    r.inside = r.inside && (off[d] >= 0 || lo) && (off[d] <= 0 || hi);
  }
  r.k = (rank[0] * na[1] + rank[1]) * na[2] + rank[2];
  r.count = na[0] * na[1] * na[2];
  r.c = r.inside ? lat_id(g, p[0] + off[0], p[1] + off[1], p[2] + off[2]) : -1;
  return r;
}

__global__ void __launch_bounds__(kMgBlock) mg_check_lattice(const int32_t* __restrict__ rowptr,
                                                             const int32_t* __restrict__ colidx,
                                                             const int32_t* __restrict__ blkpos,
                                                             const int32_t* __restrict__ sell_pos,
                                                             const int32_t* __restrict__ slice_ptr, int R, Lat g,
                                                             int32_t* __restrict__ bad) {
  const int64_t n = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  const int o = blockIdx.y;
  if (n >= g.n) return;
  const LatSlot ls = lat_slot(g, n, o);
  const int32_t r0 = rowptr[n], len = rowptr[n + 1] - r0;
  if (o == 13 && len != ls.count) bad[0] = 1;
  if (!ls.inside) return;
  if (ls.k >= len) {
    bad[0] = 1;
    return;
  }
  const int32_t q = sell_pos[n];
  const int64_t pos = ((int64_t)(slice_ptr[q / R] + ls.k) * R + q % R) * 9;
  if (colidx[r0 + ls.k] != (int32_t)ls.c || blkpos[r0 + ls.k] != pos) bad[0] = 1;
}

template <bool HALF>
__global__ void __launch_bounds__(kMgBlock) mg_fine_stencil(const int32_t* __restrict__ rowptr,
                                                            const int32_t* __restrict__ sell_pos,
                                                            const int32_t* __restrict__ slice_ptr, int R,
                                                            const float* __restrict__ vals,
                                                            const float* __restrict__ mask, Lat g,
                                                            float* __restrict__ S) {
  const int64_t n = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  const int so = blockIdx.y;                 // stored slot
  const int o = HALF ? 13 + so : so;         // its offset
  if (n >= g.n) return;
  const LatSlot ls = lat_slot(g, n, o);
  float out[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) out[e] = 0.f;
  if (ls.inside && ls.k < rowptr[n + 1] - rowptr[n]) {
    const int32_t q = sell_pos[n];
    const float* v = vals + ((int64_t)(slice_ptr[q / R] + ls.k) * R + q % R) * 9;
    const float mr[3] = {mask[3 * n], mask[3 * n + 1], mask[3 * n + 2]};
    const float mc[3] = {mask[3 * ls.c], mask[3 * ls.c + 1], mask[3 * ls.c + 2]};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) out[3 * i + j] = v[3 * j + i] * mr[i] * mc[j];
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) S[(int64_t)(so * 9 + e) * g.n + n] = out[e];
}

// The level-0 half-stored stencil from the cell operator (hf_mg_set_cell_operator),
// thread per (node, slot = blockIdx.y): block (n, 13 + slot) = sum over the cells
// holding n (corner a) and its neighbour at that offset (corner b) of scale[c]
// K(a, b); the diagonal block is minus the sum of the row's others, as the
// matrix-free rows imply.  Slots that leave the lattice hold zeros.
__global__ void __launch_bounds__(kMgBlock) mg_assemble_cellop(const MfOp* __restrict__ op,
                                                               const float* __restrict__ cs, Lat g,
                                                               float* __restrict__ S) {
  const int64_t n = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  if (n >= g.n) return;
  const int slot = blockIdx.y;
  const int o = 13 + slot;
  const int off[3] = {o / 9 - 1, (o / 3) % 3 - 1, o % 3 - 1};
  int p[3];
  lat_xyz(g, n, p[0], p[1], p[2]);
  const int np[3] = {g.nx, g.ny, g.nz};
  bool inside = true;
  for (int d = 0; d < 3; ++d) inside = inside && p[d] + off[d] >= 0 && p[d] + off[d] < np[d];
  float B[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (inside) {
    for (int c = 0; c < 8; ++c) {
      const int cb[3] = {c >> 2, (c >> 1) & 1, c & 1};
      bool cin = true;
      int ab[3];
      for (int d = 0; d < 3; ++d) {
        const int pc = p[d] - 1 + cb[d];
        cin = cin && pc >= 0 && pc < np[d] - 1;
        ab[d] = 1 - cb[d];   // the node's corner in this cell
      }
      if (!cin) continue;
      const float sc = cs[((int64_t)(p[0] - 1 + cb[0]) * (g.ny - 1) + (p[1] - 1 + cb[1])) * (g.nz - 1) +
                          (p[2] - 1 + cb[2])];
      const int a = 4 * ab[0] + 2 * ab[1] + ab[2];
      for (int b = 0; b < 8; ++b) {
        if (b == a) continue;
        const int bb[3] = {b >> 2, (b >> 1) & 1, b & 1};
        // slot 0 (the diagonal): minus every other block of the row; else the block
        // whose corner is the neighbour
        const bool take = slot == 0 || (bb[0] - ab[0] == off[0] && bb[1] - ab[1] == off[1] &&
                                        bb[2] - ab[2] == off[2]);
        if (!take) continue;
        const float w = slot == 0 ? -sc : sc;
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) B[3 * r + cc] += w * mf_sign(a, b, r, cc) * op->b[(a ^ b) * 6 + mf_tri(r, cc)];
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) S[(int64_t)(slot * 9 + e) * g.n + n] = B[e];
}

// The same stencil on a box mesh whose cells are the lattice's (cell (i, j, k) =
// (i * ncy + j) * ncz + k, corner (dx, dy, dz) of cell c at local index
// corner[dx * 4 + dy * 2 + dz]): the contributing cells, their order and both
// local indices follow from the coordinates, so a thread issues all its block
// loads at once.  Cells are visited in ascending id, as the generic kernel does.
struct CornerTable {
  int8_t idx[8];
};
__global__ void __launch_bounds__(kMgBlock) mg_assemble_hex8_box(Lat g, int64_t ncell, CornerTable ct,
                                                                 const float* __restrict__ elem,
                                                                 float* __restrict__ S) {
  const int64_t n = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  if (n >= g.n) return;
  const int so = blockIdx.y, o = 13 + so;
  int p[3];
  lat_xyz(g, n, p[0], p[1], p[2]);
  const int off[3] = {o / 9 - 1, (o / 3) % 3 - 1, o % 3 - 1};
  const int nc[3] = {g.nx - 1, g.ny - 1, g.nz - 1};
  // per dimension the candidate cells c_d in {p_d - 1, p_d} holding both p_d and p_d + off_d
  int lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = max(max(p[d] - 1, p[d] + off[d] - 1), 0);
    hi[d] = min(min(p[d], p[d] + off[d]), nc[d] - 1);
  }
  float acc[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) acc[e] = 0.f;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int cx = p[0] - 1 + i, cy = p[1] - 1 + j, cz = p[2] - 1 + k;
        if (cx < lo[0] || cx > hi[0] || cy < lo[1] || cy > hi[1] || cz < lo[2] || cz > hi[2]) continue;
        const int64_t c = ((int64_t)cx * nc[1] + cy) * nc[2] + cz;
        const int a = ct.idx[(p[0] - cx) * 4 + (p[1] - cy) * 2 + (p[2] - cz)];
        const int b = ct.idx[(p[0] + off[0] - cx) * 4 + (p[1] + off[1] - cy) * 2 + (p[2] + off[2] - cz)];
        const float* src = elem + (int64_t)((a * 8 + b) * 9) * ncell + c;
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[e] += src[(int64_t)e * ncell];
      }
#pragma unroll
  for (int e = 0; e < 9; ++e) S[(int64_t)(so * 9 + e) * g.n + n] = acc[e];
}

// Level-0 stencil of a HEX8 tangent straight from the element Jacobians (the
// upper blocks, cell-innermost: elem[((a * 8 + b) * 9 + e) * ncell + c], see
// hf_hex8_jacobian_upper_f32), without the SELL matrix: the diagonal and the 13
// upper blocks of every node, Dirichlet rows included (the multigrid masks them
// where it must).  A thread per (node, stored slot; the slot is blockIdx.y)
// walks the node's cells in ascending order (the diagonal block's contribution
// list), finds the neighbour among the cell's nodes and adds that block: the
// same FP32 additions in the same order as the SELL assembly of the block.
// Threads of consecutive nodes read consecutive cells.
__global__ void __launch_bounds__(kMgBlock) mg_assemble_hex8(const int32_t* __restrict__ diag,
                                                             const int32_t* __restrict__ seg,
                                                             const int32_t* __restrict__ contrib,
                                                             const int32_t* __restrict__ cells, Lat g,
                                                             int64_t ncell, const float* __restrict__ elem,
                                                             float* __restrict__ S) {
  const int64_t n = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  if (n >= g.n) return;
  const int so = blockIdx.y, o = 13 + so;
  int x, y, z;
  lat_xyz(g, n, x, y, z);
  const int xr = x + o / 9 - 1, yr = y + (o / 3) % 3 - 1, zr = z + o % 3 - 1;
  float acc[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) acc[e] = 0.f;
  const int32_t db = diag[n];
  if (db >= 0 && xr < g.nx && yr < g.ny && zr < g.nz && xr >= 0 && yr >= 0 && zr >= 0) {
    const int32_t m = (int32_t)lat_id(g, xr, yr, zr);
    for (int32_t s = seg[db]; s < seg[db + 1]; ++s) {
      const int32_t id = contrib[s];  // (cell * 8 + a) * 8 + a
      const int64_t c = id / 64;
      const int a = (id / 8) % 8;
      const int4 c0 = *reinterpret_cast<const int4*>(cells + c * 8);
      const int4 c1 = *reinterpret_cast<const int4*>(cells + c * 8 + 4);
      const int b = c0.x == m ? 0 : c0.y == m ? 1 : c0.z == m ? 2 : c0.w == m ? 3 :
                    c1.x == m ? 4 : c1.y == m ? 5 : c1.z == m ? 6 : c1.w == m ? 7 : -1;
      if (b < 0) continue;  // the cell does not hold the neighbour
      const float* src = elem + (int64_t)((a * 8 + b) * 9) * ncell + c;
#pragma unroll
      for (int e = 0; e < 9; ++e) acc[e] += src[(int64_t)e * ncell];
    }
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) S[(int64_t)(so * 9 + e) * g.n + n] = acc[e];
}


// the full 27-slot stencil of a half-stored one (read-back)
__global__ void __launch_bounds__(kMgBlock) mg_expand_half(const float* __restrict__ Sh, Lat g,
                                                           float* __restrict__ S) {
  const int64_t n = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  const int o = blockIdx.y;
  if (n >= g.n) return;
  int x, y, z;
  lat_xyz(g, n, x, y, z);
  const int xr = x + o / 9 - 1, yr = y + (o / 3) % 3 - 1, zr = z + o % 3 - 1;
  const bool valid = xr >= 0 && xr < g.nx && yr >= 0 && yr < g.ny && zr >= 0 && zr < g.nz;
  const int64_t m = lat_id(g, clampi(xr, 0, g.nx - 1), clampi(yr, 0, g.ny - 1), clampi(zr, 0, g.nz - 1));
  float B[9];
  load_block<true>(Sh, g.n, n, o, m, valid, B);
#pragma unroll
  for (int e = 0; e < 9; ++e) S[(int64_t)(o * 9 + e) * g.n + n] = valid ? B[e] : 0.f;
}

// Inverse of each node's diagonal block (slot 13); masked DOFs (mask 0) get a
// unit diagonal.  A block that is not safely invertible falls back to the
// inverse of its positive diagonal entries (0 elsewhere).
// 1 / diag(A) of the fine level (the scalar Jacobi the learned smoother's network
// was trained with); 1 on Dirichlet DOFs, whose residual entries are masked to 0
template <bool HALF>
__global__ void __launch_bounds__(kMgBlock) mg_scalar_dinv(const float* __restrict__ S, const float* __restrict__ mask,
                                                           int64_t n, float* __restrict__ dinv) {
  const int64_t node = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  if (node >= n) return;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float a = S[(int64_t)(diag_slot(HALF) * 9 + 4 * i) * n + node];
    const bool free = !mask || mask[3 * node + i] != 0.f;
    dinv[3 * node + i] = (free && a != 0.f) ? 1.f / a : 1.f;
  }
}

template <bool HALF>
__global__ void __launch_bounds__(kMgBlock) mg_block_dinv(const float* __restrict__ S, const float* __restrict__ mask,
                                                          int64_t n, float* __restrict__ Dinv) {
  const int64_t node = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  if (node >= n) return;
  double a[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) a[e] = (double)S[(int64_t)(diag_slot(HALF) * 9 + e) * n + node];
  float mk[3] = {1.f, 1.f, 1.f};
  if (mask) {
    // the block of M A M (the stored block may hold the Dirichlet rows), with 1 on
    // the masked diagonal so that it inverts; the masked rows and columns of the
    // inverse are zeroed below
#pragma unroll
    for (int i = 0; i < 3; ++i) mk[i] = mask[3 * node + i];
#pragma unroll
    for (int e = 0; e < 9; ++e) a[e] *= (double)(mk[e / 3] * mk[e % 3]);
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (mk[i] == 0.f) a[4 * i] += 1.0;
  }
  double c[9];
  c[0] = a[4] * a[8] - a[5] * a[7];
  c[1] = a[2] * a[7] - a[1] * a[8];
  c[2] = a[1] * a[5] - a[2] * a[4];
  c[3] = a[5] * a[6] - a[3] * a[8];
  c[4] = a[0] * a[8] - a[2] * a[6];
  c[5] = a[2] * a[3] - a[0] * a[5];
  c[6] = a[3] * a[7] - a[4] * a[6];
  c[7] = a[1] * a[6] - a[0] * a[7];
  c[8] = a[0] * a[4] - a[1] * a[3];
  const double det = a[0] * c[0] + a[1] * c[3] + a[2] * c[6];
  const double sc = fmax(fabs(a[0]), fmax(fabs(a[4]), fabs(a[8])));
  if (det > 1e-12 * sc * sc * sc && isfinite(det)) {
#pragma unroll
    for (int e = 0; e < 9; ++e) Dinv[(int64_t)e * n + node] = (float)(c[e] / det) * mk[e / 3] * mk[e % 3];
  } else {
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const bool d = (e % 4) == 0;
      Dinv[(int64_t)e * n + node] = (d && a[e] > 0.0) ? (float)(1.0 / a[e]) * mk[e / 3] : 0.f;
    }
  }
}

// Galerkin product P^T A P, one dimension at a time: P = Px (x) Py (x) Pz, so it is
// (Pz^T (Py^T (Px^T A Px) Py) Pz), and each factor keeps the operator 27-point
// (linear interpolation of a 3-point coupling along one dimension stays 3-point).
// One stage coarsens dimension `dim` of lattice f into lattice c (the other two
// dimensions unchanged).  A thread per (output node, offsets q of the two other
// dimensions) forms the node's three output slots along dim from the fine blocks
// (row cpos(K) + a, offset d along dim), at most 3 x 3 of them; the 27 slots of
// a node are 9 threads in 9 grid planes (blockIdx.y = q), each reading its own
// blocks.  The intermediate lattices are stored as the input level is (half-
// stored for a symmetric operator).
template <bool HALF>
__global__ void __launch_bounds__(kMgBlock) mg_rap_dim(const float* __restrict__ S, Lat f, Lat c, int dim,
                                                       const float* __restrict__ mask, float* __restrict__ Out,
                                                       bool out_half) {
  const int64_t n = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  if (n >= c.n) return;
  const int q = blockIdx.y;
  int cc[3];
  lat_xyz(c, n, cc[0], cc[1], cc[2]);
  const int fdim[3] = {f.nx, f.ny, f.nz};
  const int nf = fdim[dim] - 1;
  const int nc = (dim == 0 ? c.nx : dim == 1 ? c.ny : c.nz) - 1;
  const int d1 = dim == 0 ? 1 : 0, d2 = dim == 2 ? 1 : 2;  // the two other dimensions, in order
  int off[3];                                               // slot offsets (0..2) per dimension
  off[d1] = q / 3;
  off[d2] = q % 3;
  const int K = cc[dim];
  const int pK = cpos(K, nf);
  float acc[3][9];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int e = 0; e < 9; ++e) acc[k][e] = 0.f;
#pragma unroll
  for (int a = -1; a <= 1; ++a) {
    const int i = pK + a;
    if (i < 0 || i > nf) continue;
    const float wi = p1(i, K, nf);
    if (wi == 0.f) continue;
    int rc[3] = {cc[0], cc[1], cc[2]};
    rc[dim] = i;
    const int64_t row = lat_id(f, rc[0], rc[1], rc[2]);
#pragma unroll
    for (int d = -1; d <= 1; ++d) {
      const int j = i + d;
      if (j < 0 || j > nf) continue;
      float w[3];
      bool any = false;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int Kp = K + k - 1;
        w[k] = (Kp >= 0 && Kp <= nc) ? wi * p1(j, Kp, nf) : 0.f;
        any |= (w[k] != 0.f);
      }
      if (!any) continue;
      off[dim] = d + 1;
      const int o = off[0] * 9 + off[1] * 3 + off[2];
      int nb[3] = {rc[0] + off[0] - 1, rc[1] + off[1] - 1, rc[2] + off[2] - 1};
      const bool valid = nb[0] >= 0 && nb[0] < f.nx && nb[1] >= 0 && nb[1] < f.ny && nb[2] >= 0 && nb[2] < f.nz;
      const int64_t m = lat_id(f, clampi(nb[0], 0, f.nx - 1), clampi(nb[1], 0, f.ny - 1), clampi(nb[2], 0, f.nz - 1));
      float B[9];
      load_block<HALF>(S, f.n, row, o, m, valid, B);
      if (mask) {  // the fine level's M A M
#pragma unroll
        for (int e = 0; e < 9; ++e) B[e] *= mask[3 * row + e / 3] * mask[3 * m + e % 3];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int e = 0; e < 9; ++e) acc[k][e] += w[k] * B[e];
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    off[dim] = k;
    const int O = off[0] * 9 + off[1] * 3 + off[2];
    if (out_half && O < 13) continue;  // a half-stored level keeps the diagonal and upper slots
    const int so = out_half ? O - 13 : O;
#pragma unroll
    for (int e = 0; e < 9; ++e) Out[(int64_t)(so * 9 + e) * c.n + n] = acc[k][e];
  }
}

// Node count of the two intermediate lattices of the dimension-wise product
// coarsening f into c (x coarsened, then x and y).
inline int64_t rap_dim_scratch_nodes(const Lat& f, const Lat& c) {
  return (int64_t)c.nx * f.ny * f.nz + (int64_t)c.nx * c.ny * f.nz;
}

// A_c = P^T B from B = A P held as 27 coarse-column slots per fine row (slot s
// of row i: coarse column floor(i/2) + s - 1 per dimension; the SELL fine level's
// first product), one thread per (coarse node, output slot O = blockIdx.y).
__global__ void __launch_bounds__(kMgBlock) mg_rap_ptap(const float* __restrict__ B, Lat f, Lat c,
                                                        float* __restrict__ Sc) {
  const int64_t I = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  const int O = blockIdx.y;
  if (I >= c.n) return;
  int ci[3];
  lat_xyz(c, I, ci[0], ci[1], ci[2]);
  const int nf[3] = {f.nx - 1, f.ny - 1, f.nz - 1};
  const int nc[3] = {c.nx - 1, c.ny - 1, c.nz - 1};
  const int J[3] = {ci[0] + O / 9 - 1, ci[1] + (O / 3) % 3 - 1, ci[2] + O % 3 - 1};
  float acc[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) acc[e] = 0.f;
  const bool inside = J[0] >= 0 && J[0] <= nc[0] && J[1] >= 0 && J[1] <= nc[1] && J[2] >= 0 && J[2] <= nc[2];
  if (inside) {
    // fine rows in the support of I per dimension, their weights and B slots
    int fr[3][3], sl[3][3];
    float wr[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int p = cpos(ci[d], nf[d]);
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int fidx = p + a - 1;
        const bool ok = fidx >= 0 && fidx <= nf[d];
        const int s = ok ? J[d] - fidx / 2 : 5;
        fr[d][a] = ok ? fidx : p;
        sl[d][a] = s + 1;
        wr[d][a] = (ok && s >= -1 && s <= 1) ? p1(fidx, ci[d], nf[d]) : 0.f;
      }
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
      for (int ay = 0; ay < 3; ++ay)
#pragma unroll
        for (int az = 0; az < 3; ++az) {
          const float w = wr[0][ax] * wr[1][ay] * wr[2][az];
          if (w == 0.f) continue;
          const int slot = sl[0][ax] * 9 + sl[1][ay] * 3 + sl[2][az];
          const int64_t i = lat_id(f, fr[0][ax], fr[1][ay], fr[2][az]);
          const float* bp = B + (int64_t)(slot * 9) * f.n + i;
#pragma unroll
          for (int e = 0; e < 9; ++e) acc[e] += w * bp[(int64_t)e * f.n];
        }
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) Sc[(int64_t)(O * 9 + e) * c.n + I] = acc[e];
}

// Dense FP64 inverse of the coarsest level: one workgroup, matrix in LDS,
// symmetrised, Gauss-Jordan without pivoting (SPD).  A non-positive pivot
// (a coarse DOF with no support) drops that DOF: its row and column of the
// inverse are 0; bad[1] counts such pivots.
constexpr int kDenseThreads = 1024;
__global__ void __launch_bounds__(kDenseThreads) mg_dense_setup(const float* __restrict__ S, Lat g,
                                                                double* __restrict__ Ainv, int32_t* __restrict__ bad) {
  constexpr int kCols = kDenseThreads / kDenseMax;  // threads per row (11): row ti, columns tj, tj + kCols, ...
  __shared__ double A[kDenseMax * kDenseMax];
  __shared__ double col[2 * kDenseMax];
  __shared__ double prow[2 * kDenseMax];
  const int nd = (int)(3 * g.n);
  const int tid = threadIdx.x;
  const int ti = tid / kCols, tj = tid - (tid / kCols) * kCols;
  const bool owner = ti < nd;
  if (owner)
    for (int j = tj; j < nd; j += kCols) A[ti * nd + j] = 0.0;
  __syncthreads();
  for (int q = tid; q < (int)g.n * 27; q += kDenseThreads) {
    const int I = q / 27, O = q - (q / 27) * 27;
    int x, y, z;
    lat_xyz(g, I, x, y, z);
    const int jx = x + O / 9 - 1, jy = y + (O / 3) % 3 - 1, jz = z + O % 3 - 1;
    if (jx < 0 || jx >= g.nx || jy < 0 || jy >= g.ny || jz < 0 || jz >= g.nz) continue;
    const int J = (int)lat_id(g, jx, jy, jz);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) A[(3 * I + a) * nd + 3 * J + b] = (double)S[(int64_t)(O * 9 + 3 * a + b) * g.n + I];
  }
  __syncthreads();
  // symmetrise: each owner averages its entries above the diagonal with their mirrors
  double sym[(kDenseMax + kCols - 1) / kCols];
#pragma unroll
  for (int m = 0; m < (kDenseMax + kCols - 1) / kCols; ++m) {
    const int j = tj + m * kCols;
    sym[m] = (owner && j < nd) ? 0.5 * (A[ti * nd + j] + A[j * nd + ti]) : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < (kDenseMax + kCols - 1) / kCols; ++m) {
    const int j = tj + m * kCols;
    if (owner && j < nd) A[ti * nd + j] = sym[m];
  }
  __syncthreads();
  // Gauss-Jordan with each owner's row segment in registers.  Step k publishes
  // column k (fc) and the unscaled pivot row (pr) in LDS, double-buffered by the
  // step's parity, so one barrier per step suffices: a buffer is rewritten two
  // steps later, after every thread has passed the barrier between.  Columns
  // past nd are zero padding up to kSeg * kCols and stay zero, so the step has
  // no per-column branches (its 8 pivot-row reads issue together).
  constexpr int kSeg = (kDenseMax + kCols - 1) / kCols;
  static_assert(kSeg * kCols <= kDenseMax, "padded pivot row exceeds its buffer");
  double R[kSeg];
#pragma unroll
  for (int m = 0; m < kSeg; ++m) {
    const int j = tj + m * kCols;
    R[m] = (owner && j < nd) ? A[ti * nd + j] : 0.0;
  }
  int dropped = 0;
  for (int k = 0; k < nd; ++k) {
    double* fc = col + (k & 1) * kDenseMax;
    double* pr = prow + (k & 1) * kDenseMax;
    const int kt = k % kCols, km = k / kCols;
    if (owner && tj == kt) {
      double v = R[0];
#pragma unroll
      for (int m = 1; m < kSeg; ++m) v = (m == km) ? R[m] : v;
      fc[ti] = v;
    }
    if (ti == k) {
#pragma unroll
      for (int m = 0; m < kSeg; ++m) pr[tj + m * kCols] = R[m];
    }
    __syncthreads();
    const double piv = fc[k];
    const double pinv = (piv > 0.0 && isfinite(piv)) ? 1.0 / piv : 0.0;
    dropped += (pinv == 0.0);
    const bool pivot_row = (ti == k);
    const double f = (owner && !pivot_row) ? fc[ti] * pinv : 0.0;
#pragma unroll
    for (int m = 0; m < kSeg; ++m) {
      const bool diag = (tj == kt) && (m == km);
      const double p = pr[tj + m * kCols];
      const double other = diag ? -f : R[m] - f * p;
      const double scaled = (diag ? 1.0 : R[m]) * pinv;
      R[m] = pivot_row ? scaled : other;
    }
  }
  if (owner) {
#pragma unroll
    for (int m = 0; m < kSeg; ++m) {
      const int j = tj + m * kCols;
      if (j < nd) Ainv[j * nd + ti] = R[m];  // column-major: a row's reads in the solve coalesce
    }
  }
  if (tid == 0) bad[1] = dropped;
}

// ---------------------------------------------------------------------------
// V-cycle kernels (sc: the outer solve's scalars; nullptr for a single apply)

// Smoothing is a polynomial in Dinv A applied to the level's residual, run as
// the three-term recurrence (Chebyshev; damped Jacobi is the case c1 = 0):
//   d_0 = s0 Dinv r_0,  d_j = c1_j d_{j-1} + c2_j Dinv r_j,  r_j = r_{j-1} - A d_{j-1},
//   e = sum_j d_j.
// Each step is one pass over the stencil (the product A d_{j-1} needs the
// previous direction at the neighbours, so directions alternate buffers).
// The per-node bodies below are shared by the grid kernels (large levels) and
// by mg_small_cycle (every level from the first small one down, one workgroup).

// pre-smoother from e = 0: d = s0 Dinv r, e = d
__device__ __forceinline__ void poly_first_node(const float* __restrict__ Dinv, int64_t n, float s0,
                                                const float* __restrict__ r, float* __restrict__ d,
                                                float* __restrict__ e, int64_t node) {
  const float v[3] = {r[3 * node], r[3 * node + 1], r[3 * node + 2]};
  float o[3];
  block_apply(Dinv, n, node, v, o);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    d[3 * node + i] = s0 * o[i];
    e[3 * node + i] = s0 * o[i];
  }
}

// res = res_in - A d_in;  d_out = c1 d_in + c2 Dinv res;  e += d_out
// (res_in may alias res_out: each node reads and writes only its own entries)
template <bool HALF, typename ST>
__device__ __forceinline__ void poly_step_node(const ST* __restrict__ S, const float* __restrict__ scl,
                                               const float* __restrict__ Dinv,
                                               const Lat& g, float c1, float c2, const float* res_in,
                                               const float* __restrict__ d_in, float* res_out,
                                               float* __restrict__ d_out, float* __restrict__ e, int64_t node) {
  int x, y, z;
  lat_xyz(g, node, x, y, z);
  float ad[3];
  stencil_row<HALF>(S, g, node, x, y, z, d_in, ad, scl);
  const float res[3] = {res_in[3 * node] - ad[0], res_in[3 * node + 1] - ad[1], res_in[3 * node + 2] - ad[2]};
  float dr[3];
  block_apply(Dinv, g.n, node, res, dr);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float dn = c1 * d_in[3 * node + i] + c2 * dr[i];
    res_out[3 * node + i] = res[i];
    d_out[3 * node + i] = dn;
    e[3 * node + i] += dn;
  }
}

// t = res_in - A d_in (the residual after the last pre-smoothing step); with a
// mask (the fine level, whose stencil may hold the Dirichlet rows) t = M (...):
// the restriction must not see those rows
template <bool HALF, typename ST>
__device__ __forceinline__ void poly_resid_node(const ST* __restrict__ S, const float* __restrict__ scl,
                                                const Lat& g,
                                                const float* __restrict__ res_in, const float* __restrict__ d_in,
                                                const float* __restrict__ mask, float* __restrict__ t, int64_t node) {
  int x, y, z;
  lat_xyz(g, node, x, y, z);
  float ad[3];
  stencil_row<HALF>(S, g, node, x, y, z, d_in, ad, scl);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float v = res_in[3 * node + i] - ad[i];
    t[3 * node + i] = mask ? v * mask[3 * node + i] : v;
  }
}

// post-smoother start: res = r - A e_in;  d = s0 Dinv res;  e_out = e_in + d
template <bool HALF, typename ST>
__device__ __forceinline__ void post_first_node(const ST* __restrict__ S, const float* __restrict__ scl,
                                                const float* __restrict__ Dinv,
                                                const Lat& g, float s0, const float* __restrict__ r,
                                                const float* __restrict__ e_in, float* __restrict__ res,
                                                float* __restrict__ d, float* __restrict__ e_out, int64_t node) {
  int x, y, z;
  lat_xyz(g, node, x, y, z);
  float ae[3];
  stencil_row<HALF>(S, g, node, x, y, z, e_in, ae, scl);
  const float rr[3] = {r[3 * node] - ae[0], r[3 * node + 1] - ae[1], r[3 * node + 2] - ae[2]};
  float dr[3];
  block_apply(Dinv, g.n, node, rr, dr);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    res[3 * node + i] = rr[i];
    d[3 * node + i] = s0 * dr[i];
    e_out[3 * node + i] = e_in[3 * node + i] + s0 * dr[i];
  }
}

// rc = P^T t at coarse node I; with Dc given, also the coarse pre-smoother's
// start dc = ec = s0 Dc rc
__device__ __forceinline__ void restrict_node(const float* __restrict__ t, const Lat& f, const Lat& c,
                                              float* __restrict__ rc, const float* __restrict__ Dc, float s0,
                                              float* __restrict__ dc, float* __restrict__ ec, int64_t I) {
  const int nf[3] = {f.nx - 1, f.ny - 1, f.nz - 1};
  int ci[3];
  lat_xyz(c, I, ci[0], ci[1], ci[2]);
  int fr[3][3];
  float wr[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int p = cpos(ci[d], nf[d]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int fidx = p + a - 1;
      const bool ok = fidx >= 0 && fidx <= nf[d];
      fr[d][a] = ok ? fidx : p;
      wr[d][a] = ok ? p1(fidx, ci[d], nf[d]) : 0.f;
    }
  }
  float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
#pragma unroll
    for (int ay = 0; ay < 3; ++ay)
#pragma unroll
      for (int az = 0; az < 3; ++az) {
        const float w = wr[0][ax] * wr[1][ay] * wr[2][az];
        const int64_t i = lat_id(f, fr[0][ax], fr[1][ay], fr[2][az]) * 3;
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += w * t[i + k];
      }
#pragma unroll
  for (int k = 0; k < 3; ++k) rc[3 * I + k] = acc[k];
  if (Dc) {
    float o[3];
    block_apply(Dc, c.n, I, acc, o);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      dc[3 * I + k] = s0 * o[k];
      ec[3 * I + k] = s0 * o[k];
    }
  }
}

// e += mask * P ec at fine node i
__device__ __forceinline__ void prolong_node(const float* __restrict__ ec, const Lat& f, const Lat& c,
                                             const float* __restrict__ mask, float* __restrict__ e, int64_t i) {
  const int nf[3] = {f.nx - 1, f.ny - 1, f.nz - 1};
  const int nc[3] = {c.nx - 1, c.ny - 1, c.nz - 1};
  int fi[3];
  lat_xyz(f, i, fi[0], fi[1], fi[2]);
  int pc[3][2];
  float pw[3][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int v = fi[d];
    if (v % 2 == 0) {
      pc[d][0] = v / 2; pw[d][0] = 1.f; pc[d][1] = v / 2; pw[d][1] = 0.f;
    } else if (v == nf[d]) {
      pc[d][0] = nc[d]; pw[d][0] = 1.f; pc[d][1] = nc[d]; pw[d][1] = 0.f;
    } else {
      pc[d][0] = (v - 1) / 2; pw[d][0] = 0.5f; pc[d][1] = (v + 1) / 2; pw[d][1] = 0.5f;
    }
  }
  float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const float w = pw[0][a] * pw[1][b] * pw[2][cc];
        const int64_t J = lat_id(c, pc[0][a], pc[1][b], pc[2][cc]) * 3;
#pragma unroll
        for (int k = 0; k < 3; ++k) acc[k] += w * ec[J + k];
      }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float m = mask ? mask[3 * i + k] : 1.f;
    e[3 * i + k] += m * acc[k];
  }
}

#define MG_NODE_LOOP(N, ...)                                                           \
  do {                                                                                 \
    int64_t i0_, i1_;                                                                  \
    node_range((N), i0_, i1_);                                                         \
    for (int64_t node = i0_ + threadIdx.x; node < i1_; node += kMgBlock) __VA_ARGS__;  \
  } while (0)

__global__ void __launch_bounds__(kMgBlock) mg_poly_first(const float* __restrict__ Dinv, int64_t n,
                                                          const float* __restrict__ cf, const float* __restrict__ r,
                                                          float* __restrict__ d, float* __restrict__ e,
                                                          const Scalars* sc) {
  if (sc && sc->done) return;
  const float s0 = cf[0];
  MG_NODE_LOOP(n, poly_first_node(Dinv, n, s0, r, d, e, node));
}

// The smoothing passes take the level's operator as an FP32 stencil (ST = float,
// scl null), on the fine level with HF_MG_FINE_F16 as FP16 words (ST = __half2,
// scl[1] the inverse of their scale), or as the matrix-free cell operator (ST =
// MfOp, scl the per-cell scales).

template <bool HALF, typename ST = float>
__global__ void __launch_bounds__(kMgBlock) mg_poly_step(const ST* __restrict__ S, const float* __restrict__ scl,
                                                         const float* __restrict__ Dinv,
                                                         Lat g, const float* __restrict__ cf, int j,
                                                         const float* res_in, const float* __restrict__ d_in,
                                                         float* res_out, float* __restrict__ d_out,
                                                         float* __restrict__ e, const Scalars* sc) {
  if (sc && sc->done) return;
  const float c1 = cf[1 + j], c2 = cf[1 + kMaxDegree + j];
  MG_NODE_LOOP(g.n, poly_step_node<HALF>(S, scl, Dinv, g, c1, c2, res_in, d_in, res_out, d_out, e, node));
}

template <bool HALF, typename ST = float>
__global__ void __launch_bounds__(kMgBlock) mg_poly_resid(const ST* __restrict__ S, const float* __restrict__ scl,
                                                          Lat g, const float* __restrict__ res_in,
                                                          const float* __restrict__ d_in,
                                                          const float* __restrict__ mask, float* __restrict__ t,
                                                          const Scalars* sc) {
  if (sc && sc->done) return;
  MG_NODE_LOOP(g.n, poly_resid_node<HALF>(S, scl, g, res_in, d_in, mask, t, node));
}

template <bool HALF, typename ST = float>
__global__ void __launch_bounds__(kMgBlock) mg_post_first(const ST* __restrict__ S, const float* __restrict__ scl,
                                                          const float* __restrict__ Dinv,
                                                          Lat g, const float* __restrict__ cf,
                                                          const float* __restrict__ r, const float* __restrict__ e_in,
                                                          float* __restrict__ res, float* __restrict__ d,
                                                          float* __restrict__ e_out, const Scalars* sc) {
  if (sc && sc->done) return;
  const float s0 = cf[0];
  MG_NODE_LOOP(g.n, post_first_node<HALF>(S, scl, Dinv, g, s0, r, e_in, res, d, e_out, node));
}

__global__ void __launch_bounds__(kMgBlock) mg_restrict(const float* __restrict__ t, Lat f, Lat c,
                                                        float* __restrict__ rc, const float* __restrict__ Dc,
                                                        const float* __restrict__ cf, float* __restrict__ dc,
                                                        float* __restrict__ ec, const Scalars* sc) {
  if (sc && sc->done) return;
  const float s0 = cf[0];
  MG_NODE_LOOP(c.n, restrict_node(t, f, c, rc, Dc, s0, dc, ec, node));
}

__global__ void __launch_bounds__(kMgBlock) mg_prolong(const float* __restrict__ ec, Lat f, Lat c,
                                                       const float* __restrict__ mask, float* __restrict__ e,
                                                       const Scalars* sc) {
  if (sc && sc->done) return;
  MG_NODE_LOOP(f.n, prolong_node(ec, f, c, mask, e, node));
}

// Gershgorin bound of Dinv A: per-workgroup max over nodes of the largest
// absolute row sum of the node's block row of Dinv A, a thread per node (the
// large levels: each node's block inverse is read once; mg_gersh_lanes below
// takes the small ones).
template <bool HALF>
__global__ void __launch_bounds__(kMgBlock) mg_gersh_nodes(const float* __restrict__ S, const float* __restrict__ Dinv, Lat g,
                                                     const float* __restrict__ mask, float* __restrict__ part) {
  __shared__ float sm[kMgBlock / kWave];
  const int64_t n = g.n;
  float best = 0.f;
  for (int64_t node = (int64_t)blockIdx.x * kMgBlock + threadIdx.x; node < n; node += (int64_t)gridDim.x * kMgBlock) {
    float M[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) M[e] = Dinv[(int64_t)e * n + node];
    int x, y, z;
    lat_xyz(g, node, x, y, z);
    float rs[3] = {0.f, 0.f, 0.f};
    for (int t = 0; t < 27; ++t) {
      const int o = pair_slot(t);
      const int xr = x + o / 9 - 1, yr = y + (o / 3) % 3 - 1, zr = z + o % 3 - 1;
      const bool valid = xr >= 0 && xr < g.nx && yr >= 0 && yr < g.ny && zr >= 0 && zr < g.nz;
      const int64_t m = lat_id(g, clampi(xr, 0, g.nx - 1), clampi(yr, 0, g.ny - 1), clampi(zr, 0, g.nz - 1));
      float B[9];
      load_block<HALF>(S, n, node, o, m, valid, B);
      // columns of Dirichlet DOFs (M A M; the rows are zero in Dinv already)
      float cm[3] = {1.f, 1.f, 1.f};
      if (mask) {
#pragma unroll
        for (int j = 0; j < 3; ++j) cm[j] = mask[3 * m + j];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          rs[i] += fabsf(M[3 * i] * B[j] + M[3 * i + 1] * B[3 + j] + M[3 * i + 2] * B[6 + j]) * cm[j];
    }
    best = fmaxf(best, fmaxf(rs[0], fmaxf(rs[1], rs[2])));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o, 64));
  if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = sm[0];
    for (int k = 1; k < kMgBlock / kWave; ++k) m = fmaxf(m, sm[k]);
    part[blockIdx.x] = m;
  }
}

__global__ void __launch_bounds__(kMgBlock) mg_max_final(const float* __restrict__ part, int np,
                                                         float* __restrict__ out) {
  __shared__ float sm[kMgBlock / kWave];
  float best = 0.f;
  for (int k = threadIdx.x; k < np; k += kMgBlock) best = fmaxf(best, part[k]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o, 64));
  if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = sm[0];
    for (int k = 1; k < kMgBlock / kWave; ++k) m = fmaxf(m, sm[k]);
    out[0] = m;
  }
}

// FP16 copy of the fine stencil (HF_MG_FINE_F16), in three launches on the
// set-up stream: per-workgroup maxima of |S|, then (mg_max_final into
// scl[2]) the scale, then the conversion.
__global__ void __launch_bounds__(kMgBlock) mg_absmax(const float* __restrict__ S, int64_t count,
                                                      float* __restrict__ part) {
  __shared__ float sm[kMgBlock / kWave];
  float best = 0.f;
  const int64_t n4 = count / 4;
  const float4* S4 = reinterpret_cast<const float4*>(S);
  for (int64_t i = (int64_t)blockIdx.x * kMgBlock + threadIdx.x; i < n4; i += (int64_t)gridDim.x * kMgBlock) {
    const float4 v = S4[i];
    best = fmaxf(best, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
  }
  if (blockIdx.x == 0)
    for (int64_t i = 4 * n4 + threadIdx.x; i < count; i += kMgBlock) best = fmaxf(best, fabsf(S[i]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o, 64));
  if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = sm[0];
    for (int k = 1; k < kMgBlock / kWave; ++k) m = fmaxf(m, sm[k]);
    part[blockIdx.x] = m;
  }
}

// scl[0] = 2^(14 - floor(log2 max|S|)) (1 for a zero or non-finite maximum),
// scl[1] = 1 / scl[0]; scl[2] holds the maximum
__global__ void mg_f16_scale(float* __restrict__ scl) {
  if (threadIdx.x != 0) return;
  const float m = scl[2];
  float sc = 1.f;
  if (m > 0.f && isfinite(m)) sc = ldexpf(1.f, 14 - ilogbf(m));
  scl[0] = sc;
  scl[1] = 1.f / sc;
}


// coarsest level: e = Ainv r (FP64 accumulation), one workgroup
// e = Ainv r for the dense coarsest level (Ainv column-major FP64, nd <= kDenseMax)
// by one workgroup of T threads: thread t takes row t % nd and every ng-th
// column from t / nd (ng = T / nd groups), so its few loads are in flight
// together instead of nd dependent ones; the groups' partial sums meet in LDS
// and are added in a fixed order.
template <int T>
__device__ __forceinline__ void dense_apply_wg(const double* __restrict__ Ainv, int nd, const float* __restrict__ r,
                                               float* __restrict__ e) {
  static_assert(T >= kDenseMax, "one row per thread at least");
  __shared__ double rs[kDenseMax];
  __shared__ double part[T];
  const int tid = threadIdx.x;
  if (tid < nd) rs[tid] = (double)r[tid];
  __syncthreads();
  const int ng = T / nd, i = tid % nd, jg = tid / nd;
  double acc = 0.0;
  if (jg < ng)
    for (int j = jg; j < nd; j += ng) acc += Ainv[(int64_t)j * nd + i] * rs[j];
  part[tid] = acc;
  __syncthreads();
  if (tid < nd) {
    double v = 0.0;
    for (int k = 0; k < ng; ++k) v += part[k * nd + tid];
    e[tid] = (float)v;
  }
}

constexpr int kDenseApplyThreads = 1024;
__global__ void __launch_bounds__(kDenseApplyThreads) mg_dense_apply(const double* __restrict__ Ainv, int nd,
                                                                     const float* __restrict__ r,
                                                                     float* __restrict__ e, const Scalars* sc) {
  if (sc && sc->done) return;
  dense_apply_wg<kDenseApplyThreads>(Ainv, nd, r, e);
}

// The levels from the first small one down run in one workgroup
// (mg_small_cycle): the whole sub-cycle with workgroup barriers between its
// phases and every level's data in LDS.  The stencils (in the form the level
// holds them), the block inverses and the coefficients are read from global
// memory once per cycle, in one round of loads that are all in flight
// together; a phase is then a round of LDS reads, not of dependent global
// loads.  A stencil product uses 8 lanes per node (slots o = q, q+8, ... on lane
// q, summed with shuffles).  The first level's right-hand side and pre-smoother
// start were written by the restriction above it; its result goes to its e2
// buffer (or, with the dense level alone, to its e).
constexpr int kSmallThreads = 1024;
constexpr int kGroup = 8;                             // lanes per node
constexpr int kSmallNodes = kSmallThreads / kGroup;   // 128 nodes a round
// the cycle takes in levels of at most kSmallStart nodes, at most kSmallMax of
// them (the dense one included), while their LDS arena fits kSmallArenaWords
constexpr int kSmallStart = 512;
constexpr int kSmallMax = 5;
constexpr int kSmallUnroll = 1;
// 148 KB of arena + dense_apply_wg's 8.7 KB of LDS: within gfx950's 160 KB per workgroup
constexpr int kSmallArenaWords = 148 * 1024 / 4;
// stencil forms of a level in the small cycle
constexpr int kSmallF16 = 0;     // FP16 words: 5 half2 per half-stored block (the grid levels' S16)
constexpr int kSmallHalf = 1;    // FP32, diagonal and 13 upper slots
constexpr int kSmallFull = 2;    // FP32, 27 slots
struct SmallLevel {
  Lat g;
  int kind;          // stencil form (smoothed levels)
  int words;         // its 4-byte words (0 for the dense level)
  const void* S;     // the stencil (global)
  const float* scl;  // FP16 form: the scale pair (scl[1] multiplies the products)
  const float* Dinv; // block inverses (global, smoothed levels)
  const float* cf;   // smoothing coefficients (global)
  int oS, oD, oC, oV;  // arena offsets (words): stencil, block inverses, node coordinates, vectors
};
struct SmallCycle {
  int nlev;     // levels here, the last one dense
  int degree;
  const double* Ainv;
  int nd;
  const float *r0, *e0, *d0;   // first level: right-hand side, pre-smoother start (e = d)
  float* out;                  // first level's result
  SmallLevel lv[kSmallMax];
};
// Arena words of one small-cycle level.  Vectors are float4 per node (one
// 16-byte LDS read per neighbour); a smoothed level holds its stencil, its block
// inverses, its node coordinates (one packed word per node) and six vectors (r,
// two e, two d, res); the dense one its coordinates, r and e.
// (+3 per array: each starts on a 16-byte boundary)
__host__ __device__ inline int64_t small_level_words(int kind, int64_t n, bool dense) {
  if (dense) return 9 * n + 6;
  const int64_t st = kind == kSmallF16 ? (int64_t)kHalfSlots * kF16Words * n
                     : kind == kSmallHalf ? (int64_t)kHalfSlots * 9 * n : (int64_t)kSt * n;
  return st + 9 * n + n + 24 * n + 12;
}

// product of node n's stencil row with v, by the 8 lanes of a group (q = lane
// in group); every lane of the group ends with the whole sum.  UNROLL (1, 2 or
// 4) is how many of a lane's 3-4 blocks have their loads in flight together:
// 4 in the grid kernels; in the 1024-thread small cycle (128 VGPRs a lane) all
// four spill.
template <bool HALF, int UNROLL, typename ST = float>
__device__ __forceinline__ void stencil_row_group(const ST* __restrict__ S, const Lat& g, int64_t n,
                                                  const float* __restrict__ v, int q, float (&out)[3],
                                                  const float* __restrict__ scl = nullptr) {
  int x, y, z;
  lat_xyz(g, n, x, y, z);
  if constexpr (std::is_same<ST, MfOp>::value) {
    // lane q takes cell q of the node (its 8 corners' values), as mf_row
    const int i = q >> 2, j = (q >> 1) & 1, k = q & 1;
    const int ncx = g.nx - 1, ncy = g.ny - 1, ncz = g.nz - 1;
    const int px = x - 1 + i, py = y - 1 + j, pz = z - 1 + k;
    const bool in = q < 8 && px >= 0 && px < ncx && py >= 0 && py < ncy && pz >= 0 && pz < ncz;
    const float sc = in ? scl[((int64_t)px * ncy + py) * ncz + pz] : 0.f;
    const int qa = 7 - (q & 7);
    float t[3] = {0.f, 0.f, 0.f};
    const float w0 = v[3 * n], w1 = v[3 * n + 1], w2 = v[3 * n + 2];
#pragma unroll
    for (int qb = 0; qb < 8; ++qb) {
      const int64_t m = lat_id(g, clampi(px + (qb >> 2), 0, g.nx - 1), clampi(py + ((qb >> 1) & 1), 0, g.ny - 1),
                               clampi(pz + (qb & 1), 0, g.nz - 1));
      // differences from the node's own value (the diagonal block implied, as mf_row)
      const float uu[3] = {v[3 * m] - w0, v[3 * m + 1] - w1, v[3 * m + 2] - w2};
      const float* B = S->b + (qa ^ qb) * 6;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) t[r] += (mf_sign(qa, qb, r, cc) * B[mf_tri(r, cc)]) * uu[cc];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) out[r] = sc * t[r];
  } else {
    out[0] = out[1] = out[2] = 0.f;
#pragma unroll UNROLL
    for (int t = 0; t < (27 + kGroup - 1) / kGroup; ++t) {
      const int o = q + t * kGroup;
      if (o >= 27) break;
      const int xr = x + o / 9 - 1, yr = y + (o / 3) % 3 - 1, zr = z + o % 3 - 1;
      const int xx = clampi(xr, 0, g.nx - 1), yy = clampi(yr, 0, g.ny - 1), zz = clampi(zr, 0, g.nz - 1);
      const int64_t m = lat_id(g, xx, yy, zz);
      const float v0 = v[3 * m], v1 = v[3 * m + 1], v2 = v[3 * m + 2];
      float B[9];
      load_block<HALF>(S, g.n, n, o, m, xr == xx && yr == yy && zr == zz, B);
#pragma unroll
      for (int i = 0; i < 3; ++i) out[i] += B[3 * i] * v0 + B[3 * i + 1] * v1 + B[3 * i + 2] * v2;
    }
  }
#pragma unroll
  for (int w = 1; w < kGroup; w <<= 1)
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] += __shfl_xor(out[i], w, kWave);
  if constexpr (is_f16<ST>()) {
    const float sinv = scl[1];
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] *= sinv;
  }
}

#define MG_GROUP_LOOP(N, ...)                                                                           \
  do {                                                                                                  \
    const int q = threadIdx.x % kGroup;                                                                 \
    for (int64_t node = threadIdx.x / kGroup; node < (N); node += kSmallThreads / kGroup) __VA_ARGS__;  \
    __syncthreads();                                                                                    \
  } while (0)
#define MG_WG_LOOP(N, ...)                                                                   \
  do {                                                                                       \
    for (int64_t node = threadIdx.x; node < (N); node += kSmallThreads) __VA_ARGS__;         \
    __syncthreads();                                                                         \
  } while (0)

// n 4-byte words from global memory into LDS (both 16-byte aligned): 16-byte
// loads, four in flight per thread, then the tail word by word
__device__ __forceinline__ void lds_fill(float* __restrict__ dst, const float* __restrict__ src, int n) {
  constexpr int U = 4;
  const int n4 = n >> 2;
  const float4* s4 = reinterpret_cast<const float4*>(src);
  float4* d4 = reinterpret_cast<float4*>(dst);
  for (int base = threadIdx.x; base < n4; base += U * kSmallThreads) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kSmallThreads;
      v[u] = i < n4 ? s4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kSmallThreads;
      if (i < n4) d4[i] = v[u];
    }
  }
  const int i = 4 * n4 + (int)threadIdx.x;
  if (i < n) dst[i] = src[i];
}
// a 3-vector of n nodes (global, 3 floats a node) into LDS float4 slots
__device__ __forceinline__ void lds_fill3(float* __restrict__ dst, const float* __restrict__ src, int n) {
  constexpr int U = 4;
  const int n3 = 3 * n;
  for (int base = threadIdx.x; base < n3; base += U * kSmallThreads) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kSmallThreads;
      v[u] = i < n3 ? src[i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * kSmallThreads;
      if (i < n3) dst[4 * (i / 3) + i % 3] = v[u];
    }
  }
}

// sum over the 8 lanes of a group, every lane ending with it: quad swaps by DPP
// (lanes 1,0,3,2 then 2,3,0,1), then each lane adds the lane mirrored in its half
// row (i <-> 7 - i), which holds the other quad's sum
__device__ __forceinline__ float group8_sum(float v) {
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
  v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x141, 0xF, 0xF, true));
  return v;
}

__device__ __forceinline__ void small_xyz(uint32_t w, int& x, int& y, int& z) {
  x = (int)(w & 1023u);
  y = (int)((w >> 10) & 1023u);
  z = (int)(w >> 20);
}

// Node n's stencil row (level V, stencil St in LDS, form KIND) times the float4
// vector v, by the 8 lanes of a group: lane q takes slots q, q + 8, q + 16, q + 24.
// Slots below 13 of a half-stored stencil are the neighbour's upper slot 13 - o
// transposed; the slot of a neighbour outside the lattice is skipped (its
// stored block is zero, or it is a lower block the node does not have).
template <int KIND>
__device__ __forceinline__ void small_row(const float* __restrict__ St, int N, const Lat& g, int x, int y, int z,
                                          int node, const float4* __restrict__ v, int q, float sinv,
                                          float (&out)[3]) {
  float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int o = q + 8 * t;
    if (o >= 27) break;
    const int ox = o / 9, oy = (o / 3) % 3, oz = o % 3;
    const int xr = x + ox - 1, yr = y + oy - 1, zr = z + oz - 1;
    if (xr < 0 || xr >= g.nx || yr < 0 || yr >= g.ny || zr < 0 || zr >= g.nz) continue;
    const int m = (xr * g.ny + yr) * g.nz + zr;
    const float4 w = v[m];
    const float wv[3] = {w.x, w.y, w.z};
    float T[10];
    if constexpr (KIND == kSmallFull) {
#pragma unroll
      for (int e = 0; e < 9; ++e) T[e] = St[(o * 9 + e) * N + node];
#pragma unroll
      for (int i = 0; i < 3; ++i) acc[i] += T[3 * i] * wv[0] + T[3 * i + 1] * wv[1] + T[3 * i + 2] * wv[2];
    } else {
      // slots 0..7 are lower, 16.. upper; 8..15 either (a lane-dependent branch)
      const bool own = t >= 2 || (t == 1 && o >= 13);
      const int sl = own ? o - 13 : 13 - o;
      const int at = own ? node : m;
      if constexpr (KIND == kSmallF16) {
        const __half2* W = reinterpret_cast<const __half2*>(St);
#pragma unroll
        for (int k = 0; k < kF16Words; ++k) {
          const float2 f = __half22float2(W[(sl * kF16Words + k) * N + at]);
          T[2 * k] = f.x;
          T[2 * k + 1] = f.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 9; ++e) T[e] = St[(sl * 9 + e) * N + at];
      }
      if (own) {
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[i] += T[3 * i] * wv[0] + T[3 * i + 1] * wv[1] + T[3 * i + 2] * wv[2];
      } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[i] += T[i] * wv[0] + T[3 + i] * wv[1] + T[6 + i] * wv[2];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) out[i] = group8_sum(acc[i]) * sinv;
}

// rc = P^T t at coarse node I (float4 vectors in LDS); with Dc given, also the
// coarse pre-smoother's start dc = ec = s0 Dc rc (restrict_node's transfer)
__device__ __forceinline__ void small_restrict(const float4* __restrict__ t, const Lat& f, const Lat& c,
                                               const uint32_t* __restrict__ cxyz, float4* __restrict__ rc,
                                               const float* __restrict__ Dc, float s0, float4* __restrict__ dc,
                                               float4* __restrict__ ec, int I) {
  const int nf[3] = {f.nx - 1, f.ny - 1, f.nz - 1};
  int ci[3];
  small_xyz(cxyz[I], ci[0], ci[1], ci[2]);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  int fr[3][3];
  float wr[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int p = cpos(ci[d], nf[d]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int fidx = p + a - 1;
      const bool ok = fidx >= 0 && fidx <= nf[d];
      fr[d][a] = ok ? fidx : p;
      wr[d][a] = ok ? p1(fidx, ci[d], nf[d]) : 0.f;
    }
  }
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
#pragma unroll
    for (int ay = 0; ay < 3; ++ay)
#pragma unroll
      for (int az = 0; az < 3; ++az) {
        const float w = wr[0][ax] * wr[1][ay] * wr[2][az];
        const float4 tv = t[(fr[0][ax] * f.ny + fr[1][ay]) * f.nz + fr[2][az]];
        acc.x += w * tv.x;
        acc.y += w * tv.y;
        acc.z += w * tv.z;
      }
  rc[I] = acc;
  if (Dc) {
    const int n = (int)c.n;
    const float a3[3] = {acc.x, acc.y, acc.z};
    float o[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      o[i] = s0 * (Dc[(3 * i) * n + I] * a3[0] + Dc[(3 * i + 1) * n + I] * a3[1] + Dc[(3 * i + 2) * n + I] * a3[2]);
    const float4 ov = make_float4(o[0], o[1], o[2], 0.f);
    dc[I] = ov;
    ec[I] = ov;
  }
}

// e += P ec at fine node i (float4 vectors in LDS; prolong_node's transfer)
__device__ __forceinline__ void small_prolong(const float4* __restrict__ ec, const Lat& f, const Lat& c,
                                              const uint32_t* __restrict__ fxyz, float4* __restrict__ e, int i) {
  const int nf[3] = {f.nx - 1, f.ny - 1, f.nz - 1};
  const int nc[3] = {c.nx - 1, c.ny - 1, c.nz - 1};
  int fi[3];
  small_xyz(fxyz[i], fi[0], fi[1], fi[2]);
  int pc[3][2];
  float pw[3][2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int v = fi[d];
    if (v % 2 == 0) {
      pc[d][0] = v / 2; pw[d][0] = 1.f; pc[d][1] = v / 2; pw[d][1] = 0.f;
    } else if (v == nf[d]) {
      pc[d][0] = nc[d]; pw[d][0] = 1.f; pc[d][1] = nc[d]; pw[d][1] = 0.f;
    } else {
      pc[d][0] = (v - 1) / 2; pw[d][0] = 0.5f; pc[d][1] = (v + 1) / 2; pw[d][1] = 0.5f;
    }
  }
  float4 acc = e[i];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const float w = pw[0][a] * pw[1][b] * pw[2][cc];
        const float4 v = ec[(pc[0][a] * c.ny + pc[1][b]) * c.nz + pc[2][cc]];
        acc.x += w * v.x;
        acc.y += w * v.y;
        acc.z += w * v.z;
      }
  e[i] = acc;
}

// One smoothing phase of a small-cycle level over its nodes, 8 lanes a node; after
// the row product ad = A din, lane i < 3 of the group finishes component i.
//   MODE 0 (a Chebyshev step):  rr = res - ad, dn = c1 din + c2 D rr, res_out = rr,
//                               dout = dn, e += dn
//   MODE 1 (the residual):      dout = res - ad
//   MODE 2 (post start, din = e): rr = r - ad, res_out = rr, dout = s0 D rr,
//                               e_out = e + s0 D rr
template <int KIND, int MODE>
__device__ __forceinline__ void small_phase(const SmallLevel& V, const float* __restrict__ St,
                                            const float* __restrict__ D, const uint32_t* __restrict__ xyz,
                                            float sinv, const float4* res, const float4* din, float4* res_out,
                                            float4* dout, float4* e, const float4* ein, float c1, float c2) {
  const int N = (int)V.g.n;
  const int q = threadIdx.x % kGroup;
  for (int node = threadIdx.x / kGroup; node < N; node += kSmallNodes) {
    int x, y, z;
    small_xyz(xyz[node], x, y, z);
    float ad[3];
    small_row<KIND>(St, N, V.g, x, y, z, node, din, q, sinv, ad);
    if (q < 3) {
      const float4 rv = res[node];
      const float rr[3] = {rv.x - ad[0], rv.y - ad[1], rv.z - ad[2]};
      float* dv = reinterpret_cast<float*>(dout + node);
      if (MODE == 1) {
        dv[q] = rr[q];
      } else {
        const float dr = D[(3 * q) * N + node] * rr[0] + D[(3 * q + 1) * N + node] * rr[1] +
                         D[(3 * q + 2) * N + node] * rr[2];
        float* ev = reinterpret_cast<float*>(e + node);
        reinterpret_cast<float*>(res_out + node)[q] = rr[q];
        if (MODE == 0) {
          const float dn = c1 * reinterpret_cast<const float*>(din + node)[q] + c2 * dr;
          dv[q] = dn;
          ev[q] += dn;
        } else {
          dv[q] = c1 * dr;
          ev[q] = reinterpret_cast<const float*>(ein + node)[q] + c1 * dr;
        }
      }
    }
  }
  __syncthreads();
}

template <int MODE>
__device__ __forceinline__ void small_phase_any(const SmallLevel& V, const float* A, float sinv, const float4* res,
                                                const float4* din, float4* res_out, float4* dout, float4* e,
                                                const float4* ein, float c1, float c2) {
  const float* St = A + V.oS;
  const float* D = A + V.oD;
  const uint32_t* xyz = reinterpret_cast<const uint32_t*>(A + V.oC);
  if (V.kind == kSmallF16)
    small_phase<kSmallF16, MODE>(V, St, D, xyz, sinv, res, din, res_out, dout, e, ein, c1, c2);
  else if (V.kind == kSmallHalf)
    small_phase<kSmallHalf, MODE>(V, St, D, xyz, 1.f, res, din, res_out, dout, e, ein, c1, c2);
  else
    small_phase<kSmallFull, MODE>(V, St, D, xyz, 1.f, res, din, res_out, dout, e, ein, c1, c2);
}

// e = Ainv r for the dense level (dense_apply_wg's scheme), r and e read and
// written through the index maps ri(i) and ei(i) (global stride 3 or LDS float4)
template <typename RI, typename EI>
__device__ __forceinline__ void small_dense(const double* __restrict__ Ainv, int nd, RI ri, EI ei, double* rs,
                                            double* part) {
  const int tid = threadIdx.x;
  if (tid < nd) rs[tid] = (double)ri(tid);
  __syncthreads();
  const int ng = kSmallThreads / nd, i = tid % nd, jg = tid / nd;
  double acc = 0.0;
  if (jg < ng)
    for (int j = jg; j < nd; j += ng) acc += Ainv[(int64_t)j * nd + i] * rs[j];
  part[tid] = acc;
  __syncthreads();
  if (tid < nd) {
    double v = 0.0;
    for (int kk = 0; kk < ng; ++kk) v += part[kk * nd + tid];
    ei(tid) = (float)v;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kSmallThreads) mg_small_cycle(SmallCycle sc_, const Scalars* sc) {
  if (sc && sc->done) return;
  const int L = sc_.nlev, k = sc_.degree;
  __shared__ __align__(16) float A[kSmallArenaWords];
  __shared__ float cfl[kSmallMax][kCoefStride];
  __shared__ float sinv[kSmallMax];
  __shared__ double rs[kDenseMax];
  __shared__ double part[kSmallThreads];
  // the level descriptors in LDS (indexed by level at run time; the kernel
  // argument is read with constant indices only)
  __shared__ SmallLevel slv[kSmallMax];
#pragma unroll
  for (int l = 0; l < kSmallMax; ++l)
    if (threadIdx.x == l) slv[l] = sc_.lv[l];
  __syncthreads();
  if (L == 1) {   // the dense level alone
    const float* r0 = sc_.r0;
    float* out = sc_.out;
    small_dense(sc_.Ainv, sc_.nd, [&](int i) { return r0[i]; }, [&](int i) -> float& { return out[i]; }, rs, part);
    return;
  }
  // stage: every smoothed level's stencil, block inverses and coefficients, every
  // level's node coordinates, the first level's r, e and d starts
  for (int l = 0; l < L; ++l) {
    const SmallLevel& V = slv[l];
    if (l + 1 < L) {
      lds_fill(A + V.oS, static_cast<const float*>(V.S), V.words);
      lds_fill(A + V.oD, V.Dinv, (int)(9 * V.g.n));
      if (threadIdx.x < kCoefStride) cfl[l][threadIdx.x] = V.cf[threadIdx.x];
      if (threadIdx.x == 0) sinv[l] = V.kind == kSmallF16 ? V.scl[1] : 1.f;
    }
    uint32_t* xyz = reinterpret_cast<uint32_t*>(A + V.oC);
    for (int n = threadIdx.x; n < (int)V.g.n; n += kSmallThreads) {
      const int z = n % V.g.nz, qq = n / V.g.nz;
      xyz[n] = (uint32_t)(qq / V.g.ny) | ((uint32_t)(qq % V.g.ny) << 10) | ((uint32_t)z << 20);
    }
  }
  {
    const SmallLevel& F = slv[0];
    const int n = (int)F.g.n;
    float* v0 = A + F.oV;
    lds_fill3(v0, sc_.r0, n);            // r
    lds_fill3(v0 + 4 * n, sc_.e0, n);    // e (the start)
    lds_fill3(v0 + 12 * n, sc_.d0, n);   // d (the start)
  }
  __syncthreads();
  // level l's vectors: r, e (two buffers), d (two buffers), res; the dense level r, e
  auto vec = [&](int l, int i) { return reinterpret_cast<float4*>(A + slv[l].oV) + i * slv[l].g.n; };
  int ecur[kSmallMax];   // which e buffer (1 or 2) holds each level's current e
#pragma unroll
  for (int l = 0; l < kSmallMax; ++l) ecur[l] = 1;
  for (int l = 0; l + 1 < L; ++l) {
    const SmallLevel& V = slv[l];
    const float* cf = cfl[l];
    const float4* res = vec(l, 0);
    float4* din = vec(l, 3);
    float4* dout = vec(l, 4);
    float4* e = vec(l, 1);
    for (int j = 1; j < k; ++j) {
      small_phase_any<0>(V, A, sinv[l], res, din, vec(l, 5), dout, e, nullptr, cf[1 + j], cf[1 + kMaxDegree + j]);
      res = vec(l, 5);
      float4* tmp = din; din = dout; dout = tmp;
    }
    small_phase_any<1>(V, A, sinv[l], res, din, nullptr, dout, nullptr, nullptr, 0.f, 0.f);
    const SmallLevel& C = slv[l + 1];
    const bool smooth_c = (l + 2 < L);
    const uint32_t* cxyz = reinterpret_cast<const uint32_t*>(A + C.oC);
    for (int I = threadIdx.x; I < (int)C.g.n; I += kSmallThreads)
      small_restrict(dout, V.g, C.g, cxyz, vec(l + 1, 0), smooth_c ? A + C.oD : nullptr, smooth_c ? cfl[l + 1][0] : 0.f,
                     vec(l + 1, 3), vec(l + 1, 1), I);
    __syncthreads();
  }
  {
    // dense level: e = Ainv r (FP64 accumulation) on its float4 slots
    const float* r = reinterpret_cast<const float*>(vec(L - 1, 0));
    float* e = reinterpret_cast<float*>(vec(L - 1, 1));
    small_dense(sc_.Ainv, sc_.nd, [&](int i) { return r[4 * (i / 3) + i % 3]; },
                [&](int i) -> float& { return e[4 * (i / 3) + i % 3]; }, rs, part);
  }
  for (int l = L - 2; l >= 0; --l) {
    const SmallLevel& V = slv[l];
    const SmallLevel& C = slv[l + 1];
    const float* cf = cfl[l];
    float4* ein = vec(l, ecur[l]);
    float4* eout = vec(l, 3 - ecur[l]);
    const float4* ec = vec(l + 1, ecur[l + 1]);
    const uint32_t* fxyz = reinterpret_cast<const uint32_t*>(A + V.oC);
    for (int i = threadIdx.x; i < (int)V.g.n; i += kSmallThreads) small_prolong(ec, V.g, C.g, fxyz, ein, i);
    __syncthreads();
    small_phase_any<2>(V, A, sinv[l], vec(l, 0), ein, vec(l, 5), vec(l, 3), eout, ein, cf[0], 0.f);
    ecur[l] = 3 - ecur[l];
    float4* din = vec(l, 3);
    float4* dout = vec(l, 4);
    for (int j = 1; j < k; ++j) {
      small_phase_any<0>(V, A, sinv[l], vec(l, 5), din, vec(l, 5), dout, eout, nullptr, cf[1 + j],
                         cf[1 + kMaxDegree + j]);
      float4* t2 = din; din = dout; dout = t2;
    }
  }
  // the first level's result back to global memory
  const float* e0 = reinterpret_cast<const float*>(vec(0, ecur[0]));
  const int n3 = (int)(3 * slv[0].g.n);
  for (int i = threadIdx.x; i < n3; i += kSmallThreads) sc_.out[i] = e0[4 * (i / 3) + i % 3];
}

// Levels of up to kLaneNodes nodes above the small cycle run the stencil passes
// with kGroup lanes per node as well: a thread then issues 3-4 block loads
// instead of 27 dependent ones, and a level of 33^3 nodes launches 8x the
// workgroups (one lane per node leaves half the SIMDs of the chip idle there).
// After the shuffle reduction lane i < 3 finishes component i of the node.
constexpr int64_t kLaneNodes = 65536;

#define MG_LANE_LOOP(N, ...)                                                                   \
  do {                                                                                         \
    const int q = threadIdx.x % kGroup;                                                        \
    constexpr int per_ = kMgBlock / kGroup;                                                    \
    for (int64_t node = (int64_t)logical_block() * per_ + threadIdx.x / kGroup; node < (N);   \
         node += (int64_t)gridDim.x * per_)                                                    \
      __VA_ARGS__;                                                                             \
  } while (0)

inline int lane_grid(int64_t n) {
  int64_t g = ceil_div(n * kGroup, kMgBlock);
  if (g >= 64) g = ceil_div(g, kNumXcd) * kNumXcd;  // whole XCDs, rounded up: no workgroup gets a second round
  if (g > kMaxGrid) g = kMaxGrid;
  return (int)(g < 1 ? 1 : g);
}

// block-wide max of v (non-negative) into the workgroup's partial
__device__ __forceinline__ void block_max_to(float v, float* __restrict__ part) {
  __shared__ float sm[kMgBlock / kWave];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = sm[0];
    for (int k = 1; k < kMgBlock / kWave; ++k) m = fmaxf(m, sm[k]);
    *part = m;
  }
}

// Gershgorin bound of Dinv A on a level of at most kLaneNodes nodes (as
// mg_gersh_nodes; Dirichlet columns of the fine level masked: M A M, its rows
// are zero in Dinv already).  8 lanes a node,
// each summing the |(Dinv B)_ij| of its 3-4 slots' blocks, the group then adding
// its lanes' sums: the loads of a node's 27 blocks are in flight together.
template <bool HALF>
__global__ void __launch_bounds__(kMgBlock) mg_gersh_lanes(const float* __restrict__ S, const float* __restrict__ Dinv,
                                                           Lat g, const float* __restrict__ mask,
                                                           float* __restrict__ part) {
  const int64_t n = g.n;
  float best = 0.f;
  const int q = threadIdx.x % kGroup;
  constexpr int per = kMgBlock / kGroup;
  for (int64_t node = (int64_t)logical_block() * per + threadIdx.x / kGroup; node < n;
       node += (int64_t)gridDim.x * per) {
    float M[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) M[e] = Dinv[(int64_t)e * n + node];
    int x, y, z;
    lat_xyz32(g, (uint32_t)node, x, y, z);
    float rs[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int o = q + 8 * t;
      if (o >= 27) break;
      const int xr = x + o / 9 - 1, yr = y + (o / 3) % 3 - 1, zr = z + o % 3 - 1;
      const bool valid = xr >= 0 && xr < g.nx && yr >= 0 && yr < g.ny && zr >= 0 && zr < g.nz;
      const int64_t m = lat_id(g, clampi(xr, 0, g.nx - 1), clampi(yr, 0, g.ny - 1), clampi(zr, 0, g.nz - 1));
      float B[9];
      load_block<HALF>(S, n, node, o, m, valid, B);
      float cm[3] = {1.f, 1.f, 1.f};
      if (mask) {
#pragma unroll
        for (int j = 0; j < 3; ++j) cm[j] = mask[3 * m + j];
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          rs[i] += fabsf(M[3 * i] * B[j] + M[3 * i + 1] * B[3 + j] + M[3 * i + 2] * B[6 + j]) * cm[j];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) rs[i] = group8_sum(rs[i]);
    best = fmaxf(best, fmaxf(rs[0], fmaxf(rs[1], rs[2])));
  }
  block_max_to(best, part + blockIdx.x);
}

// FP16 copies of several half-stored levels in two launches: per-workgroup
// maxima of |S| (mg_f16_max_all), then each workgroup of mg_f16_convert_all
// reduces its level's maxima itself, forms the scale (mg_f16_scale's rule) and
// converts its nodes' slots: W[(s*5 + k)*n + node] = half2(sc S[(s*9 + 2k)*n +
// node], sc S[(s*9 + 2k + 1)*n + node]), the fifth word's upper half zero.
struct F16Levels {
  int nlev;
  const float* S[kMgMaxLevels];   // FP32 half-stored stencils
  __half2* W[kMgMaxLevels];       // their FP16 words
  float* scl[kMgMaxLevels];       // scale pairs (scl[2]: the maximum)
  int64_t n[kMgMaxLevels];        // nodes
  int p0[kMgMaxLevels + 1];       // first partial (workgroup of mg_f16_max_all) of each level
  int c0[kMgMaxLevels + 1];       // first workgroup of mg_f16_convert_all of each level
};
constexpr int kF16MaxParts = 256;   // mg_f16_max_all workgroups per level at most

__device__ __forceinline__ int f16_level_of(const int* first, int nlev, int b) {
  int l = 0;
  while (l + 1 < nlev && b >= first[l + 1]) ++l;
  return l;
}

__global__ void __launch_bounds__(kMgBlock) mg_f16_max_all(F16Levels J, float* __restrict__ part) {
  const int b = blockIdx.x;
  const int l = f16_level_of(J.p0, J.nlev, b);
  const int nb = J.p0[l + 1] - J.p0[l], bl = b - J.p0[l];
  const int64_t count = (int64_t)kHalfSlots * 9 * J.n[l];
  const float* S = J.S[l];
  const int64_t n4 = count / 4;
  const float4* S4 = reinterpret_cast<const float4*>(S);
  float best = 0.f;
  for (int64_t i = (int64_t)bl * kMgBlock + threadIdx.x; i < n4; i += (int64_t)nb * kMgBlock) {
    const float4 v = S4[i];
    best = fmaxf(best, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
  }
  if (bl == 0)
    for (int64_t i = 4 * n4 + threadIdx.x; i < count; i += kMgBlock) best = fmaxf(best, fabsf(S[i]));
  block_max_to(best, part + b);
}

__global__ void __launch_bounds__(kMgBlock) mg_f16_convert_all(F16Levels J, const float* __restrict__ part) {
  __shared__ float sm[kMgBlock / kWave];
  const int b = blockIdx.x;
  const int l = f16_level_of(J.c0, J.nlev, b);
  // the level's maximum from its partials (at most kF16MaxParts)
  float m = 0.f;
  for (int i = J.p0[l] + threadIdx.x; i < J.p0[l + 1]; i += kMgBlock) m = fmaxf(m, part[i]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & (kWave - 1)) == 0) sm[threadIdx.x / kWave] = m;
  __syncthreads();
  m = sm[0];
  for (int k = 1; k < kMgBlock / kWave; ++k) m = fmaxf(m, sm[k]);
  const float sc = (m > 0.f && isfinite(m)) ? ldexpf(1.f, 14 - ilogbf(m)) : 1.f;
  if (b == J.c0[l] && threadIdx.x == 0) {
    J.scl[l][0] = sc;
    J.scl[l][1] = 1.f / sc;
    J.scl[l][2] = m;
  }
  const int64_t n = J.n[l];
  const int64_t t = (int64_t)(b - J.c0[l]) * kMgBlock + threadIdx.x;   // (slot, node), node fastest
  if (t >= n * kHalfSlots) return;
  const int sl = (int)(t / n);
  const int64_t node = t - (int64_t)sl * n;
  const float* S = J.S[l];
  float v[2 * kF16Words];
#pragma unroll
  for (int e = 0; e < 9; ++e) v[e] = sc * S[(int64_t)(sl * 9 + e) * n + node];
  v[9] = 0.f;
#pragma unroll
  for (int k = 0; k < kF16Words; ++k)
    J.W[l][(int64_t)(sl * kF16Words + k) * n + node] = __floats2half2_rn(v[2 * k], v[2 * k + 1]);
}

// row q of the node's 3x3 block of D times v
__device__ __forceinline__ float block_row(const float* __restrict__ D, int64_t n, int64_t node, int q,
                                           const float (&v)[3]) {
  return D[(3 * q) * n + node] * v[0] + D[(3 * q + 1) * n + node] * v[1] + D[(3 * q + 2) * n + node] * v[2];
}

// as mg_poly_step (res_in may alias res_out: a group reads all three entries of
// its node before any of its lanes stores one)
template <bool HALF, typename ST = float>
__global__ void __launch_bounds__(kMgBlock) mg_poly_step_lanes(const ST* __restrict__ S,
                                                               const float* __restrict__ scl,
                                                               const float* __restrict__ Dinv, Lat g,
                                                               const float* __restrict__ cf, int j,
                                                               const float* res_in, const float* __restrict__ d_in,
                                                               float* res_out, float* __restrict__ d_out,
                                                               float* __restrict__ e, const Scalars* sc) {
  if (sc && sc->done) return;
  const float c1 = cf[1 + j], c2 = cf[1 + kMaxDegree + j];
  MG_LANE_LOOP(g.n, {
    float ad[3];
    stencil_row_group<HALF, 4>(S, g, node, d_in, q, ad, scl);
    const float rr[3] = {res_in[3 * node] - ad[0], res_in[3 * node + 1] - ad[1], res_in[3 * node + 2] - ad[2]};
    if (q < 3) {
      const float dn = c1 * d_in[3 * node + q] + c2 * block_row(Dinv, g.n, node, q, rr);
      res_out[3 * node + q] = rr[q];
      d_out[3 * node + q] = dn;
      e[3 * node + q] += dn;
    }
  });
}

template <bool HALF, typename ST = float>
__global__ void __launch_bounds__(kMgBlock) mg_poly_resid_lanes(const ST* __restrict__ S,
                                                                const float* __restrict__ scl, Lat g,
                                                                const float* __restrict__ res_in,
                                                                const float* __restrict__ d_in,
                                                                const float* __restrict__ mask,
                                                                float* __restrict__ t, const Scalars* sc) {
  if (sc && sc->done) return;
  MG_LANE_LOOP(g.n, {
    float ad[3];
    stencil_row_group<HALF, 4>(S, g, node, d_in, q, ad, scl);
    if (q < 3) {
      const float v = res_in[3 * node + q] - ad[q];
      t[3 * node + q] = mask ? v * mask[3 * node + q] : v;
    }
  });
}

template <bool HALF, typename ST = float>
__global__ void __launch_bounds__(kMgBlock) mg_post_first_lanes(const ST* __restrict__ S,
                                                                const float* __restrict__ scl,
                                                                const float* __restrict__ Dinv, Lat g,
                                                                const float* __restrict__ cf,
                                                                const float* __restrict__ r,
                                                                const float* __restrict__ e_in,
                                                                float* __restrict__ res, float* __restrict__ d,
                                                                float* __restrict__ e_out, const Scalars* sc) {
  if (sc && sc->done) return;
  const float s0 = cf[0];
  MG_LANE_LOOP(g.n, {
    float ae[3];
    stencil_row_group<HALF, 4>(S, g, node, e_in, q, ae, scl);
    const float rr[3] = {r[3 * node] - ae[0], r[3 * node + 1] - ae[1], r[3 * node + 2] - ae[2]};
    if (q < 3) {
      const float dq = s0 * block_row(Dinv, g.n, node, q, rr);
      res[3 * node + q] = rr[q];
      d[3 * node + q] = dq;
      e_out[3 * node + q] = e_in[3 * node + q] + dq;
    }
  });
}

// ---------------------------------------------------------------------------
// Fine level kept as the tangent's SELL matrix (elements whose nodes couple more
// than one lattice step apart, e.g. HEX27 on its (2n+1)^3 node lattice).  The
// row-eliminated matrix acts on vectors that are zero on Dirichlet DOFs exactly
// as M A M does, so the smoother multiplies by it directly; the 3x3 diagonal
// blocks and the Gershgorin bound are formed from the masked blocks.  A smoothing
// step is a SELL product y = A v (the Krylov sweep) followed by a node-wise update.

// masked diagonal blocks (BC DOFs: unit diagonal) -> inverse, as mg_block_dinv
__global__ void __launch_bounds__(kMgBlock) mg_sell_dinv(const int32_t* __restrict__ dpos,
                                                         const float* __restrict__ vals, const float* __restrict__ mask,
                                                         int64_t n, float* __restrict__ Dinv) {
  const int64_t node = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  if (node >= n) return;
  double a[9];
  const int32_t b = dpos[node];
  const float m[3] = {mask[3 * node], mask[3 * node + 1], mask[3 * node + 2]};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      a[3 * i + j] = b >= 0 ? (double)vals[(int64_t)b + 3 * j + i] * m[i] * m[j] : 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (m[i] == 0.f) a[4 * i] += 1.0;
  double c[9];
  c[0] = a[4] * a[8] - a[5] * a[7];
  c[1] = a[2] * a[7] - a[1] * a[8];
  c[2] = a[1] * a[5] - a[2] * a[4];
  c[3] = a[5] * a[6] - a[3] * a[8];
  c[4] = a[0] * a[8] - a[2] * a[6];
  c[5] = a[2] * a[3] - a[0] * a[5];
  c[6] = a[3] * a[7] - a[4] * a[6];
  c[7] = a[1] * a[6] - a[0] * a[7];
  c[8] = a[0] * a[4] - a[1] * a[3];
  const double det = a[0] * c[0] + a[1] * c[3] + a[2] * c[6];
  const double sc = fmax(fabs(a[0]), fmax(fabs(a[4]), fabs(a[8])));
  if (det > 1e-12 * sc * sc * sc && isfinite(det)) {
#pragma unroll
    for (int e = 0; e < 9; ++e) Dinv[(int64_t)e * n + node] = (float)(c[e] / det) * m[e / 3] * m[e % 3];
  } else {
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      const bool d = (e % 4) == 0;
      Dinv[(int64_t)e * n + node] = (d && a[e] > 0.0) ? (float)(1.0 / a[e]) * m[e / 3] : 0.f;
    }
  }
}

// Gershgorin bound of Dinv (M A M) over the SELL rows (per-workgroup maxima, as mg_gersh)
// A wavefront per slice and a lane per (row, column component) as in the Krylov
// sweep: lane (r, i) reads column i of its row's blocks (3 contiguous floats,
// coalesced across the slice), masks it by the column DOF and adds
// |Dinv col| into the row's three sums (Dinv is zero on masked rows); the row's
// three lanes are summed through LDS.
__global__ void __launch_bounds__(kBlock) mg_sell_gersh(MatArgs m, const float* __restrict__ vals,
                                                        const float* __restrict__ mask,
                                                        const float* __restrict__ Dinv, int64_t n,
                                                        float* __restrict__ part) {
  using Sh = SellShape<3>;
  __shared__ float red[kWavesPerBlock][kWave * 3];
  __shared__ float sm[kWavesPerBlock];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int r = lane / 3, i = lane - (lane / 3) * 3;
  const bool in_slice = lane < Sh::L;
  int64_t s0, s1;
  row_range(m.nslices, s0, s1);
  float best = 0.f;
  for (int64_t s = s0 + w; s < s1; s += kWavesPerBlock) {
    const int32_t node = in_slice ? m.sell_row[s * Sh::R + r] : -1;
    float rs[3] = {0.f, 0.f, 0.f};
    if (node >= 0) {
      float M[9];
#pragma unroll
      for (int e = 0; e < 9; ++e) M[e] = Dinv[(int64_t)e * n + node];
      const int32_t t0 = m.slice_ptr[s], t1 = m.slice_ptr[s + 1];
#pragma unroll 4
      for (int32_t t = t0; t < t1; ++t) {
        const int64_t col = m.sell_col[(int64_t)t * Sh::R + r];
        const float mc = mask[3 * col + i];
        const float* vc = vals + ((int64_t)t * Sh::L + lane) * 3;  // entries (k, i), k = 0..2
        const float b0 = vc[0] * mc, b1 = vc[1] * mc, b2 = vc[2] * mc;
#pragma unroll
        for (int p = 0; p < 3; ++p) rs[p] += fabsf(M[3 * p] * b0 + M[3 * p + 1] * b1 + M[3 * p + 2] * b2);
      }
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) red[w][lane * 3 + p] = rs[p];
    wave_lds_sync();
    if (node >= 0 && i == 0) {
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const float tot = red[w][(r * 3) * 3 + p] + red[w][(r * 3 + 1) * 3 + p] + red[w][(r * 3 + 2) * 3 + p];
        best = fmaxf(best, tot);
      }
    }
    wave_lds_sync();  // the next slice rewrites red
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o, 64));
  if (lane == 0) sm[w] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    float mx = sm[0];
    for (int k = 1; k < kWavesPerBlock; ++k) mx = fmaxf(mx, sm[k]);
    part[blockIdx.x] = mx;
  }
}

// Stage 1 of the Galerkin product from the SELL rows: B = (M A M) P in 27
// coarse-column slots per fine row (slot s: coarse column floor(i/2) + s - 1 per
// dimension; mg_rap_ptap forms P^T B).  A wavefront per slice and a lane per
// (row, column component i), as in the Krylov sweep: the lane adds column i of
// each of its row's blocks (coalesced across the slice), weighted by the coarse
// parents of the block's column node, into its slot accumulators in LDS.  The
// slice's three wavefronts take the coarse x-planes sx = 0, 1, 2 of the 27 slots
// (9 slots x 3 entries per lane), reading the same blocks, so that the
// accumulators leave room for 8 waves per SIMD.  A parent outside the 27 slots
// sets bad[0] (the pattern is not supported).
__global__ void __launch_bounds__(3 * kWave) mg_sell_rap_ap(MatArgs m, const float* __restrict__ vals,
                                                            const float* __restrict__ mask, Lat f, Lat c,
                                                            float* __restrict__ B, int32_t* __restrict__ bad) {
  using Sh = SellShape<3>;
  __shared__ float acc[3 * kWave * 27];
  const int lane = threadIdx.x & (kWave - 1), plane = threadIdx.x / kWave;
  const int r = lane / 3, i = lane - (lane / 3) * 3;
  const int64_t sl = blockIdx.x;
  float* my = acc + (plane * kWave + lane) * 27;  // [slot in plane][k]: entry (k, i)
  for (int e = 0; e < 27; ++e) my[e] = 0.f;
  const int32_t node = lane < Sh::L ? m.sell_row[sl * Sh::R + r] : -1;
  if (node < 0) return;
  int fi[3];
  lat_xyz(f, node, fi[0], fi[1], fi[2]);
  const int nf[3] = {f.nx - 1, f.ny - 1, f.nz - 1};
  const int nc[3] = {c.nx - 1, c.ny - 1, c.nz - 1};
  const float mr[3] = {mask[3 * (int64_t)node], mask[3 * (int64_t)node + 1], mask[3 * (int64_t)node + 2]};
  bool broken = false;
  const int32_t t0 = m.slice_ptr[sl], t1 = m.slice_ptr[sl + 1];
#pragma unroll 4
  for (int32_t t = t0; t < t1; ++t) {
    const int64_t col = m.sell_col[(int64_t)t * Sh::R + r];
    int fj[3];
    lat_xyz32(f, (uint32_t)col, fj[0], fj[1], fj[2]);
    const float mc = mask[3 * col + i];
    const float* vc = vals + ((int64_t)t * Sh::L + lane) * 3;  // entries (k, i), k = 0..2
    const float b[3] = {vc[0] * mr[0] * mc, vc[1] * mr[1] * mc, vc[2] * mr[2] * mc};
    // coarse columns of the fine column j: per dimension the parent(s) of j
    int ps[3][2];
    float pw[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int v2 = fj[d];
      int c0, c1w;
      float w0, w1;
      if (v2 % 2 == 0) { c0 = v2 / 2; w0 = 1.f; c1w = c0; w1 = 0.f; }
      else if (v2 == nf[d]) { c0 = nc[d]; w0 = 1.f; c1w = c0; w1 = 0.f; }
      else { c0 = (v2 - 1) / 2; w0 = 0.5f; c1w = (v2 + 1) / 2; w1 = 0.5f; }
      pw[d][0] = w0; pw[d][1] = w1;
      ps[d][0] = c0 - fi[d] / 2 + 1;
      ps[d][1] = c1w - fi[d] / 2 + 1;
    }
#pragma unroll
    for (int ax = 0; ax < 2; ++ax)
#pragma unroll
      for (int ay = 0; ay < 2; ++ay)
#pragma unroll
        for (int az = 0; az < 2; ++az) {
          const float w = pw[0][ax] * pw[1][ay] * pw[2][az];
          if (w == 0.f) continue;
          const int sx = ps[0][ax], sy = ps[1][ay], sz = ps[2][az];
          if (sx < 0 || sx > 2 || sy < 0 || sy > 2 || sz < 0 || sz > 2) {
            broken = true;
            continue;
          }
          if (sx != plane) continue;
          float* dst = my + (sy * 3 + sz) * 3;
#pragma unroll
          for (int k = 0; k < 3; ++k) dst[k] += w * b[k];
        }
  }
  if (broken) bad[0] = 1;
  for (int q = 0; q < 9; ++q)
#pragma unroll
    for (int k = 0; k < 3; ++k) B[(int64_t)((plane * 9 + q) * 9 + 3 * k + i) * f.n + node] = my[q * 3 + k];
}

// y = A v over the SELL rows (the Krylov sweep), node-major output
__global__ void __launch_bounds__(kBlock) mg_sell_spmv(MatArgs m, const float* __restrict__ vals,
                                                       const float* __restrict__ v, float* __restrict__ y,
                                                       const Scalars* sc) {
  if (sc && sc->done) return;
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return v[c]; };
  auto epi = [&](int64_t n, int i, float o) { y[n * 3 + i] = o; };
  sweep<float, 3, false>(m, r0, r1, vals, xv, epi);
}

// The SELL fine level's values in FP16 for the smoothing products (HF_MG_FINE_F16):
// slot t's value k of lane l at W[(t*3 + k)*L + l] (vals[(t*L + l)*3 + k] times the
// power-of-two scale scl[0]), so a wavefront's loads of one (t, k) are 126
// contiguous bytes.
constexpr int kSellL = SellShape<3>::L;
__global__ void __launch_bounds__(kMgBlock) mg_sell_to_f16(const float* __restrict__ vals, int64_t nslots,
                                                           const float* __restrict__ scl, __half* __restrict__ W) {
  const int64_t idx = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;   // t * L + lane
  if (idx >= nslots * kSellL) return;
  const int64_t t = idx / kSellL;
  const int lane = (int)(idx - t * kSellL);
  const float sc = scl[0];
#pragma unroll
  for (int k = 0; k < 3; ++k) W[(t * 3 + k) * kSellL + lane] = __float2half_rn(sc * vals[idx * 3 + k]);
}

// y = A v as mg_sell_spmv (the same lanes, sums and order) from the FP16 values,
// scaled back by scl[1]
__global__ void __launch_bounds__(kBlock) mg_sell_spmv16(MatArgs m, const __half* __restrict__ W,
                                                         const float* __restrict__ scl, const float* __restrict__ v,
                                                         float* __restrict__ y, const Scalars* sc) {
  if (sc && sc->done) return;
  using S = SellShape<3>;
  int64_t s0, s1;
  row_range(m.nslices, s0, s1);
  const float sinv = scl[1];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  const int r = lane / 3, i = lane - r * 3;
  const bool in_slice = lane < S::L;
  __shared__ float red[kWavesPerBlock][kWave * 3];
  for (int64_t sl = s0 + w; sl < s1; sl += kWavesPerBlock) {
    const int32_t t0 = m.slice_ptr[sl], t1 = m.slice_ptr[sl + 1];
    float yk[3] = {0.f, 0.f, 0.f};
    if (in_slice) {
#pragma unroll 8
      for (int32_t t = t0; t < t1; ++t) {
        const int64_t cr = (int64_t)t * S::R + r;
        const float x = v[(int64_t)m.sell_col[cr] * 3 + i];
        const __half* wt = W + (int64_t)t * 3 * S::L + lane;
#pragma unroll
        for (int k = 0; k < 3; ++k) yk[k] += __half2float(wt[k * S::L]) * x;
      }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) red[w][lane * 3 + k] = yk[k];
    wave_lds_sync();
    if (in_slice) {
      float yi = 0.f;
#pragma unroll
      for (int j = 0; j < 3; ++j) yi += red[w][(r * 3 + j) * 3 + i];
      const int32_t n = m.sell_row[sl * S::R + r];
      if (n >= 0) y[(int64_t)n * 3 + i] = yi * sinv;
    }
    wave_lds_sync();  // the next slice rewrites red
  }
}

// node-wise halves of the fine-level smoothing passes, with y = A v from mg_sell_spmv:
//   step:  res = res_in - y (v = d_in); d_out = c1 d_in + c2 Dinv res; e += d_out
//   resid: t = res_in - y (v = d_in)
//   post:  res = r - y (v = e_in); d = s0 Dinv res; e_out = e_in + d
__global__ void __launch_bounds__(kMgBlock) mg_sell_step_update(const float* __restrict__ Dinv, int64_t n,
                                                                const float* __restrict__ cf, int j,
                                                                const float* res_in, const float* __restrict__ y,
                                                                const float* __restrict__ d_in, float* res_out,
                                                                float* __restrict__ d_out, float* __restrict__ e,
                                                                const Scalars* sc) {
  if (sc && sc->done) return;
  const float c1 = cf[1 + j], c2 = cf[1 + kMaxDegree + j];
  MG_NODE_LOOP(n, {
    const float res[3] = {res_in[3 * node] - y[3 * node], res_in[3 * node + 1] - y[3 * node + 1],
                          res_in[3 * node + 2] - y[3 * node + 2]};
    float dr[3];
    block_apply(Dinv, n, node, res, dr);
    for (int i = 0; i < 3; ++i) {
      const float dn = c1 * d_in[3 * node + i] + c2 * dr[i];
      res_out[3 * node + i] = res[i];
      d_out[3 * node + i] = dn;
      e[3 * node + i] += dn;
    }
  });
}

__global__ void __launch_bounds__(kMgBlock) mg_sell_resid_update(int64_t n, const float* __restrict__ res_in,
                                                                 const float* __restrict__ y, float* __restrict__ t,
                                                                 const Scalars* sc) {
  if (sc && sc->done) return;
  MG_NODE_LOOP(n, {
    for (int i = 0; i < 3; ++i) t[3 * node + i] = res_in[3 * node + i] - y[3 * node + i];
  });
}

// 1 / diag(A) of a SELL fine level (the learned smoother's scalar Jacobi, as
// mg_scalar_dinv); 1 on Dirichlet DOFs
__global__ void __launch_bounds__(kMgBlock) mg_sell_scalar_dinv(const int32_t* __restrict__ dpos,
                                                                const float* __restrict__ vals,
                                                                const float* __restrict__ mask, int64_t n,
                                                                float* __restrict__ dinv) {
  const int64_t node = (int64_t)blockIdx.x * kMgBlock + threadIdx.x;
  if (node >= n) return;
  const int32_t b = dpos[node];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float a = b >= 0 ? vals[(int64_t)b + 4 * i] : 0.f;
    const bool free = !mask || mask[3 * node + i] != 0.f;
    dinv[3 * node + i] = (free && a != 0.f) ? 1.f / a : 1.f;
  }
}

// res = mask (r - y): the fine residual the learned step reads, y = A e by the SELL product
__global__ void __launch_bounds__(kMgBlock) mg_sell_masked_resid(int64_t n, const float* __restrict__ r,
                                                                 const float* __restrict__ y,
                                                                 const float* __restrict__ mask,
                                                                 float* __restrict__ res, const Scalars* sc) {
  if (sc && sc->done) return;
  MG_NODE_LOOP(n, {
    for (int i = 0; i < 3; ++i) res[3 * node + i] = mask[3 * node + i] * (r[3 * node + i] - y[3 * node + i]);
  });
}

__global__ void __launch_bounds__(kMgBlock) mg_sell_post_update(const float* __restrict__ Dinv, int64_t n,
                                                                const float* __restrict__ cf,
                                                                const float* __restrict__ r,
                                                                const float* __restrict__ y,
                                                                const float* __restrict__ e_in, float* __restrict__ res,
                                                                float* __restrict__ d, float* __restrict__ e_out,
                                                                const Scalars* sc) {
  if (sc && sc->done) return;
  const float s0 = cf[0];
  MG_NODE_LOOP(n, {
    const float rr[3] = {r[3 * node] - y[3 * node], r[3 * node + 1] - y[3 * node + 1], r[3 * node + 2] - y[3 * node + 2]};
    float dr[3];
    block_apply(Dinv, n, node, rr, dr);
    for (int i = 0; i < 3; ++i) {
      res[3 * node + i] = rr[i];
      d[3 * node + i] = s0 * dr[i];
      e_out[3 * node + i] = e_in[3 * node + i] + s0 * dr[i];
    }
  });
}

// ---------------------------------------------------------------------------
// outer flexible CG kernels

// workspace: scalars, partials A (<= 2 per workgroup) and B (<= 3), then the vectors
constexpr int kPartB = 2 * kMaxGrid;
constexpr size_t kMgHeader = 4096 + (size_t)5 * kMaxGrid * sizeof(double);

// r = b - A x, p = q = 0; partials (A) r.r, b.b
// Sum of nparts partials in a fixed order by one kMgBlock workgroup; every
// thread receives the totals, and every workgroup that runs it gets the same
// bits (same data, same order).
template <int NV>
__device__ __forceinline__ void sum_partials_wg(const double* __restrict__ partials, int nparts, double (&tot)[NV]) {
  __shared__ double sm[kMgBlock / kWave][NV];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  double acc[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) acc[j] = 0;
  for (int b = threadIdx.x; b < nparts; b += kMgBlock)
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] += partials[(int64_t)b * NV + j];
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    const double v = wave_sum(acc[j]);
    if (lane == 0) sm[w][j] = v;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double v = 0;
    for (int k = 0; k < kMgBlock / kWave; ++k) v += sm[k][j];
    tot[j] = v;
  }
}

__global__ void __launch_bounds__(kBlock) mgcg_init_kernel(MatArgs m, const float* __restrict__ vals,
                                                           const float* __restrict__ b, const float* __restrict__ x,
                                                           float* __restrict__ r, float* __restrict__ p,
                                                           float* __restrict__ q, double* partials) {
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return x[c]; };
  double loc[2] = {0, 0};
  auto epi = [&](int64_t n, int i, float ax) {
    const int64_t d = n * 3 + i;
    const float rr = b[d] - ax;
    r[d] = rr;
    p[d] = 0.f;
    q[d] = 0.f;
    loc[0] += (double)rr * (double)rr;
    loc[1] += (double)b[d] * (double)b[d];
  };
  sweep<float, 3, false>(m, r0, r1, vals, xv, epi);
  block_partials<2>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) mgcg_init_scalars(Scalars* sc, const double* partA, int nA,
                                                                  double rtol2, double atol2, int rel_r0,
                                                                  int maxiter) {
  double a[2];
  sum_partials<2>(partA, nA, a);
  if (threadIdx.x != 0) return;
  sc->rr = a[0];
  sc->bb = a[1];
  sc->thresh = fmax(rtol2 * (rel_r0 ? a[0] : a[1]), atol2);
  sc->iter = 0;
  sc->status = 0;
  sc->maxiter = maxiter;
  sc->alpha = 0.0;
  sc->omega = 0.0;
  sc->pq = 0.0;
  sc->rho = 0.0;
  sc->done = (a[0] <= sc->thresh) ? 1 : 0;
  if (!isfinite(a[0])) { sc->done = 1; sc->status = 3; }
  else if (!sc->done && maxiter <= 0) { sc->done = 1; sc->status = 1; }
}

// w = A z; partials (B) r.z, z.w, z.q
__global__ void __launch_bounds__(kBlock) mgcg_spmv_kernel(MatArgs m, const float* __restrict__ vals,
                                                           const float* __restrict__ z, const float* __restrict__ r,
                                                           const float* __restrict__ q, float* __restrict__ w,
                                                           const Scalars* sc, double* partials) {
  if (sc->done) return;
  int64_t r0, r1;
  row_range(m.nslices, r0, r1);
  auto xv = [&](int64_t c) { return z[c]; };
  double loc[3] = {0, 0, 0};
  auto epi = [&](int64_t n, int i, float v) {
    const int64_t d = n * 3 + i;
    w[d] = v;
    const double zz = (double)z[d];
    loc[0] += (double)r[d] * zz;
    loc[1] += zz * (double)v;
    loc[2] += zz * (double)q[d];
  };
  sweep<float, 3, false>(m, r0, r1, vals, xv, epi);
  block_partials<3>(loc, partials);
}

// The same two phases with the tangent as a half-stored lattice stencil (Dirichlet
// rows included): A v = M (S v) + (1 - M) v, the row-eliminated operator.
template <typename ST = float>
__global__ void __launch_bounds__(kMgBlock) mgcg_init_stencil(const ST* __restrict__ S, const float* __restrict__ scl,
                                                              Lat g,
                                                              const float* __restrict__ mask,
                                                              const float* __restrict__ b,
                                                              const float* __restrict__ x, float* __restrict__ r,
                                                              float* __restrict__ p, float* __restrict__ q,
                                                              double* partials) {
  double loc[2] = {0, 0};
  MG_NODE_LOOP(g.n, {
    int cx, cy, cz;
    lat_xyz(g, node, cx, cy, cz);
    float ax[3];
    stencil_row<true>(S, g, node, cx, cy, cz, x, ax, scl);
    for (int i = 0; i < 3; ++i) {
      const int64_t d = node * 3 + i;
      const float mk = mask[d];
      const float rr = b[d] - (mk * ax[i] + (1.f - mk) * x[d]);
      r[d] = rr;
      p[d] = 0.f;
      q[d] = 0.f;
      loc[0] += (double)rr * (double)rr;
      loc[1] += (double)b[d] * (double)b[d];
    }
  });
  block_partials<2>(loc, partials);
}

template <typename ST = float>
__global__ void __launch_bounds__(kMgBlock) mgcg_spmv_stencil(const ST* __restrict__ S, const float* __restrict__ scl,
                                                              Lat g,
                                                              const float* __restrict__ mask,
                                                              const float* __restrict__ z,
                                                              const float* __restrict__ r,
                                                              const float* __restrict__ q, float* __restrict__ w,
                                                              const Scalars* sc, double* partials) {
  if (sc->done) return;
  double loc[3] = {0, 0, 0};
  MG_NODE_LOOP(g.n, {
    int cx, cy, cz;
    lat_xyz(g, node, cx, cy, cz);
    float az[3];
    stencil_row<true>(S, g, node, cx, cy, cz, z, az, scl);
    for (int i = 0; i < 3; ++i) {
      const int64_t d = node * 3 + i;
      const float mk = mask[d];
      const float zz = z[d];
      const float v = mk * az[i] + (1.f - mk) * zz;
      w[d] = v;
      loc[0] += (double)r[d] * (double)zz;
      loc[1] += (double)zz * (double)v;
      loc[2] += (double)zz * (double)q[d];
    }
  });
  block_partials<3>(loc, partials);
}

// y = A x for the stencil form (tests and the fallback)
template <typename ST = float>
__global__ void __launch_bounds__(kMgBlock) mg_stencil_matvec(const ST* __restrict__ S, const float* __restrict__ scl,
                                                              Lat g,
                                                              const float* __restrict__ mask,
                                                              const float* __restrict__ x, float* __restrict__ y) {
  MG_NODE_LOOP(g.n, {
    int cx, cy, cz;
    lat_xyz(g, node, cx, cy, cz);
    float ax[3];
    stencil_row<true>(S, g, node, cx, cy, cz, x, ax, scl);
    for (int i = 0; i < 3; ++i) {
      const int64_t d = node * 3 + i;
      y[d] = mask[d] * ax[i] + (1.f - mask[d]) * x[d];
    }
  });
}

// beta = -zeta / p.q_old, p.q = delta + 2 beta zeta + beta^2 p.q_old, alpha =
// gamma / p.q (beta = 0 on the first step) from d = (gamma, delta, zeta) =
// (r.z, z.w, z.q); returns 0, or the breakdown status (2: p.q <= 0, 3: non-finite).
__device__ __forceinline__ int dir_step(const double (&d)[3], int it, double pq_old, double& alpha, double& beta,
                                        double& pq) {
  const double gamma = d[0], delta = d[1], zeta = d[2];
  beta = 0.0;
  pq = delta;
  if (it > 0) {
    beta = -zeta / pq_old;
    pq = delta + 2.0 * beta * zeta + beta * beta * pq_old;
  }
  alpha = gamma / pq;
  if (!(pq > 0.0) || !isfinite(pq) || !isfinite(gamma)) return (isfinite(pq) && isfinite(gamma)) ? 2 : 3;
  return 0;
}
// p.q alternates between two slots by iteration parity, so no reader of the old
// value can see it replaced within the kernel that forms the new one
__device__ __forceinline__ void dir_record(Scalars* sc, int st, int it, double gamma, double alpha, double beta,
                                           double pq) {
  if (st) {
    sc->status = st;
    sc->done = 1;
    return;
  }
  sc->pqs[(it + 1) & 1] = pq;
  sc->omega = beta;
  sc->rho = gamma;
  sc->alpha = alpha;
}

// The direction scalars by one workgroup, for products with many partials
// (a grid of more than kFusedParts workgroups), where every workgroup of the
// update summing them all would cost more than this launch.
constexpr int kFusedParts = 2048;
__global__ void __launch_bounds__(kScalarBlock) mgcg_dir_scalars(Scalars* sc, const double* partB, int nB) {
  if (sc->done) return;
  double d[3];
  sum_partials<3>(partB, nB, d);
  if (threadIdx.x != 0) return;
  const int it = sc->iter;
  double alpha, beta, pq;
  const int st = dir_step(d, it, it > 0 ? sc->pqs[it & 1] : 0.0, alpha, beta, pq);
  dir_record(sc, st, it, d[0], alpha, beta, pq);
}

// One outer step, node by node: with nB partials (at most kFusedParts) every
// workgroup first sums the product's partials (B: gamma = r.z, delta = z.w,
// zeta = z.q) in the same fixed order and forms the direction scalars itself
// (dir_step; workgroup 0 records them); with nB = 0 mgcg_dir_scalars has formed
// them.  Then p = z + beta p, q = w + beta q, x += alpha p, r -= alpha q, the fine
// level's pre-smoother start from e = 0 on the new residual (d = e = s0 Dinv r;
// Dinv null: a one-level hierarchy, whose cycle is the dense solve of r), and
// partials (A) r.r for the convergence test.
__global__ void __launch_bounds__(kMgBlock) mgcg_update_first(int64_t n, const float* __restrict__ z,
                                                              const float* __restrict__ w, float* __restrict__ p,
                                                              float* __restrict__ q, float* __restrict__ x,
                                                              float* __restrict__ r, const float* __restrict__ Dinv,
                                                              const float* __restrict__ cf, float* __restrict__ d0,
                                                              float* __restrict__ e0, Scalars* sc,
                                                              const double* __restrict__ partB, int nB,
                                                              double* partials) {
  if (sc->done) return;
  double alpha_d, beta;
  if (nB > 0) {
    double dv[3];
    sum_partials_wg<3>(partB, nB, dv);
    double pq;
    const int it = sc->iter;
    const int st = dir_step(dv, it, it > 0 ? sc->pqs[it & 1] : 0.0, alpha_d, beta, pq);
    if (blockIdx.x == 0 && threadIdx.x == 0) dir_record(sc, st, it, dv[0], alpha_d, beta, pq);
    if (st) return;
  } else {
    // mgcg_dir_scalars formed them (many partials: one workgroup sums them once)
    alpha_d = sc->alpha;
    beta = sc->omega;
  }
  const float alpha = (float)alpha_d, betaf = (float)beta;
  const float s0 = Dinv ? cf[0] : 0.f;
  double loc[1] = {0};
  MG_NODE_LOOP(n, {
    float rn[3];
    for (int i = 0; i < 3; ++i) {
      const int64_t k = 3 * node + i;
      const float pp = z[k] + betaf * p[k];
      const float qq = w[k] + betaf * q[k];
      p[k] = pp;
      q[k] = qq;
      x[k] += alpha * pp;
      rn[i] = r[k] - alpha * qq;
      r[k] = rn[i];
      loc[0] += (double)rn[i] * (double)rn[i];
    }
    if (Dinv) {
      float o[3];
      block_apply(Dinv, n, node, rn, o);
      for (int i = 0; i < 3; ++i) {
        d0[3 * node + i] = s0 * o[i];
        e0[3 * node + i] = s0 * o[i];
      }
    }
  });
  block_partials<1>(loc, partials);
}

__global__ void __launch_bounds__(kScalarBlock) mgcg_conv_scalars(Scalars* sc, const double* partA, int nA) {
  if (sc->done) return;
  double a[1];
  sum_partials<1>(partA, nA, a);
  if (threadIdx.x != 0) return;
  sc->iter += 1;
  sc->rr = a[0];
  if (!isfinite(a[0])) { sc->done = 1; sc->status = 3; }
  else if (a[0] <= sc->thresh) { sc->done = 1; sc->status = 0; }
  else if (sc->iter >= sc->maxiter) { sc->done = 1; sc->status = 1; }
}

// Smoothing coefficients of levels 0..nl-1 from their Gershgorin bounds, one
// thread per level, in FP64 as the recurrence defines them (mg_poly_step):
// Chebyshev on [lam / param, lam]: theta = (hi + lo) / 2, delta = (hi - lo) / 2,
// sigma = theta / delta, s0 = 1 / theta, rho_1 = 1 / sigma, and for j >= 1
// rho_{j+1} = 1 / (2 sigma - rho_j), c1_j = rho_{j+1} rho_j, c2_j = 2 rho_{j+1} / delta;
// damped block Jacobi: s0 = c2_j = omega, c1_j = 0.
__global__ void mg_coef(const float* __restrict__ lam, int nl, int kind, float param, float* __restrict__ coef) {
  const int l = threadIdx.x;
  if (l >= nl) return;
  float* cf = coef + (int64_t)l * kCoefStride;
  if (kind == HF_MG_JACOBI) {
    cf[0] = param;
    for (int j = 0; j < kMaxDegree; ++j) {
      cf[1 + j] = 0.f;
      cf[1 + kMaxDegree + j] = param;
    }
    return;
  }
  const double hi = lam[l] > 0.f ? (double)lam[l] : 1.0, lo = hi / (double)param;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  double rho = 1.0 / sigma;
  cf[0] = (float)(1.0 / theta);
  cf[1] = 0.f;
  cf[1 + kMaxDegree] = 0.f;
  for (int j = 1; j < kMaxDegree; ++j) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    cf[1 + j] = (float)(rn * rho);
    cf[1 + kMaxDegree + j] = (float)(2.0 * rn / delta);
    rho = rn;
  }
}

}  // namespace
}  // namespace hf

// ---------------------------------------------------------------------------
// host side
using namespace hf;

struct MgLevel {
  Lat g;
  bool half = false;      // stencil half-stored (diagonal + 13 upper slots: 126 n floats)
  float* S = nullptr;     // 243 n, or 126 n when half
  float* Dinv = nullptr;  // 9 n
  float* r = nullptr;     // 3 n (levels >= 1: the restricted residual)
  float* e = nullptr;     // 3 n
  float* e2 = nullptr;    // 3 n
  float* da = nullptr;    // 3 n smoother directions (two, alternating)
  float* db = nullptr;    // 3 n
  float* res = nullptr;   // 3 n smoother residual / restriction input
  __half2* S16 = nullptr; // half-stored level with HF_MG_FINE_F16: 70 n words (mg_f16_convert_all)
  bool s16_ok = false;    // S16 made by the last set-up (level 0 skips it while a cell operator is set)
};

struct hf_mg {
  int nlev = 0;
  MgLevel lv[kMgMaxLevels];
  float* B = nullptr;        // Galerkin-product scratch, reused by every level
  double* Ainv = nullptr;    // coarsest dense inverse
  int nd = 0;
  int32_t* flags = nullptr;  // [0] pattern not on the lattice, [1] dropped coarse pivots
  float* part = nullptr;     // kMaxGrid per-workgroup maxima + kMgMaxLevels bounds
  float* coef = nullptr;     // kMgMaxLevels x kCoefStride smoothing coefficients (mg_coef)
  bool half = false;         // fine-level stencil half-stored (symmetric operator)
  bool f16 = false;          // fine-level smoothing passes read an FP16 copy of the stencil
  float* scl = nullptr;      // per level (4 floats each): scale, inverse scale, |S| maximum (mg_f16_scale)
  const float* s0 = nullptr; // the FP32 fine stencil of the last set-up (the caller's, or lv[0].S)
  // matrix-free fine level (hf_mg_set_cell_operator): the cell matrix on the device,
  // its host copy, and the per-cell scales (the caller's; null: off)
  MfOp* mf = nullptr;
  float mf_host[8 * 6] = {0};
  bool mf_loaded = false;
  const float* mf_scale = nullptr;
  bool sell = false;         // fine level is the tangent's SELL matrix (no fine stencil)
  float* y = nullptr;        // 3 n0: SELL products of the fine level
  const hf_pattern* sp = nullptr;  // the fine level's pattern and values (last set-up, caller's memory)
  const float* svals = nullptr;
  __half* svals16 = nullptr;       // their FP16 copy (HF_MG_FINE_F16), scale in scl[0..2]
  int64_t svals16_cap = 0;
  bool svals16_ok = false;
  int kind = HF_MG_CHEBYSHEV;
  int degree = 2;
  int degree_post = 0;       // post-smoothing steps on the grid levels (0: degree)
  float param = 8.f;         // Chebyshev: lam / lower end of the interval; Jacobi: omega
  const float* mask = nullptr;  // level-0 Dirichlet mask of the last set-up (caller's memory)
  const hf_pattern* checked = nullptr;  // pattern whose lattice structure was verified
  int64_t checked_nblocks = -1;
  bool ready = false;
  // hf_mg_solve_f32's poll batch as a captured graph, replayed while its key
  // (the cycle's shape and every pointer and size the batch uses) matches; the
  // level data and coefficients are read from device memory, so a new set-up
  // reuses the graph
  // learned fine-level smoother (hf_mg_set_learned): mode 1 replaces the fine
  // post-smoothing by one network step, mode 2 adds the step after it
  int lmode = 0;
  LpcSmoother lsm;
  float* sdinv = nullptr;           // 3 n0: 1 / diag of the fine operator (each set-up)
  uint64_t gen = 0;                 // bumped when the smoother's kind or degree changes
  hipStream_t cap = nullptr;        // capture stream (non-blocking, default priority)
  hipGraphExec_t gexec = nullptr;
  uintptr_t gkey[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

static const float* coef_of(const hf_mg* mg, int l) { return mg->coef + (int64_t)l * kCoefStride; }

// A small-cycle level's stencil form (as the level holds it after a set-up).
static int small_kind(const MgLevel& V) {
  return V.S16 ? kSmallF16 : V.half ? kSmallHalf : kSmallFull;
}

// First level run inside mg_small_cycle (the levels from it down): the coarsest
// levels, from the dense one up, while a level has at most kSmallStart nodes, the
// cycle at most kSmallMax levels and their LDS arena fits; nlev - 1 (only the dense
// solve) when no level above the dense one qualifies.  The fine level never.
static int small_cycle_start(const hf_mg* mg) {
  const int L = mg->nlev;
  if (L == 1) return 0;
  int lb = L - 1;
  int64_t words = small_level_words(0, mg->lv[L - 1].g.n, true);
  for (int l = L - 2; l >= 1; --l) {
    const MgLevel& V = mg->lv[l];
    if (V.g.n > kSmallStart || L - l > kSmallMax) break;
    words += small_level_words(small_kind(V), V.g.n, false);
    if (words > kSmallArenaWords) break;
    lb = l;
  }
  return lb;
}

static int mg_free(hf_mg* mg) {
  for (int l = 0; l < mg->nlev; ++l) {
    MgLevel& L = mg->lv[l];
    float* bufs[] = {L.S, L.Dinv, L.r, L.e, L.e2, L.da, L.db, L.res};
    for (float* b : bufs)
      if (b) (void)hipFree(b);
  }
  if (mg->B) (void)hipFree(mg->B);
  if (mg->y) (void)hipFree(mg->y);
  if (mg->gexec) (void)hipGraphExecDestroy(mg->gexec);
  if (mg->cap) (void)hipStreamDestroy(mg->cap);
  if (mg->Ainv) (void)hipFree(mg->Ainv);
  if (mg->flags) (void)hipFree(mg->flags);
  if (mg->part) (void)hipFree(mg->part);
  if (mg->coef) (void)hipFree(mg->coef);
  if (mg->lsm.buf) (void)hipFree(mg->lsm.buf);
  if (mg->sdinv) (void)hipFree(mg->sdinv);
  if (mg->svals16) (void)hipFree(mg->svals16);
  for (int l = 0; l < mg->nlev; ++l)
    if (mg->lv[l].S16) (void)hipFree(mg->lv[l].S16);
  if (mg->scl) (void)hipFree(mg->scl);
  if (mg->mf) (void)hipFree(mg->mf);
  return 0;
}

extern "C" int hf_mg_create(int32_t nx, int32_t ny, int32_t nz, int32_t flags, hf_mg** out) {
  HF_REQUIRE(out, "hf_mg_create: null output");
  HF_REQUIRE((flags & ~(HF_MG_FINE_FULL | HF_MG_FINE_SELL | HF_MG_FINE_F16)) == 0,
             "hf_mg_create: unknown flags 0x%x", flags);
  HF_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2, "hf_mg_create: lattice %dx%dx%d needs >= 2 nodes per dimension", nx,
             ny, nz);
  HF_REQUIRE((int64_t)nx * ny * nz < (int64_t)1 << 30, "hf_mg_create: lattice too large");
  // level dimensions first (no allocation until they are known)
  int dims[kMgMaxLevels][3];
  int nlev = 0;
  int a = nx, b = ny, c = nz;
  for (;;) {
    HF_REQUIRE(nlev < kMgMaxLevels, "hf_mg_create: too many levels");
    dims[nlev][0] = a; dims[nlev][1] = b; dims[nlev][2] = c;
    ++nlev;
    if (3 * (int64_t)a * b * c <= kDenseMax) break;
    // intervals n -> ceil(n/2): nodes a -> a/2 + 1 (a lattice of 2 nodes stays)
    const int na = a / 2 + 1, nb = b / 2 + 1, nc = c / 2 + 1;
    HF_REQUIRE(!(na == a && nb == b && nc == c), "hf_mg_create: cannot coarsen %dx%dx%d below %d DOF", a, b, c,
               kDenseMax);
    a = na; b = nb; c = nc;
  }
  hf_mg* mg = new hf_mg();
  mg->nlev = nlev;
  mg->sell = (flags & HF_MG_FINE_SELL) != 0;
  // a one-level hierarchy is the dense FP64 solve, which reads the full stencil
  mg->half = !(flags & HF_MG_FINE_FULL) && !mg->sell && nlev > 1;
  // FP16 copies of the half-stored levels the grid kernels smooth (none if no level is)
  mg->f16 = (flags & HF_MG_FINE_F16) != 0;
  if (mg->sell && nlev < 2) {
    delete mg;
    HF_REQUIRE(false, "hf_mg_create: a SELL fine level needs at least one coarser level");
  }
  for (int l = 0; l < nlev; ++l) mg->lv[l].g = make_lat(dims[l][0], dims[l][1], dims[l][2]);
  // Coarse levels are half-stored too (Galerkin products of a symmetric
  // operator); the dense one keeps 27 slots (its set-up reads them).
  // HF_MG_FINE_FULL (an operator that need not be symmetric) keeps all.
  mg->lv[0].half = mg->half;
  // (a SELL fine level's level 1 comes from mg_rap_ptap, which writes 27 slots)
  for (int l = mg->sell ? 2 : 1; l + 1 < nlev; ++l) mg->lv[l].half = !(flags & HF_MG_FINE_FULL);
  auto alloc = [&](float** p, int64_t n) -> int {
    HF_CHECK_HIP(hipMalloc((void**)p, (size_t)n * sizeof(float)));
    return 0;
  };
  int rc = 0;
  for (int l = 0; l < mg->nlev && !rc; ++l) {
    MgLevel& L = mg->lv[l];
    const int64_t n = L.g.n;
    if (L.half) {
      rc |= alloc(&L.S, (int64_t)kHalfSlots * 9 * n);
    } else if (l == 0 && mg->sell) {
      rc |= alloc(&mg->y, 3 * n);
    } else {
      rc |= alloc(&L.S, kSt * n);
    }
    if (l + 1 < mg->nlev) rc |= alloc(&L.Dinv, 9 * n) | alloc(&L.da, 3 * n) | alloc(&L.db, 3 * n) |
                                alloc(&L.res, 3 * n);
    if (l > 0) rc |= alloc(&L.r, 3 * n);
    rc |= alloc(&L.e, 3 * n) | alloc(&L.e2, 3 * n);
  }
  if (!rc && mg->nlev > 1) {
    // B: the SELL fine level's A P (27 coarse-column slots per fine row), and the
    // intermediate lattices of the dimension-wise Galerkin products
    int64_t nb = mg->sell ? mg->lv[0].g.n : 0;
    for (int l = mg->sell ? 1 : 0; l + 1 < mg->nlev; ++l)
      nb = std::max(nb, rap_dim_scratch_nodes(mg->lv[l].g, mg->lv[l + 1].g));
    rc |= alloc(&mg->B, kSt * nb);
  }
  if (!rc) rc |= alloc(&mg->part, kMaxGrid + kMgMaxLevels);
  if (!rc && mg->f16) {
    // an FP16 copy for every half-stored level (its smoothing passes and
    // restricted residuals read it)
    rc |= alloc(&mg->scl, 4 * kMgMaxLevels);
    for (int l = 0; l + 1 < mg->nlev && !rc; ++l)
      if (mg->lv[l].half && hipMalloc((void**)&mg->lv[l].S16, (size_t)kHalfSlots * kF16Words * mg->lv[l].g.n *
                                                                  sizeof(__half2)) != hipSuccess) {
        set_error("hf_mg_create: hipMalloc failed");
        rc = -2;
      }
  }
  if (!rc) rc |= alloc(&mg->coef, (int64_t)kMgMaxLevels * kCoefStride);
  mg->nd = (int)(3 * mg->lv[mg->nlev - 1].g.n);
  if (!rc) {
    hipError_t e1 = hipMalloc((void**)&mg->Ainv, (size_t)mg->nd * mg->nd * sizeof(double));
    hipError_t e2 = hipMalloc((void**)&mg->flags, 4 * sizeof(int32_t));
    if (e1 != hipSuccess || e2 != hipSuccess) {
      set_error("hf_mg_create: hipMalloc failed");
      rc = -2;
    }
  }
  if (rc) {
    mg_free(mg);
    delete mg;
    return rc;
  }
  *out = mg;
  return 0;
}

extern "C" int hf_mg_destroy(hf_mg* mg) {
  if (!mg) return 0;
  mg_free(mg);
  delete mg;
  return 0;
}

extern "C" int hf_mg_levels(const hf_mg* mg, int32_t* nlev, int32_t* dims, int32_t max_levels) {
  HF_REQUIRE(mg && nlev, "hf_mg_levels: null argument");
  *nlev = mg->nlev;
  if (dims)
    for (int l = 0; l < mg->nlev && l < max_levels; ++l) {
      dims[3 * l] = mg->lv[l].g.nx;
      dims[3 * l + 1] = mg->lv[l].g.ny;
      dims[3 * l + 2] = mg->lv[l].g.nz;
    }
  return 0;
}

// Polynomial coefficients of every level from its bound lam.
// the levels' smoothing coefficients from the bounds of the last set-up, on stream s
static int mg_coefficients(hf_mg* mg, hipStream_t s) {
  const float* lam_dev = mg->part + kMaxGrid;
  if (mg->nlev > 1) mg_coef<<<1, kMgMaxLevels, 0, s>>>(lam_dev, mg->nlev - 1, mg->kind, mg->param, mg->coef);
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_mg_set_smoother(hf_mg* mg, int32_t kind, int32_t degree, float param) {
  HF_REQUIRE(mg, "hf_mg_set_smoother: null handle");
  HF_REQUIRE(kind == HF_MG_JACOBI || kind == HF_MG_CHEBYSHEV, "hf_mg_set_smoother: unknown kind %d", kind);
  HF_REQUIRE(degree >= 1 && degree <= kMaxDegree, "hf_mg_set_smoother: degree %d outside 1..%d", degree, kMaxDegree);
  if (kind == HF_MG_JACOBI)
    HF_REQUIRE(param > 0.f && param < 2.f, "hf_mg_set_smoother: omega %g outside (0, 2)", (double)param);
  else
    HF_REQUIRE(param > 1.f, "hf_mg_set_smoother: interval ratio %g must exceed 1", (double)param);
  mg->kind = kind;
  mg->degree = degree;
  mg->param = param;
  if (mg->degree_post > degree) mg->degree_post = 0;
  ++mg->gen;  // the cycle's launch sequence depends on the degree
  if (mg->ready) {
    // new coefficients for the current levels; the set-up ran on a caller's stream
    HF_CHECK_HIP(hipDeviceSynchronize());
    if (mg_coefficients(mg, nullptr)) return -2;
    HF_CHECK_HIP(hipDeviceSynchronize());
  }
  return 0;
}

extern "C" int hf_mg_set_post_degree(hf_mg* mg, int32_t degree) {
  HF_REQUIRE(mg, "hf_mg_set_post_degree: null handle");
  HF_REQUIRE(degree >= 0 && degree <= mg->degree, "hf_mg_set_post_degree: %d outside 0..%d (0: the pre-smoothing "
             "degree)", degree, mg->degree);
  if (degree != mg->degree_post) ++mg->gen;
  mg->degree_post = degree;
  return 0;
}

extern "C" int hf_mg_set_cell_operator(hf_mg* mg, const float* k, const float* cell_scale) {
  HF_REQUIRE(mg, "hf_mg_set_cell_operator: null handle");
  if (!cell_scale) {
    if (mg->mf_scale) {
      ++mg->gen;   // the cycle's kernels change
      // a set-up with the cell operator made no FP16 copy: the smoother has none until the next
      if (mg->lv[0].S16 && !mg->lv[0].s16_ok) mg->ready = false;
    }
    mg->mf_scale = nullptr;
    return 0;
  }
  HF_REQUIRE(k, "hf_mg_set_cell_operator: null cell matrix");
  HF_REQUIRE(!mg->sell && mg->nlev > 1, "hf_mg_set_cell_operator: needs a lattice-stencil fine level and a coarser "
                                        "level");
  float kmax = 0.f;
  for (int i = 0; i < 8 * 8 * 9; ++i) {
    HF_REQUIRE(std::isfinite(k[i]), "hf_mg_set_cell_operator: non-finite cell matrix entry %d", i);
    kmax = std::max(kmax, std::fabs(k[i]));
  }
  MfOp h;
  for (int q = 0; q < 8; ++q)
    for (int r = 0; r < 3; ++r)
      for (int c = r; c < 3; ++c) h.b[q * 6 + mf_tri(r, c)] = k[q * 9 + 3 * r + c];   // block (0, q)
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          const float want = k[(8 * a + b) * 9 + 3 * r + c];
          const float got = mf_sign(a, b, r, c) * h.b[(a ^ b) * 6 + mf_tri(r, c)];
          HF_REQUIRE(std::fabs(want - got) <= 1e-5f * kmax,
                     "hf_mg_set_cell_operator: the cell matrix is not symmetric under the cell's mid-plane "
                     "reflections (block (%d, %d) entry (%d, %d): %g, expected %g)", a, b, r, c, (double)want,
                     (double)got);
        }
  // block rows sum to zero (the rows are applied to neighbour differences)
  for (int a = 0; a < 8; ++a)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double sum = 0.0;
        for (int b = 0; b < 8; ++b) sum += k[(8 * a + b) * 9 + 3 * r + c];
        HF_REQUIRE(std::fabs(sum) <= 1e-5 * kmax, "hf_mg_set_cell_operator: block row %d of the cell matrix does "
                   "not sum to zero (entry (%d, %d): %g): not a translation-invariant operator", a, r, c, sum);
      }
  if (!mg->mf) HF_CHECK_HIP(hipMalloc((void**)&mg->mf, sizeof(MfOp)));
  if (!mg->mf_loaded || std::memcmp(mg->mf_host, h.b, sizeof(h.b)) != 0) {
    // a new cell matrix (once per geometry and law): no cycle may be reading the old one
    HF_CHECK_HIP(hipDeviceSynchronize());
    std::memcpy(mg->mf_host, h.b, sizeof(h.b));
    HF_CHECK_HIP(hipMemcpy(mg->mf, &h, sizeof(MfOp), hipMemcpyHostToDevice));
    mg->mf_loaded = true;
  }
  if (mg->mf_scale != cell_scale) ++mg->gen;   // a kernel argument of the captured batch
  mg->mf_scale = cell_scale;
  return 0;
}

// Set-up from the SELL tangent (vals) or from a half-stored level-0 stencil (S0).
static int mg_setup(hf_mg* mg, const hf_pattern* p, const float* vals, const float* S0, const float* mask,
                    int32_t* dropped, hf_stream_t stream) {
  HF_REQUIRE(mg && p && (vals || S0) && mask, "hf_mg_setup_f32: null argument");
  HF_REQUIRE(!S0 || (mg->half && !mg->sell), "hf_mg_setup_stencil_f32: the hierarchy does not keep a half-stored "
                                              "fine stencil (HF_MG_FINE_FULL / HF_MG_FINE_SELL, or one level)");
  HF_REQUIRE(pattern_vec(p) == 3, "hf_mg_setup_f32: vec=%d, the multigrid is built for vec=3", pattern_vec(p));
  const Lat g0 = mg->lv[0].g;
  HF_REQUIRE(pattern_nnode(p) == g0.n, "hf_mg_setup_f32: pattern has %lld nodes, lattice %lld",
             (long long)pattern_nnode(p), (long long)g0.n);
  hipStream_t s = (hipStream_t)stream;
  mg->ready = false;
  HF_CHECK_HIP(hipMemsetAsync(mg->flags, 0, 4 * sizeof(int32_t), s));
  const unsigned gb0 = (unsigned)ceil_div(g0.n, kMgBlock);
  float* lam_dev0 = mg->part + kMaxGrid;
  if (mg->sell) {
    // fine level = the SELL tangent: block inverses, bound and A P straight from its rows
    const Lat c = mg->lv[1].g;
    mg_sell_dinv<<<gb0, kMgBlock, 0, s>>>(pattern_dpos(p), vals, mask, g0.n, mg->lv[0].Dinv);
    const MatArgs ma = mat_args(p);
    const int gs = rows_grid(ma.nslices);
    mg_sell_gersh<<<gs, kBlock, 0, s>>>(ma, vals, mask, mg->lv[0].Dinv, g0.n, mg->part);
    mg_max_final<<<1, kMgBlock, 0, s>>>(mg->part, gs, lam_dev0);
    mg_sell_rap_ap<<<(unsigned)ma.nslices, 3 * kWave, 0, s>>>(ma, vals, mask, g0, c, mg->B, mg->flags);
    mg_rap_ptap<<<dim3((unsigned)ceil_div(c.n, kMgBlock), 27), kMgBlock, 0, s>>>(mg->B, g0, c, mg->lv[1].S);
    mg->sp = p;
    mg->svals = vals;
  }
  const bool check = !mg->sell && !(mg->checked == p && mg->checked_nblocks == pattern_nblocks(p));
  if (check)
    mg_check_lattice<<<dim3(gb0, 27), kMgBlock, 0, s>>>(pattern_rowptr(p), pattern_colidx(p), pattern_blkpos(p),
                                                        pattern_sell_pos(p), pattern_slice_ptr(p), pattern_R(p), g0,
                                                        mg->flags);
  // the FP32 fine stencil the set-up reads: with the FP16 copy serving the
  // smoother, the caller's own (no copy); otherwise the hierarchy's copy
  const float* S0p = (S0 && mg->f16) ? S0 : mg->lv[0].S;
  if (S0 && !mg->f16)
    HF_CHECK_HIP(hipMemcpyAsync(mg->lv[0].S, S0, (size_t)kHalfSlots * 9 * g0.n * sizeof(float),
                                hipMemcpyDeviceToDevice, s));
  else if (S0)
    HF_REQUIRE(((uintptr_t)S0 & 15) == 0, "hf_mg_setup_stencil_f32: stencil not 16-byte aligned");
  else if (mg->half)
    mg_fine_stencil<true><<<dim3(gb0, kHalfSlots), kMgBlock, 0, s>>>(pattern_rowptr(p), pattern_sell_pos(p),
                                                                     pattern_slice_ptr(p), pattern_R(p), vals, mask,
                                                                     g0, mg->lv[0].S);
  else if (!mg->sell)
    mg_fine_stencil<false><<<dim3(gb0, 27), kMgBlock, 0, s>>>(pattern_rowptr(p), pattern_sell_pos(p),
                                                              pattern_slice_ptr(p), pattern_R(p), vals, mask, g0,
                                                              mg->lv[0].S);
  float* lam_dev = lam_dev0;
  for (int l = mg->sell ? 1 : 0; l + 1 < mg->nlev; ++l) {
    const Lat f = mg->lv[l].g, c = mg->lv[l + 1].g;
    const bool glanes = f.n <= kLaneNodes;
    // the Gershgorin kernel's workgroups (partials)
    const int gl = glanes ? lane_grid(f.n) : (int)std::min<int64_t>(ceil_div(f.n, kMgBlock), kMaxGrid);
    auto level = [&](auto half_tag) {
      constexpr bool H = decltype(half_tag)::value;
      const float* Sp = l == 0 ? S0p : mg->lv[l].S;
      mg_block_dinv<H><<<(unsigned)ceil_div(f.n, kMgBlock), kMgBlock, 0, s>>>(Sp, l == 0 ? mask : nullptr, f.n,
                                                                              mg->lv[l].Dinv);
      auto gersh = glanes ? mg_gersh_lanes<H> : mg_gersh_nodes<H>;
      gersh<<<gl, kMgBlock, 0, s>>>(Sp, mg->lv[l].Dinv, f, l == 0 ? mask : nullptr, mg->part);
      // P^T A P: x, then y, then z; the two intermediates share the B buffer
      const Lat bx = make_lat(c.nx, f.ny, f.nz), bxy = make_lat(c.nx, c.ny, f.nz);
      float* T1 = mg->B;
      float* T2 = mg->B + (int64_t)kSt * bx.n;
      // the intermediates P_x^T A P_x and P_y^T (.) P_y are symmetric when A is: with a
      // half-stored level they are half-stored too (half the bytes written and re-read)
      mg_rap_dim<H><<<dim3((unsigned)ceil_div(bx.n, kMgBlock), 9), kMgBlock, 0, s>>>(Sp, f, bx, 0,
                                                                                   l == 0 ? mask : nullptr, T1, H);
      mg_rap_dim<H><<<dim3((unsigned)ceil_div(bxy.n, kMgBlock), 9), kMgBlock, 0, s>>>(T1, bx, bxy, 1, nullptr,
                                                                                     T2, H);
      mg_rap_dim<H><<<dim3((unsigned)ceil_div(c.n, kMgBlock), 9), kMgBlock, 0, s>>>(T2, bxy, c, 2, nullptr,
                                                                                  mg->lv[l + 1].S,
                                                                                  mg->lv[l + 1].half);
    };
    if (mg->lv[l].half) level(std::true_type{});
    else level(std::false_type{});
    mg_max_final<<<1, kMgBlock, 0, s>>>(mg->part, gl, lam_dev + l);
  }
  const Lat gl = mg->lv[mg->nlev - 1].g;
  mg_dense_setup<<<1, kDenseThreads, 0, s>>>(mg->lv[mg->nlev - 1].S, gl, mg->Ainv, mg->flags);
  if (mg_coefficients(mg, s)) return -2;
  // FP16 copy of a SELL fine level's values (scale in level 0's slot, unused by it)
  mg->svals16_ok = false;
  if (mg->f16 && mg->sell) {
    const int64_t nv = hf_matrix_values(p);
    if (nv > mg->svals16_cap) {
      if (mg->svals16) HF_CHECK_HIP(hipFree(mg->svals16));
      mg->svals16 = nullptr;
      mg->svals16_cap = 0;
      HF_CHECK_HIP(hipMalloc((void**)&mg->svals16, (size_t)nv * sizeof(__half)));
      mg->svals16_cap = nv;
      ++mg->gen;   // a kernel argument of the captured iteration
    }
    const int ga = (int)std::min<int64_t>(std::max<int64_t>(ceil_div(nv / 4, kMgBlock), 1), kMaxGrid);
    mg_absmax<<<ga, kMgBlock, 0, s>>>(vals, nv, mg->part);
    mg_max_final<<<1, kMgBlock, 0, s>>>(mg->part, ga, mg->scl + 2);
    mg_f16_scale<<<1, kWave, 0, s>>>(mg->scl);
    const int64_t nslots = nv / (kSellL * 3);
    mg_sell_to_f16<<<(unsigned)ceil_div(nslots * kSellL, kMgBlock), kMgBlock, 0, s>>>(vals, nslots, mg->scl,
                                                                                     mg->svals16);
    mg->svals16_ok = true;
  }
  // FP16 copies of the half-stored levels (level 0 only without a cell operator,
  // whose matrix-free rows replace its stencil in the cycle), all in two launches
  {
    F16Levels J;
    memset(&J, 0, sizeof(J));
    int np = 0, nc = 0;
    for (int l = 0; l < mg->nlev; ++l) {
      MgLevel& V = mg->lv[l];
      V.s16_ok = false;
      if (!V.S16 || (l == 0 && mg->mf_scale)) continue;
      const int j = J.nlev++;
      J.S[j] = l == 0 ? S0p : V.S;
      J.W[j] = V.S16;
      J.scl[j] = mg->scl + 4 * l;
      J.n[j] = V.g.n;
      J.p0[j] = np;
      J.c0[j] = nc;
      const int64_t cnt4 = (int64_t)kHalfSlots * 9 * V.g.n / 4;
      np += (int)std::min<int64_t>(std::max<int64_t>(ceil_div(cnt4, kMgBlock), 1), kF16MaxParts);
      nc += (int)ceil_div((int64_t)kHalfSlots * V.g.n, kMgBlock);
      V.s16_ok = true;
    }
    J.p0[J.nlev] = np;
    J.c0[J.nlev] = nc;
    if (J.nlev) {
      mg_f16_max_all<<<np, kMgBlock, 0, s>>>(J, mg->part);
      mg_f16_convert_all<<<nc, kMgBlock, 0, s>>>(J, mg->part);
    }
  }
  if (mg->lmode) {
    const MgLevel& F = mg->lv[0];
    if (mg->sell) {
      mg_sell_scalar_dinv<<<(unsigned)ceil_div(F.g.n, kMgBlock), kMgBlock, 0, s>>>(pattern_dpos(p), vals, mask,
                                                                                   F.g.n, mg->sdinv);
    } else {
      auto sd = F.half ? mg_scalar_dinv<true> : mg_scalar_dinv<false>;
      sd<<<(unsigned)ceil_div(F.g.n, kMgBlock), kMgBlock, 0, s>>>(S0p, mask, F.g.n, mg->sdinv);
    }
  }
  mg->s0 = mg->sell ? nullptr : S0p;
  HF_LAUNCH_CHECK();
  mg->mask = mask;
  // The host waits for the set-up only to check it: a pattern's first set-up (its
  // lattice structure), a SELL fine level (couplings beyond one coarse step), or a
  // caller asking for the dropped pivot count.  Otherwise it returns at once and
  // the solve's launches queue behind the set-up.
  if (!(check || mg->sell || dropped)) {
    mg->ready = true;
    return 0;
  }
  int32_t hflags[4] = {0, 0, 0, 0};
  float lam[kMgMaxLevels] = {0};
  HF_CHECK_HIP(hipMemcpyAsync(hflags, mg->flags, sizeof(hflags), hipMemcpyDeviceToHost, s));
  if (mg->nlev > 1)
    HF_CHECK_HIP(hipMemcpyAsync(lam, lam_dev, (mg->nlev - 1) * sizeof(float), hipMemcpyDeviceToHost, s));
  HF_CHECK_HIP(hipStreamSynchronize(s));
  HF_REQUIRE(hflags[0] == 0, mg->sell ? "hf_mg_setup_f32: the pattern couples lattice nodes whose coarse parents "
                                        "are more than one coarse step apart (not a HEX8/HEX27 box lattice)"
                                      : "hf_mg_setup_f32: the pattern couples nodes that are not lattice neighbours "
                                        "(the multigrid needs a HEX8 box lattice)");
  if (!mg->sell) {
    mg->checked = p;
    mg->checked_nblocks = pattern_nblocks(p);
  }
  for (int l = 0; l + 1 < mg->nlev; ++l)
    HF_REQUIRE(std::isfinite(lam[l]), "hf_mg_setup_f32: non-finite operator on level %d", l);
  if (dropped) *dropped = hflags[1];
  mg->ready = true;
  return 0;
}

extern "C" int hf_mg_setup_f32(hf_mg* mg, const hf_pattern* p, const float* vals, const float* mask,
                               int32_t* dropped, hf_stream_t stream) {
  HF_REQUIRE(vals, "hf_mg_setup_f32: null argument");
  return mg_setup(mg, p, vals, nullptr, mask, dropped, stream);
}

extern "C" int hf_mg_setup_stencil_f32(hf_mg* mg, const hf_pattern* p, const float* S, const float* mask,
                                       int32_t* dropped, hf_stream_t stream) {
  HF_REQUIRE(S, "hf_mg_setup_stencil_f32: null argument");
  return mg_setup(mg, p, nullptr, S, mask, dropped, stream);
}

extern "C" size_t hf_mg_stencil_size(const hf_mg* mg) {
  return mg ? (size_t)kHalfSlots * 9 * mg->lv[0].g.n : 0;
}

extern "C" int hf_mg_assemble_hex8_f32(const hf_mg* mg, const hf_pattern* p, const int32_t* cells, int64_t ncell,
                                       const int32_t* corner, const float* elem, float* S, hf_stream_t stream) {
  HF_REQUIRE(mg && p && cells && elem && S, "hf_mg_assemble_hex8_f32: null argument");
  const Lat g = mg->lv[0].g;
  HF_REQUIRE(pattern_vec(p) == 3 && pattern_nn(p) == 8 && pattern_nnode(p) == g.n,
             "hf_mg_assemble_hex8_f32: needs the HEX8 vec-3 pattern of the hierarchy's lattice");
  HF_REQUIRE(g.n < INT32_MAX, "hf_mg_assemble_hex8_f32: lattice too large");
  const dim3 grid((unsigned)ceil_div(g.n, kMgBlock), kHalfSlots);
  if (corner) {
    HF_REQUIRE(ncell == (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1),
               "hf_mg_assemble_hex8_f32: %lld cells are not the lattice's", (long long)ncell);
    CornerTable ct;
    bool seen[8] = {false, false, false, false, false, false, false, false};
    for (int q = 0; q < 8; ++q) {
      HF_REQUIRE(corner[q] >= 0 && corner[q] < 8 && !seen[corner[q]],
                 "hf_mg_assemble_hex8_f32: corner table is not a permutation of 0..7");
      seen[corner[q]] = true;
      ct.idx[q] = (int8_t)corner[q];
    }
    mg_assemble_hex8_box<<<grid, kMgBlock, 0, (hipStream_t)stream>>>(g, ncell, ct, elem, S);
  } else {
    mg_assemble_hex8<<<grid, kMgBlock, 0, (hipStream_t)stream>>>(pattern_diag(p), pattern_seg(p),
                                                                pattern_contrib(p), cells, g, ncell, elem, S);
  }
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_mg_assemble_cellop_f32(const hf_mg* mg, float* S, hf_stream_t stream) {
  HF_REQUIRE(mg && S, "hf_mg_assemble_cellop_f32: null argument");
  HF_REQUIRE(mg->mf_scale, "hf_mg_assemble_cellop_f32: no cell operator (hf_mg_set_cell_operator)");
  const Lat g = mg->lv[0].g;
  mg_assemble_cellop<<<dim3((unsigned)ceil_div(g.n, kMgBlock), kHalfSlots), kMgBlock, 0, (hipStream_t)stream>>>(
      mg->mf, mg->mf_scale, g, S);
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_mg_stencil_matvec_f32(const hf_mg* mg, const float* S, const float* x, float* y,
                                        hf_stream_t stream) {
  HF_REQUIRE(mg && x && y, "hf_mg_stencil_matvec_f32: null argument");
  HF_REQUIRE(S || mg->mf_scale, "hf_mg_stencil_matvec_f32: no stencil and no cell operator");
  HF_REQUIRE(mg->mask, "hf_mg_stencil_matvec_f32: no set-up (the Dirichlet mask comes from it)");
  const Lat g = mg->lv[0].g;
  if (S)
    mg_stencil_matvec<<<node_grid(g.n), kMgBlock, 0, (hipStream_t)stream>>>(S, nullptr, g, mg->mask, x, y);
  else
    mg_stencil_matvec<<<node_grid(g.n), kMgBlock, 0, (hipStream_t)stream>>>(mg->mf, mg->mf_scale, g, mg->mask, x,
                                                                             y);
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_mg_bounds(const hf_mg* mg, float* lam, int32_t max_levels) {
  HF_REQUIRE(mg && lam, "hf_mg_bounds: null argument");
  HF_REQUIRE(mg->ready, "hf_mg_bounds: hf_mg_setup_f32 has not run");
  float dev[kMgMaxLevels] = {0};
  // the set-up may still run on the caller's stream
  HF_CHECK_HIP(hipDeviceSynchronize());
  if (mg->nlev > 1)
    HF_CHECK_HIP(hipMemcpy(dev, mg->part + kMaxGrid, (mg->nlev - 1) * sizeof(float), hipMemcpyDeviceToHost));
  for (int l = 0; l < mg->nlev && l < max_levels; ++l) lam[l] = dev[l];
  return 0;
}

// Enqueue one V-cycle: level-0 input r0, result in *out (one of the two level-0
// buffers z0 / z1 given by the caller).  Levels with more than kSmallNodes
// nodes run as grid kernels; the rest (from level lb down) in mg_small_cycle.
// timed: the level-0 pre-smoother's first stencil step (mg_poly_step, one pass
// over the fine stencil; the SELL product on a SELL fine level) is launched
// between the two profiling events (HIP-event timing for roofline reporting,
// see hf_profile_enable).
// first_done: the fine level's pre-smoother start (d = e = s0 Dinv r) was
// written by the caller's kernel (the outer iteration's update)
static int vcycle(hf_mg* mg, const float* r0, float* z0, float* z1, float** out, const Scalars* sc, hipStream_t s,
                  bool timed = false, bool skip_fine_post = false, bool first_done = false) {
  const int L = mg->nlev;
  const int k = mg->degree;
  // post-smoothing: the first kp steps of the same recurrence (the Chebyshev
  // polynomial of degree kp on the same interval: its coefficients depend only on
  // the step, not on the degree)
  const int kp = mg->degree_post > 0 ? mg->degree_post : k;
  float* cur[kMgMaxLevels];
  float* oth[kMgMaxLevels];
  for (int l = 0; l < L; ++l) {
    cur[l] = l == 0 ? z0 : mg->lv[l].e;
    oth[l] = l == 0 ? z1 : mg->lv[l].e2;
  }
  const int lb = small_cycle_start(mg);
  auto rin = [&](int l) -> const float* { return l == 0 ? r0 : mg->lv[l].r; };
  for (int l = 0; l < lb; ++l) {
    const MgLevel& V = mg->lv[l];
    const int gg = node_grid(V.g.n);
    // pre-smoothing from e = 0 (its start on levels >= 1 is fused into the
    // restriction that produced rin(l)), then the residual for the coarse level
    const float* cf = coef_of(mg, l);
    if (l == 0 && !first_done) mg_poly_first<<<gg, kMgBlock, 0, s>>>(V.Dinv, V.g.n, cf, rin(l), V.da, cur[l], sc);
    const float* res = rin(l);
    float* din = V.da;
    float* dout = V.db;
    const bool lanes = V.g.n <= kLaneNodes;
    const int gl = lane_grid(V.g.n);
    int derr = 0;
    // Sx: the level's stencil, FP32 or (the fine level with HF_MG_FINE_F16) FP16
    // words with scale scl
    auto down = [&](auto half_tag, const auto* Sx, const float* scl) {
      constexpr bool H = decltype(half_tag)::value;
      using ST = std::remove_cv_t<std::remove_pointer_t<decltype(Sx)>>;
      for (int j = 1; j < k; ++j) {
        if (timed && l == 0 && j == 1) {
          hipEvent_t e0, e1;
          if (prof_events(&e0, &e1)) { derr = -2; return; }
          auto step = lanes ? mg_poly_step_lanes<H, ST> : mg_poly_step<H, ST>;
          hipExtLaunchKernelGGL(step, dim3(lanes ? gl : gg), dim3(kMgBlock), 0, s, e0, e1, 0, Sx, scl,
                                (const float*)V.Dinv, V.g, cf, j, res, (const float*)din, V.res, dout, cur[l], sc);
        } else if (lanes) {
          mg_poly_step_lanes<H, ST><<<gl, kMgBlock, 0, s>>>(Sx, scl, V.Dinv, V.g, cf, j, res, din, V.res, dout,
                                                            cur[l], sc);
        } else {
          mg_poly_step<H, ST><<<gg, kMgBlock, 0, s>>>(Sx, scl, V.Dinv, V.g, cf, j, res, din, V.res, dout,
                                                      cur[l], sc);
        }
        res = V.res;
        float* tmp = din; din = dout; dout = tmp;
      }
      const float* rmask = l == 0 ? mg->mask : nullptr;
      if (lanes)
        mg_poly_resid_lanes<H, ST><<<gl, kMgBlock, 0, s>>>(Sx, scl, V.g, res, din, rmask, dout, sc);
      else
        mg_poly_resid<H, ST><<<gg, kMgBlock, 0, s>>>(Sx, scl, V.g, res, din, rmask, dout, sc);
    };
    if (l == 0 && mg->sell) {
      // SELL fine level: each pass is y = A v, then the node-wise update
      const MatArgs m = mat_args(mg->sp);
      const int gr = rows_grid(m.nslices);
      // y = A v with the FP16 values when the set-up made them
      auto sell_spmv = [&](const float* vv) {
        if (mg->svals16_ok)
          mg_sell_spmv16<<<gr, kBlock, 0, s>>>(m, mg->svals16, mg->scl, vv, mg->y, sc);
        else
          mg_sell_spmv<<<gr, kBlock, 0, s>>>(m, mg->svals, vv, mg->y, sc);
      };
      for (int j = 1; j < k; ++j) {
        sell_spmv(din);
        mg_sell_step_update<<<gg, kMgBlock, 0, s>>>(V.Dinv, V.g.n, cf, j, res, mg->y, din, V.res, dout,
                                                    cur[l], sc);
        res = V.res;
        float* tmp = din; din = dout; dout = tmp;
      }
      sell_spmv(din);
      mg_sell_resid_update<<<gg, kMgBlock, 0, s>>>(V.g.n, res, mg->y, dout, sc);
    } else if (l == 0 && mg->mf_scale) {
      down(std::true_type{}, (const MfOp*)mg->mf, mg->mf_scale);
    } else if (V.s16_ok) {
      down(std::true_type{}, (const __half2*)V.S16, (const float*)(mg->scl + 4 * l));
    } else if (V.half) {
      down(std::true_type{}, (const float*)V.S, (const float*)nullptr);
    } else {
      down(std::false_type{}, (const float*)V.S, (const float*)nullptr);
    }
    if (derr) return derr;
    const MgLevel& C = mg->lv[l + 1];
    const bool smooth_c = (l + 2 < L);
    mg_restrict<<<node_grid(C.g.n), kMgBlock, 0, s>>>(dout, V.g, C.g, C.r, smooth_c ? C.Dinv : nullptr, coef_of(mg, l + 1),
                                                      C.da, cur[l + 1], sc);
  }
  if (L == 1) {
    mg_dense_apply<<<1, kDenseApplyThreads, 0, s>>>(mg->Ainv, mg->nd, rin(0), cur[0], sc);
  } else {
    SmallCycle c;
    memset(&c, 0, sizeof(c));
    c.nlev = L - lb;
    c.degree = k;
    c.Ainv = mg->Ainv;
    c.nd = mg->nd;
    const MgLevel& F = mg->lv[lb];
    c.r0 = F.r;
    c.e0 = F.e;
    c.d0 = F.da;
    c.out = (L - lb >= 2) ? F.e2 : F.e;
    int64_t at = 0;   // arena words
    for (int l = lb; l < L; ++l) {
      const MgLevel& V = mg->lv[l];
      SmallLevel& q = c.lv[l - lb];
      const bool dense = l == L - 1;
      q.g = V.g;
      q.kind = V.s16_ok ? kSmallF16 : V.half ? kSmallHalf : kSmallFull;
      q.words = dense ? 0 : (int)(small_level_words(q.kind, V.g.n, false) - 34 * V.g.n - 12);
      q.S = V.s16_ok ? (const void*)V.S16 : (const void*)V.S;
      q.scl = mg->scl ? mg->scl + 4 * l : nullptr;
      q.Dinv = V.Dinv;
      q.cf = coef_of(mg, l);
      // every array 16-byte aligned (the staging copies 16-byte words)
      auto up4 = [](int64_t w) { return (w + 3) & ~(int64_t)3; };
      q.oS = (int)at;
      q.oD = (int)up4(at + q.words);
      q.oC = dense ? (int)at : (int)up4(q.oD + 9 * V.g.n);
      q.oV = (int)up4(q.oC + V.g.n);
      at = up4(q.oV + (dense ? 8 : 24) * V.g.n);
    }
    HF_REQUIRE(at <= kSmallArenaWords, "vcycle: small-cycle levels need %lld LDS words, the arena holds %d",
               (long long)at, kSmallArenaWords);
    mg_small_cycle<<<1, kSmallThreads, 0, s>>>(c, sc);
    cur[lb] = (L - lb >= 2) ? mg->lv[lb].e2 : mg->lv[lb].e;
  }
  for (int l = lb - 1; l >= 0; --l) {
    const MgLevel& V = mg->lv[l];
    const MgLevel& C = mg->lv[l + 1];
    const float* cf = coef_of(mg, l);
    const int gg = node_grid(V.g.n);
    mg_prolong<<<gg, kMgBlock, 0, s>>>(cur[l + 1], V.g, C.g, l == 0 ? mg->mask : nullptr, cur[l], sc);
    if (l == 0 && skip_fine_post) break;
    // the learned step: res = M (r - A e), then e += D^-1 res + s D^-1/2 mask N(res)
    auto learned_step = [&]() -> int {
      if (mg->sell) {
        const MatArgs m = mat_args(mg->sp);
        const int gr = rows_grid(m.nslices);
        if (mg->svals16_ok)
          mg_sell_spmv16<<<gr, kBlock, 0, s>>>(m, mg->svals16, mg->scl, cur[0], mg->y, sc);
        else
          mg_sell_spmv<<<gr, kBlock, 0, s>>>(m, mg->svals, cur[0], mg->y, sc);
        mg_sell_masked_resid<<<gg, kMgBlock, 0, s>>>(V.g.n, rin(0), mg->y, mg->mask, V.res, sc);
      } else if (mg->mf_scale) {
        mg_poly_resid<true, MfOp><<<gg, kMgBlock, 0, s>>>(mg->mf, mg->mf_scale, V.g, rin(0), cur[0], mg->mask,
                                                          V.res, sc);
      } else if (V.s16_ok) {
        mg_poly_resid<true, __half2><<<gg, kMgBlock, 0, s>>>(V.S16, mg->scl, V.g, rin(0), cur[0], mg->mask, V.res,
                                                             sc);
      } else {
        auto resid = V.half ? mg_poly_resid<true> : mg_poly_resid<false>;
        resid<<<gg, kMgBlock, 0, s>>>(V.S, nullptr, V.g, rin(0), cur[0], mg->mask, V.res, sc);
      }
      return lpc_smoother_apply(mg->lsm, V.res, mg->sdinv, mg->mask, cur[0], sc ? &sc->done : nullptr, s);
    };
    if (l == 0 && mg->lmode == 1) {
      if (learned_step()) return -2;
      continue;
    }
    // post-smoothing: the same polynomial on the residual after the correction
    int perr = 0;
    const bool lanes = V.g.n <= kLaneNodes;
    const int gl = lane_grid(V.g.n);
    auto up = [&](auto half_tag, const auto* Sx, const float* scl) {
      constexpr bool H = decltype(half_tag)::value;
      using ST = std::remove_cv_t<std::remove_pointer_t<decltype(Sx)>>;
      auto post = lanes ? mg_post_first_lanes<H, ST> : mg_post_first<H, ST>;
      const int gp = lanes ? gl : gg;
      post<<<gp, kMgBlock, 0, s>>>(Sx, scl, V.Dinv, V.g, cf, rin(l), cur[l], V.res, V.da, oth[l], sc);
      float* tmp = cur[l]; cur[l] = oth[l]; oth[l] = tmp;
      float* din = V.da;
      float* dout = V.db;
      for (int j = 1; j < kp; ++j) {
        if (lanes)
          mg_poly_step_lanes<H, ST><<<gl, kMgBlock, 0, s>>>(Sx, scl, V.Dinv, V.g, cf, j, V.res, din, V.res, dout,
                                                            cur[l], sc);
        else
          mg_poly_step<H, ST><<<gg, kMgBlock, 0, s>>>(Sx, scl, V.Dinv, V.g, cf, j, V.res, din, V.res, dout,
                                                      cur[l], sc);
        float* t2 = din; din = dout; dout = t2;
      }
    };
    if (l == 0 && mg->sell) {
      const MatArgs m = mat_args(mg->sp);
      const int gr = rows_grid(m.nslices);
      // y = A v with the FP16 values when the set-up made them
      auto sell_spmv = [&](const float* vv) {
        if (mg->svals16_ok)
          mg_sell_spmv16<<<gr, kBlock, 0, s>>>(m, mg->svals16, mg->scl, vv, mg->y, sc);
        else
          mg_sell_spmv<<<gr, kBlock, 0, s>>>(m, mg->svals, vv, mg->y, sc);
      };
      if (timed) {
        hipEvent_t e0, e1;
        if (prof_events(&e0, &e1)) return -2;
        if (mg->svals16_ok)
          hipExtLaunchKernelGGL(mg_sell_spmv16, dim3(gr), dim3(kBlock), 0, s, e0, e1, 0, m,
                                (const __half*)mg->svals16, (const float*)mg->scl, (const float*)cur[l], mg->y, sc);
        else
          hipExtLaunchKernelGGL(mg_sell_spmv, dim3(gr), dim3(kBlock), 0, s, e0, e1, 0, m, mg->svals,
                                (const float*)cur[l], mg->y, sc);
      } else {
        sell_spmv(cur[l]);
      }
      mg_sell_post_update<<<gg, kMgBlock, 0, s>>>(V.Dinv, V.g.n, cf, rin(l), mg->y, cur[l], V.res, V.da, oth[l],
                                                  sc);
      float* tmp = cur[l]; cur[l] = oth[l]; oth[l] = tmp;
      float* din = V.da;
      float* dout = V.db;
      for (int j = 1; j < kp; ++j) {
        sell_spmv(din);
        mg_sell_step_update<<<gg, kMgBlock, 0, s>>>(V.Dinv, V.g.n, cf, j, V.res, mg->y, din, V.res, dout,
                                                    cur[l], sc);
        float* t2 = din; din = dout; dout = t2;
      }
    } else if (l == 0 && mg->mf_scale) {
      up(std::true_type{}, (const MfOp*)mg->mf, mg->mf_scale);
    } else if (V.s16_ok) {
      up(std::true_type{}, (const __half2*)V.S16, (const float*)(mg->scl + 4 * l));
    } else if (V.half) {
      up(std::true_type{}, (const float*)V.S, (const float*)nullptr);
    } else {
      up(std::false_type{}, (const float*)V.S, (const float*)nullptr);
    }
    if (perr) return perr;
    if (l == 0 && mg->lmode == 2 && learned_step()) return -2;
  }
  HF_LAUNCH_CHECK();
  *out = cur[0];
  return 0;
}

extern "C" int hf_mg_set_learned(hf_mg* mg, int32_t mode, const float* const* weights, const float* gfield,
                                 hf_stream_t stream) {
  HF_REQUIRE(mg, "hf_mg_set_learned: null hierarchy");
  HF_REQUIRE(mode >= 0 && mode <= 2, "hf_mg_set_learned: mode %d (0 off, 1 replace, 2 after)", mode);
  hipStream_t s = (hipStream_t)stream;
  if (mode == 0) {
    if (mg->lmode) ++mg->gen;
    mg->lmode = 0;
    return 0;
  }
  HF_REQUIRE(weights && gfield, "hf_mg_set_learned: null argument");
  HF_REQUIRE(mg->nlev > 1, "hf_mg_set_learned: needs a coarser level");
  const MgLevel& F = mg->lv[0];
  if (!mg->lsm.buf) {
    HF_CHECK_HIP(hipMalloc(&mg->lsm.buf, lpc_smoother_bytes(F.g.n)));
    HF_CHECK_HIP(hipMalloc((void**)&mg->sdinv, (size_t)F.g.n * 3 * sizeof(float)));
  }
  mg->lsm.nx = F.g.nx;
  mg->lsm.ny = F.g.ny;
  mg->lsm.nz = F.g.nz;
  mg->lsm.gfield = gfield;
  if (lpc_smoother_load(mg->lsm, weights, s)) return -2;
  if (mg->ready) {
    if (mg->sell) {
      mg_sell_scalar_dinv<<<(unsigned)ceil_div(F.g.n, kMgBlock), kMgBlock, 0, s>>>(pattern_dpos(mg->sp), mg->svals,
                                                                                   mg->mask, F.g.n, mg->sdinv);
    } else {
      auto sd = F.half ? mg_scalar_dinv<true> : mg_scalar_dinv<false>;
      sd<<<(unsigned)ceil_div(F.g.n, kMgBlock), kMgBlock, 0, s>>>(mg->s0, mg->mask, F.g.n, mg->sdinv);
    }
    HF_LAUNCH_CHECK();
  }
  mg->lmode = mode;
  ++mg->gen;   // the captured solve batch has another shape
  return 0;
}

extern "C" int hf_mg_apply_f32(hf_mg* mg, const float* r, float* z, int32_t flags, hf_stream_t stream) {
  HF_REQUIRE(mg && r && z, "hf_mg_apply_f32: null argument");
  HF_REQUIRE(mg->ready, "hf_mg_apply_f32: hf_mg_setup_f32 has not run");
  HF_REQUIRE((flags & ~HF_MG_SKIP_FINE_POST) == 0, "hf_mg_apply_f32: unknown flags 0x%x", flags);
  hipStream_t s = (hipStream_t)stream;
  float* res = nullptr;
  if (vcycle(mg, r, z, mg->lv[0].e2, &res, nullptr, s, false, (flags & HF_MG_SKIP_FINE_POST) != 0)) return -2;
  if (res != z) HF_CHECK_HIP(hipMemcpyAsync(z, res, (size_t)mg->lv[0].g.n * 3 * sizeof(float),
                                            hipMemcpyDeviceToDevice, s));
  return 0;
}

extern "C" int hf_mg_copy_stencil(const hf_mg* mg, int32_t level, float* dst, hf_stream_t stream) {
  HF_REQUIRE(mg && dst, "hf_mg_copy_stencil: null argument");
  HF_REQUIRE(level >= 0 && level < mg->nlev, "hf_mg_copy_stencil: level %d outside 0..%d", level, mg->nlev - 1);
  HF_REQUIRE(mg->ready, "hf_mg_copy_stencil: hf_mg_setup_f32 has not run");
  HF_REQUIRE(!(level == 0 && mg->sell), "hf_mg_copy_stencil: the fine level is the tangent itself (SELL)");
  const int64_t nv = (int64_t)kSt * mg->lv[level].g.n;
  // the fine level: the FP32 stencil of the last set-up (the caller's buffer when it gave one)
  const float* src = level == 0 ? mg->s0 : mg->lv[level].S;
  if (mg->lv[level].half) {
    const Lat gl = mg->lv[level].g;
    mg_expand_half<<<dim3((unsigned)ceil_div(gl.n, kMgBlock), 27), kMgBlock, 0, (hipStream_t)stream>>>(
        src, gl, dst);
    HF_LAUNCH_CHECK();
  } else {
    HF_CHECK_HIP(hipMemcpyAsync(dst, src, (size_t)nv * sizeof(float), hipMemcpyDeviceToDevice,
                                (hipStream_t)stream));
  }
  return 0;
}

extern "C" size_t hf_mg_workspace_size(const hf_mg* mg) {
  if (!mg) return 0;
  return kMgHeader + 7 * align256((size_t)mg->lv[0].g.n * 3 * sizeof(float));
}

// Enqueue `poll` FCG iterations (update, convergence scalars, V-cycle, w = A z,
// direction scalars) on stream s; timed marks one launch for HIP-event timing.
struct MgBatch {
  MatArgs m;
  const float* vals;  // the SELL tangent, or
  const float* S;     // the half-stored lattice stencil (Dirichlet rows masked on the fly)
  const MfOp* mf;     // or the matrix-free cell operator with per-cell scales mfs
  const float* mfs;
  const float* mask;
  Lat g;
  int64_t N;
  float *r, *z0, *z1, *w, *pv, *q, *x;
  float* z;  // the V-cycle's output buffer (the same one every cycle)
  bool split;  // direction scalars by their own one-workgroup kernel at every grid size
  Scalars* sc;
  double *partA, *partB;
  int gr, ge;
};
static int mg_enqueue_batch(hf_mg* mg, MgBatch& B, int poll, hipStream_t s, int timed_k) {
  const MgLevel& F = mg->lv[0];
  // with one level the V-cycle is the dense solve of r: no smoother start to fuse
  const bool fuse = mg->nlev > 1;
  const bool dir_fused = B.gr <= kFusedParts && !B.split;
  for (int k = 0; k < poll; ++k) {
    if (!dir_fused) mgcg_dir_scalars<<<1, kScalarBlock, 0, s>>>(B.sc, B.partB, B.gr);
    mgcg_update_first<<<B.ge, kMgBlock, 0, s>>>(F.g.n, B.z, B.w, B.pv, B.q, B.x, B.r, fuse ? F.Dinv : nullptr,
                                                coef_of(mg, 0), F.da, B.z0, B.sc, B.partB,
                                                dir_fused ? B.gr : 0, B.partA);
    mgcg_conv_scalars<<<1, kScalarBlock, 0, s>>>(B.sc, B.partA, B.ge);
    float* z = nullptr;
    if (vcycle(mg, B.r, B.z0, B.z1, &z, B.sc, s, k == timed_k, false, fuse)) return -2;
    B.z = z;
    if (B.mf)
      mgcg_spmv_stencil<<<B.gr, kMgBlock, 0, s>>>(B.mf, B.mfs, B.g, B.mask, B.z, B.r, B.q, B.w, B.sc, B.partB);
    else if (B.S)
      mgcg_spmv_stencil<<<B.gr, kMgBlock, 0, s>>>(B.S, nullptr, B.g, B.mask, B.z, B.r, B.q, B.w, B.sc, B.partB);
    else
      mgcg_spmv_kernel<<<B.gr, kBlock, 0, s>>>(B.m, B.vals, B.z, B.r, B.q, B.w, B.sc, B.partB);
  }
  return 0;
}

// Capture a poll batch into mg->gexec unless the cached one matches.  The key
// holds every value the captured launches depend on besides device memory
// contents: the set-up generation (level coefficients are kernel arguments),
// the tangent's values and pattern, the workspace and the batch length.
static int mg_batch_graph(hf_mg* mg, const hf_pattern* p, MgBatch& B, void* ws, int poll) {
  const uintptr_t key[8] = {(uintptr_t)mg->gen, (uintptr_t)p, (uintptr_t)(B.S ? B.S : B.vals), (uintptr_t)ws,
                            (uintptr_t)poll, (uintptr_t)B.N, (uintptr_t)mg->mask, (uintptr_t)B.split};
  bool same = mg->gexec != nullptr;
  for (int i = 0; i < 8 && same; ++i) same = mg->gkey[i] == key[i];
  if (same) return 0;
  if (!mg->cap) HF_CHECK_HIP(hipStreamCreateWithFlags(&mg->cap, hipStreamNonBlocking));
  if (mg->gexec) {
    (void)hipGraphExecDestroy(mg->gexec);
    mg->gexec = nullptr;
  }
  HF_CHECK_HIP(hipStreamBeginCapture(mg->cap, hipStreamCaptureModeThreadLocal));
  const int rc = mg_enqueue_batch(mg, B, poll, mg->cap, -1);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(mg->cap, &g);
  HF_REQUIRE(rc == 0 && e == hipSuccess && g, "hf_mg_solve_f32: capturing the iteration failed (%s)",
             hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&mg->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HF_CHECK_HIP(ei);
  for (int i = 0; i < 8; ++i) mg->gkey[i] = key[i];
  return 0;
}

// Flexible CG with the SELL tangent (p, vals) or the half-stored stencil S as the operator.
static int mg_solve(hf_mg* mg, const hf_pattern* p, const float* vals, const float* S, const float* b, float* x,
                    int32_t flags, int32_t maxiter, double rtol, double atol, int32_t poll, void* ws,
                    hf_krylov_info* info, hipStream_t s) {
  HF_REQUIRE(mg->ready, "hf_mg_solve_f32: hf_mg_setup_f32 has not run");
  HF_REQUIRE(!(flags & HF_KRYLOV_TRANSPOSE), "hf_mg_solve_f32: transposed solves are not supported");
  HF_REQUIRE(maxiter >= 0, "hf_mg_solve_f32: maxiter < 0");
  if (poll < 1) poll = 1;
  const int rel_r0 = (flags & HF_KRYLOV_REL_R0) ? 1 : 0;
  const MatArgs m = S ? MatArgs{} : mat_args(p);
  const int64_t N = mg->lv[0].g.n * 3;
  Scalars* sc = (Scalars*)ws;
  double* partA = (double*)((char*)ws + 4096);
  double* partB = partA + kPartB;
  float* v = (float*)((char*)ws + kMgHeader);
  const size_t stride = align256(N * sizeof(float)) / sizeof(float);
  MgBatch B;
  B.m = m;
  B.vals = vals;
  B.S = S;
  // the cell operator replaces the lattice stencil's products (a SELL fine level keeps its own)
  const bool mf = mg->mf_scale && !mg->sell;
  B.mf = mf ? mg->mf : nullptr;
  B.mfs = mf ? mg->mf_scale : nullptr;
  B.mask = mg->mask;
  B.g = mg->lv[0].g;
  B.N = N;
  B.r = v; B.z0 = v + stride; B.z1 = v + 2 * stride; B.w = v + 3 * stride; B.pv = v + 4 * stride;
  B.q = v + 5 * stride;
  B.x = v + 6 * stride;  // the iterate lives in the workspace, so a captured batch outlives the caller's x
  B.split = (flags & HF_MG_SOLVE_SPLIT_SCALARS) != 0;
  B.sc = sc;
  B.partA = partA;
  B.partB = partB;
  B.gr = (S || mf) ? node_grid(mg->lv[0].g.n) : rows_grid(m.nslices);  // the operator kernels' grid (= partials)
  B.ge = node_grid(mg->lv[0].g.n);  // the update's grid (= its partial count)
  const bool graph = !(flags & HF_MG_SOLVE_EAGER);
  HF_CHECK_HIP(hipMemcpyAsync(B.x, x, N * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (mf)
    mgcg_init_stencil<<<B.gr, kMgBlock, 0, s>>>(B.mf, B.mfs, B.g, B.mask, b, B.x, B.r, B.pv, B.q, partA);
  else if (S)
    mgcg_init_stencil<<<B.gr, kMgBlock, 0, s>>>(S, nullptr, B.g, B.mask, b, B.x, B.r, B.pv, B.q, partA);
  else
    mgcg_init_kernel<<<B.gr, kBlock, 0, s>>>(m, vals, b, B.x, B.r, B.pv, B.q, partA);
  mgcg_init_scalars<<<1, kScalarBlock, 0, s>>>(sc, partA, B.gr, rtol * rtol, atol * atol, rel_r0, maxiter);
  float* z = nullptr;
  if (vcycle(mg, B.r, B.z0, B.z1, &z, sc, s)) return -2;
  B.z = z;
  if (mf)
    mgcg_spmv_stencil<<<B.gr, kMgBlock, 0, s>>>(B.mf, B.mfs, B.g, B.mask, B.z, B.r, B.q, B.w, sc, partB);
  else if (S)
    mgcg_spmv_stencil<<<B.gr, kMgBlock, 0, s>>>(S, nullptr, B.g, B.mask, B.z, B.r, B.q, B.w, sc, partB);
  else
    mgcg_spmv_kernel<<<B.gr, kBlock, 0, s>>>(m, vals, B.z, B.r, B.q, B.w, sc, partB);
  HF_LAUNCH_CHECK();
  if (poll_init()) return -2;
  Scalars host{};
  int it = 0, batch = 0;
  bool timed_pending = false;
  int timed_batch = -1;
  double rr_prev = -1.0;
  for (;;) {
    // the batch holding the HIP-event timed launch (a steady-state one) runs
    // eagerly; the others replay the captured batch
    // (a timed launch exists when the fine level is SELL or smooths with degree >= 2)
    const bool timed = g_prof && mg->nlev > 1 && (mg->sell || mg->degree >= 2) && !timed_pending;
    if (timed || !graph) {
      if (mg_enqueue_batch(mg, B, poll, s, timed ? poll / 2 : -1)) return -2;
      if (timed) {
        timed_pending = true;
        timed_batch = batch;
      }
    } else {
      if (mg_batch_graph(mg, p, B, ws, poll)) return -2;
      HF_CHECK_HIP(hipGraphLaunch(mg->gexec, s));
    }
    it += poll;
    HF_LAUNCH_CHECK();
    if (poll_post(sc, batch & 1, s)) return -2;
    if (batch > 0) {
      if (poll_wait((batch - 1) & 1, host)) return -2;
      if (timed_pending && timed_batch == batch - 1) {
        // the timed launch did work unless the solve had finished before it
        if (!host.done || host.iter > (batch - 1) * poll + poll / 2) {
          if (prof_collect()) return -2;
        }
        timed_pending = false;
      }
      if (host.done) break;
      // When the last batch's reduction of r.r, carried forward, takes the batch just
      // queued below the threshold, wait for that batch before queuing another: a
      // batch queued behind a converged one only launches kernels that exit at once.
      const double q = rr_prev > 0.0 ? host.rr / rr_prev : 1.0;
      rr_prev = host.rr;
      if (!timed_pending && q < 1.0 && host.rr * q <= host.thresh) {
        if (poll_wait(batch & 1, host)) return -2;
        if (host.done) break;
        rr_prev = host.rr;
      }
    }
    ++batch;
    if (it >= maxiter + 2 * poll) {
      if (poll_wait((batch - 1) & 1, host)) return -2;
      break;
    }
  }
  HF_CHECK_HIP(hipMemcpyAsync(x, B.x, N * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (!host.done) {
    // stopped by the iteration cap: the batches still queued may count further
    HF_CHECK_HIP(hipStreamSynchronize(s));
    HF_CHECK_HIP(hipMemcpy(&host, sc, sizeof(Scalars), hipMemcpyDeviceToHost));
  }
  // once done is set every later launch exits at its first line, so the polled
  // scalars are final and the copy of x is left queued on the stream
  if (info) {
    info->iterations = host.iter;
    info->status = host.done ? host.status : 1;
    info->res_norm = sqrt(host.rr);
    info->b_norm = sqrt(host.bb);
  }
  return 0;
}

extern "C" int hf_mg_solve_f32(hf_mg* mg, const hf_pattern* p, const float* vals, const float* b, float* x,
                               int32_t flags, int32_t maxiter, double rtol, double atol, int32_t poll, void* ws,
                               hf_krylov_info* info, hf_stream_t stream) {
  HF_REQUIRE(mg && p && vals && b && x && ws, "hf_mg_solve_f32: null argument");
  HF_REQUIRE(pattern_vec(p) == 3 && pattern_nnode(p) == mg->lv[0].g.n, "hf_mg_solve_f32: pattern does not match");
  return mg_solve(mg, p, vals, nullptr, b, x, flags, maxiter, rtol, atol, poll, ws, info, (hipStream_t)stream);
}

extern "C" int hf_mg_solve_stencil_f32(hf_mg* mg, const float* S, const float* b, float* x, int32_t flags,
                                       int32_t maxiter, double rtol, double atol, int32_t poll, void* ws,
                                       hf_krylov_info* info, hf_stream_t stream) {
  HF_REQUIRE(mg && S && b && x && ws, "hf_mg_solve_stencil_f32: null argument");
  HF_REQUIRE(mg->half && !mg->sell, "hf_mg_solve_stencil_f32: the hierarchy has no half-stored fine stencil");
  return mg_solve(mg, nullptr, nullptr, S, b, x, flags, maxiter, rtol, atol, poll, ws, info, (hipStream_t)stream);
}
