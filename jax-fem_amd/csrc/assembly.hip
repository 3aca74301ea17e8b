// Sparsity pattern, deterministic assembly, Dirichlet row elimination.
//
// Reference: the COO pattern of Problem.__post_init__ (problem.py:86-107) is
// preallocated by PETSc (solver.py:469-475) as the sorted unique (I, J) pairs;
// setValuesCOO sums duplicates (solver.py:487) and zeroRows(rows) puts identity
// rows on Dirichlet DOFs, columns kept (solver.py:489-490).
//
// Here every (cell, a, b) node pair is one contribution; a device radix sort of
// the 64-bit keys row_node * nnode + col_node (stable, so equal keys keep
// ascending contribution = ascending cell order) gives the unique node-pair
// blocks in CSR order.  For vec components every node pair is a dense vec x vec
// block of the scalar COO pattern, so the scalar CSR view (row 3n+i, columns
// 3m+j over the sorted neighbours m) is exactly the sorted-unique (I, J).
// Assembly then is a segmented reduction over each block's contribution list,
// summed in cell order: deterministic, no float atomics.
#include "common.hpp"
#include "hipfem.h"
#include <hipcub/hipcub.hpp>

struct hf_pattern {
  int64_t nnode, ncell, nblocks, ncontrib;
  int nn, vec;
  int32_t* rowptr;   // nnode+1 block-row pointer
  int32_t* colidx;   // nblocks
  int32_t* seg;      // nblocks+1 start of each block's contribution segment
  int32_t* contrib;  // ncontrib sorted contribution ids (cell*nn + a)*nn + b
  int32_t* diag;     // nnode diagonal block index (-1 for a node in no cell)
  int32_t* trans;    // nblocks index of the transposed block
  int32_t* brow;     // nblocks block row
  int32_t max_row_blocks;
  // sliced-ELL value layout (see below)
  int R;             // node rows per slice: 64 / vec (64, 32, 21 for vec 1, 2, 3)
  int64_t nslices, nslots;
  int32_t* slice_ptr;  // nslices+1 first slot of each slice
  int32_t* sell_col;   // nslots*R block column per (slot, row); padding -> own row
  int32_t* sell_tpos;  // nslots*R value base of the transposed block, -1 for padding
  int32_t* blkpos;     // nblocks value base of each block
  int32_t* dpos;       // nnode value base of the diagonal block (-1: none)
  int32_t* sell_row;   // nslices*R node of each slice row (-1: padding row)
  int32_t* sell_pos;   // nnode slice row of each node
  uint8_t* cslot;      // HEX8 (nn = 8): ncell*8*8 slot of block (cells[c,a], cells[c,b]) in row cells[c,a]
};

namespace hf {

__global__ void make_keys(const int32_t* __restrict__ cells, int64_t ncell, int nn, int64_t nnode,
                          uint64_t* __restrict__ keys, int32_t* __restrict__ ids) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)nn * nn;
  if (k >= ncell * per) return;
  const int64_t c = k / per;
  const int r = (int)(k - c * per);
  const int a = r / nn, b = r - (r / nn) * nn;
  const uint64_t na = (uint64_t)cells[c * nn + a], nb = (uint64_t)cells[c * nn + b];
  keys[k] = na * (uint64_t)nnode + nb;
  ids[k] = (int32_t)k;
}

__global__ void head_flags(const uint64_t* __restrict__ keys, int64_t n, int32_t* __restrict__ flags) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  flags[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}

__global__ void fill_blocks(const uint64_t* __restrict__ keys, const int32_t* __restrict__ scan,
                            int64_t n, int64_t nnode, int32_t* __restrict__ seg,
                            int32_t* __restrict__ colidx, int32_t* __restrict__ brow,
                            int32_t* __restrict__ rowcount, int32_t* __restrict__ diag) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (k != 0 && keys[k] == keys[k - 1]) return;
  const int32_t b = scan[k] - 1;
  const uint64_t key = keys[k];
  const int32_t row = (int32_t)(key / (uint64_t)nnode), col = (int32_t)(key % (uint64_t)nnode);
  seg[b] = (int32_t)k;
  colidx[b] = col;
  brow[b] = row;
  atomicAdd(&rowcount[row], 1);  // integer count, order-independent
  if (row == col) diag[row] = b;
}

__global__ void transpose_map(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                              const int32_t* __restrict__ brow, int64_t nblocks,
                              int32_t* __restrict__ trans) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const int32_t r = brow[b], c = colidx[b];
  int lo = rowptr[c], hi = rowptr[c + 1] - 1;
  int found = -1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int v = colidx[mid];
    if (v == r) { found = mid; break; }
    if (v < r) lo = mid + 1; else hi = mid - 1;
  }
  trans[b] = found;
}

__global__ void row_max(const int32_t* __restrict__ rowptr, int64_t nnode, int32_t* __restrict__ mx) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  atomicMax(mx, rowptr[n + 1] - rowptr[n]);  // integer max, order-independent
}

// scalar CSR view of a vec-block pattern
__global__ void csr_view(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                         int64_t nnode, int vec, int32_t* __restrict__ srow, int32_t* __restrict__ scol) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // scalar row
  const int64_t N = nnode * vec;
  if (r > N) return;
  if (r == N) { srow[N] = rowptr[nnode] * vec * vec; return; }
  const int64_t n = r / vec;
  const int i = (int)(r - n * vec);
  const int32_t b0 = rowptr[n], b1 = rowptr[n + 1];
  const int32_t len = (b1 - b0) * vec;
  const int64_t start = (int64_t)b0 * vec * vec + (int64_t)i * len;
  srow[r] = (int32_t)start;
  for (int32_t b = b0; b < b1; ++b)
    for (int j = 0; j < vec; ++j) scol[start + (int64_t)(b - b0) * vec + j] = colidx[b] * vec + j;
}

// Matrix values: sliced ELL ("SELL-C-sigma").  Node rows are grouped in slices of
// R rows (R = 21 for vec = 3, so a slice is 63 scalar rows = one wavefront lane
// each; R = 32 for vec = 2 and 64 for vec = 1).  Within windows of 8 slices the rows are ordered by
// decreasing length (stable), so rows of similar length share a slice and the
// padding stays small when row lengths vary (the vertex / edge / face / interior
// nodes of HEX27 meshes have 125 / 75 / 45 / 27 blocks); sell_row[s*R + r] is the
// node of slice row r of slice s and sell_pos its inverse.  Slice s has width W_s = its
// longest row (in blocks); its slots are slice_ptr[s] .. slice_ptr[s+1]-1.
// Slot t holds, for each row r of the slice, the r-th row's j-th block
// (j = t - slice_ptr[s]) or zero padding; the block's vec x vec entries are
// contiguous and column-major:
//     entry (i, jj)  vals[(t*R + r)*vec*vec + jj*vec + i]   (= base + jj*vec + i,
//                    base = (t*R + r)*vec*vec)
//     column         sell_col[t*R + r]
// so lane (r, i) of a wavefront reads column i of its row's block of slot t as
// one contiguous vec-word load (the L = R*vec lanes read L*vec consecutive
// values), multiplies it by the single gathered entry x[col*vec + i] and keeps
// vec partial outputs; the vec lanes of a row are summed once per slice.
// keys ordering the rows of each window of `window` consecutive nodes by
// decreasing length: (window index, ~length)
__global__ void perm_keys(const int32_t* __restrict__ rowptr, int64_t nnode, int64_t window,
                          uint64_t* __restrict__ keys, int32_t* __restrict__ ids) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  const uint32_t len = (uint32_t)min(rowptr[n + 1] - rowptr[n], (int32_t)0xFFFFF);
  keys[n] = ((uint64_t)(n / window) << 20) | (uint64_t)(0xFFFFFu - len);
  ids[n] = (int32_t)n;
}

__global__ void sell_row_fill(const int32_t* __restrict__ order, int64_t nnode, int64_t npos,
                              int32_t* __restrict__ sell_row, int32_t* __restrict__ sell_pos) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npos) return;
  if (q < nnode) {
    const int32_t n = order[q];
    sell_row[q] = n;
    sell_pos[n] = (int32_t)q;
  } else {
    sell_row[q] = -1;
  }
}

__global__ void slice_width(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ sell_row, int R,
                            int64_t nslices, int32_t* __restrict__ W) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nslices) return;
  if (s == nslices) { W[s] = 0; return; }
  int32_t w = 0;
  for (int r = 0; r < R; ++r) {
    const int32_t n = sell_row[s * R + r];
    if (n >= 0) w = max(w, rowptr[n + 1] - rowptr[n]);
  }
  W[s] = w;
}

__global__ void sell_fill(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
                          const int32_t* __restrict__ slice_ptr, const int32_t* __restrict__ sell_row, int R,
                          int vec, int64_t nslices, int32_t* __restrict__ sell_col, int32_t* __restrict__ blkpos) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // slice row
  if (q >= nslices * R) return;
  const int64_t s = q / R;
  const int r = (int)(q - s * R);
  const int32_t n = sell_row[q];
  const int32_t t0 = slice_ptr[s], t1 = slice_ptr[s + 1];
  const int32_t b0 = n >= 0 ? rowptr[n] : 0, len = n >= 0 ? rowptr[n + 1] - b0 : 0;
  for (int32_t t = t0; t < t1; ++t) {
    const int32_t j = t - t0;
    if (j < len) {
      sell_col[(int64_t)t * R + r] = colidx[b0 + j];
      blkpos[b0 + j] = (int32_t)(((int64_t)t * R + r) * vec * vec);
    } else {
      sell_col[(int64_t)t * R + r] = n >= 0 ? n : 0;   // padding: a zero block on an existing column
    }
  }
}

__global__ void sell_tpos_fill(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ trans,
                               const int32_t* __restrict__ blkpos, const int32_t* __restrict__ slice_ptr,
                               const int32_t* __restrict__ sell_row, int R, int64_t nslices,
                               int32_t* __restrict__ sell_tpos) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nslices * R) return;
  const int64_t s = q / R;
  const int r = (int)(q - s * R);
  const int32_t n = sell_row[q];
  const int32_t t0 = slice_ptr[s], t1 = slice_ptr[s + 1];
  const int32_t b0 = n >= 0 ? rowptr[n] : 0, len = n >= 0 ? rowptr[n + 1] - b0 : 0;
  for (int32_t t = t0; t < t1; ++t) {
    const int32_t j = t - t0;
    int32_t v = -1;
    if (j < len) {
      const int32_t kt = trans[b0 + j];
      v = kt >= 0 ? blkpos[kt] : -1;
    }
    sell_tpos[(int64_t)t * R + r] = v;
  }
}

// For each (cell c, local row a, local column b): the index of the block
// (cells[c,a], cells[c,b]) within its CSR row (255 if absent or beyond 254).
__global__ void cslot_fill(const int32_t* __restrict__ cells, int64_t ncell, const int32_t* __restrict__ rowptr,
                           const int32_t* __restrict__ colidx, uint8_t* __restrict__ cslot) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // c*8 + a
  if (t >= ncell * 8) return;
  const int64_t c = t >> 3;
  const int32_t n = cells[t];
  const int32_t b0 = rowptr[n], len = rowptr[n + 1] - b0;
  for (int b = 0; b < 8; ++b) {
    const int32_t m = cells[c * 8 + b];
    int lo = 0, hi = len;   // columns of a row are sorted
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (colidx[b0 + mid] < m) lo = mid + 1;
      else hi = mid;
    }
    const int j = (lo < len && colidx[b0 + lo] == m) ? lo : 255;
    cslot[t * 8 + b] = (uint8_t)(j < 255 ? j : 255);
  }
}

__global__ void dpos_fill(const int32_t* __restrict__ diag, const int32_t* __restrict__ blkpos, int64_t nnode,
                          int32_t* __restrict__ dpos) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nnode) return;
  dpos[n] = diag[n] >= 0 ? blkpos[diag[n]] : -1;
}

// scalar CSR values (the order of hf_pattern_csr) from the SELL values
template <typename T>
__global__ void sell_to_csr_kernel(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ brow,
                                   const int32_t* __restrict__ blkpos, int64_t nblocks, int vec, int R,
                                   const T* __restrict__ vals, T* __restrict__ csr) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const int32_t n = brow[b];
  const int32_t b0 = rowptr[n], len = (rowptr[n + 1] - b0) * vec;
  const int64_t base = blkpos[b];
  (void)R;
  for (int i = 0; i < vec; ++i)
    for (int j = 0; j < vec; ++j)
      csr[(int64_t)b0 * vec * vec + (int64_t)i * len + (b - b0) * vec + j] = vals[base + j * vec + i];
}

// One thread per node row: every slot of the row's slice gets the sum of the
// block's contributions (element blocks, block-major (cell, a, b, i, k), each
// BB contiguous values) in cell order, or zeros for padding.
// kAsmLanes lanes per slice row: lane l sums the row's slots t0 + l, t0 + l +
// kAsmLanes, ..., so a HEX27 row's ~125 blocks take ~16 dependent rounds per lane
// instead of 125 for one thread.  Each block is still summed by one lane over its
// contributions in cell order: the same FP additions as one thread per row.
constexpr int kAsmLanes = 8;
template <typename T, int VEC>
__global__ void __launch_bounds__(256) assemble_kernel(const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ seg,
                                                       const int32_t* __restrict__ contrib,
                                                       const int32_t* __restrict__ slice_ptr,
                                                       const int32_t* __restrict__ sell_row, int R, int64_t nslices,
                                                       const T* __restrict__ elem, T* __restrict__ vals) {
  constexpr int BB = VEC * VEC;
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t q = gid / kAsmLanes;  // slice row
  const int lane = (int)(gid % kAsmLanes);
  if (q >= nslices * R) return;
  const int64_t s = q / R;
  const int r = (int)(q - s * R);
  const int32_t n = sell_row[q];
  const int32_t t0 = slice_ptr[s], t1 = slice_ptr[s + 1];
  const int32_t b0 = n >= 0 ? rowptr[n] : 0, len = n >= 0 ? rowptr[n + 1] - b0 : 0;
  for (int32_t t = t0 + lane; t < t1; t += kAsmLanes) {
    const int32_t j = t - t0;
    T acc[BB];
#pragma unroll
    for (int e = 0; e < BB; ++e) acc[e] = 0;
    if (j < len) {
      const int32_t b = b0 + j;
      for (int32_t c = seg[b]; c < seg[b + 1]; ++c) {
        const T* src = elem + (int64_t)contrib[c] * BB;
#pragma unroll
        for (int e = 0; e < BB; ++e) acc[e] += src[e];
      }
    }
    T* dst = vals + ((int64_t)t * R + r) * BB;
    // element blocks are row-major (i, k); matrix blocks are stored column-major
#pragma unroll
    for (int e = 0; e < BB; ++e) dst[(e % VEC) * VEC + e / VEC] = acc[e];
  }
}

// R[n*vec+i] = sum over the cells holding node n (diag-block contributions,
// cell order) of elem_res[(cell*nn + a)*vec + i]  (+ f_ext).  8 lanes per node:
// lane l fetches contribution l of each chunk of 8, so a node's loads are in
// flight together; the group then adds them in cell order (shuffles from lanes
// 0..7), the same sequence of FP64 additions as one thread walking the list.
constexpr int kGatherLanes = 8;
template <int VEC>
__global__ void __launch_bounds__(256) gather_kernel(const int32_t* __restrict__ diag, const int32_t* __restrict__ seg,
                                                     const int32_t* __restrict__ contrib, int64_t nnode, int nn,
                                                     const double* __restrict__ elem, const double* __restrict__ fext,
                                                     double* __restrict__ R) {
  const int64_t n = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kGatherLanes;
  const int cl = threadIdx.x % kGatherLanes;
  if (n >= nnode) return;  // whole groups leave together
  int32_t s0 = 0, ncl = 0;
  const int32_t b = diag[n];
  if (b >= 0) {
    s0 = seg[b];
    ncl = seg[b + 1] - s0;
  }
  double acc[VEC];
#pragma unroll
  for (int i = 0; i < VEC; ++i) acc[i] = fext ? fext[n * VEC + i] : 0.0;
  for (int cbase = 0; cbase < ncl; cbase += kGatherLanes) {
    double v[VEC];
    if (cbase + cl < ncl) {
      const int64_t ca = contrib[s0 + cbase + cl] / nn;  // = cell*nn + a  (contribution (cell,a,a))
#pragma unroll
      for (int i = 0; i < VEC; ++i) v[i] = elem[ca * VEC + i];
    } else {
#pragma unroll
      for (int i = 0; i < VEC; ++i) v[i] = 0.0;
    }
#pragma unroll
    for (int l = 0; l < kGatherLanes; ++l) {
#pragma unroll
      for (int i = 0; i < VEC; ++i) {
        const double w = __shfl(v[i], l, kGatherLanes);
        if (cbase + l < ncl) acc[i] += w;
      }
    }
  }
  if (cl < VEC) {
    double out = acc[0];
#pragma unroll
    for (int i = 1; i < VEC; ++i) out = (cl == i) ? acc[i] : out;
    R[n * VEC + cl] = out;
  }
}

// The same sums on a HEX8 vec-3 box lattice, thread per node: the node's cells
// and its local corner in each follow from its coordinates (cell (i, j, k) =
// (i*(ny-1) + j)*(nz-1) + k, corner (dx, dy, dz) at local index ct[4dx+2dy+dz]),
// so its 8 contributions load in one round; added in ascending cell order, the
// order of the pattern's contribution lists, so the result is bit-identical.
struct GatherCorners {
  int8_t idx[8];
};
__global__ void __launch_bounds__(256) gather_box_kernel(int nx, int ny, int nz, GatherCorners ct,
                                                         const double* __restrict__ elem,
                                                         const double* __restrict__ fext, double* __restrict__ R) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nnode = (int64_t)nx * ny * nz;
  if (n >= nnode) return;
  const int z = (int)(n % nz), y = (int)((n / nz) % ny), x = (int)(n / ((int64_t)nz * ny));
  const int ncx = nx - 1, ncy = ny - 1, ncz = nz - 1;
  double acc[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) acc[i] = fext ? fext[3 * n + i] : 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int cx = x - 1 + a, cy = y - 1 + b, cz = z - 1 + c;
        if (cx < 0 || cx >= ncx || cy < 0 || cy >= ncy || cz < 0 || cz >= ncz) continue;
        const int64_t cell = ((int64_t)cx * ncy + cy) * ncz + cz;
        const int loc = ct.idx[(x - cx) * 4 + (y - cy) * 2 + (z - cz)];
        const double* src = elem + (cell * 8 + loc) * 3;
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[i] += src[i];
      }
#pragma unroll
  for (int i = 0; i < 3; ++i) R[3 * n + i] = acc[i];
}

template <typename T>
__global__ void bc_rows_kernel(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ slice_ptr,
                               const int32_t* __restrict__ sell_pos, const int32_t* __restrict__ dpos,
                               const int32_t* __restrict__ dofs, int64_t nbc, int vec, int R, T* __restrict__ vals) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbc) return;
  const int32_t d = dofs[t];
  const int32_t n = d / vec, i = d - (d / vec) * vec;
  const int32_t q = sell_pos[n];
  const int64_t s = q / R;
  const int r = (int)(q - s * R);
  const int32_t len = rowptr[n + 1] - rowptr[n];
  for (int32_t j = 0; j < len; ++j) {
    T* base = vals + ((int64_t)(slice_ptr[s] + j) * R + r) * vec * vec;
    for (int jj = 0; jj < vec; ++jj) base[jj * vec + i] = (T)0;  // row i of a column-major block
  }
  if (dpos[n] >= 0) vals[(int64_t)dpos[n] + i * vec + i] = (T)1;
}

__global__ void bc_residual_kernel(const int32_t* __restrict__ dofs, const double* __restrict__ vals,
                                   int64_t nbc, double scale, const double* __restrict__ u,
                                   double* __restrict__ R) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbc) return;
  const int32_t d = dofs[t];
  R[d] = u[d] - vals[t] * scale;
}

__global__ void bc_x0_kernel(const int32_t* __restrict__ dofs, const double* __restrict__ vals,
                             int64_t nbc, const double* __restrict__ u, double* __restrict__ x0) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbc) return;
  const int32_t d = dofs[t];
  x0[d] = vals[t] - u[d];
}

template <typename T>
__global__ void jacobi_kernel(const int32_t* __restrict__ dpos, int64_t nnode, int vec, int R,
                              const T* __restrict__ vals, T* __restrict__ dinv) {
  const int64_t rr = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (rr >= nnode * vec) return;
  const int64_t n = rr / vec;
  const int i = (int)(rr - n * vec);
  const int32_t b = dpos[n];
  (void)R;
  const T d = b >= 0 ? vals[(int64_t)b + (int64_t)i * vec + i] : (T)0;
  dinv[rr] = d != (T)0 ? (T)1 / d : (T)1;   // a zero diagonal (saddle-point block) is left unscaled
}

}  // namespace hf

using namespace hf;

static inline unsigned grid_for(int64_t n, int b = 256) { return (unsigned)ceil_div(n, b); }

extern "C" int hf_pattern_build(const int32_t* cells, int64_t ncell, int32_t nn, int64_t nnode,
                                int32_t vec, hf_stream_t stream_, hf_pattern** out) {
  hipStream_t stream = (hipStream_t)stream_;
  HF_REQUIRE(out != nullptr && cells != nullptr, "hf_pattern_build: null argument");
  HF_REQUIRE(ncell > 0 && nnode > 0 && nn > 0 && vec > 0, "hf_pattern_build: empty mesh");
  HF_REQUIRE(vec <= 3, "hf_pattern_build: vec=%d (1, 2 or 3 components per node)", vec);
  const int64_t n = ncell * (int64_t)nn * nn;
  HF_REQUIRE(n < (int64_t)INT32_MAX, "hf_pattern_build: %lld contributions exceed int32", (long long)n);
  HF_REQUIRE(nnode * vec < (int64_t)INT32_MAX, "hf_pattern_build: too many DOFs");
  hf_pattern* p = new hf_pattern();
  p->nnode = nnode; p->ncell = ncell; p->nn = nn; p->vec = vec; p->ncontrib = n;
  uint64_t *keys = nullptr, *keys_s = nullptr;
  int32_t *ids = nullptr, *flags = nullptr, *scan = nullptr, *rowcount = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0, tmp2 = 0;
  int end_bit = 1;
  while (end_bit < 64 && ((uint64_t)1 << end_bit) < (uint64_t)nnode * (uint64_t)nnode) ++end_bit;
#define PB_CHECK(expr)                                                                    \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess) {                                                               \
      set_error("hf_pattern_build: %s -> %s", #expr, hipGetErrorString(_e));              \
      hipFree(keys); hipFree(keys_s); hipFree(ids); hipFree(flags); hipFree(scan);        \
      hipFree(rowcount); hipFree(tmp);                                                    \
      hf_pattern_destroy(p);                                                              \
      return -2;                                                                          \
    }                                                                                     \
  } while (0)
  p->rowptr = p->colidx = p->seg = p->contrib = p->diag = p->trans = p->brow = nullptr;
  p->slice_ptr = p->sell_col = p->sell_tpos = p->blkpos = p->dpos = p->sell_row = p->sell_pos = nullptr;
  p->cslot = nullptr;
  p->R = kWave / vec;   // node rows per slice: one wavefront holds R*vec <= 64 scalar rows
  p->nslices = (nnode + p->R - 1) / p->R;
  PB_CHECK(hipMalloc(&keys, n * sizeof(uint64_t)));
  PB_CHECK(hipMalloc(&keys_s, n * sizeof(uint64_t)));
  PB_CHECK(hipMalloc(&ids, n * sizeof(int32_t)));
  PB_CHECK(hipMalloc(&p->contrib, n * sizeof(int32_t)));
  make_keys<<<grid_for(n), 256, 0, stream>>>(cells, ncell, nn, nnode, keys, ids);
  PB_CHECK(hipGetLastError());
  PB_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_s, ids, p->contrib, (int)n, 0,
                                              end_bit, stream));
  PB_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, tmp2, flags, scan, (int)n, stream));
  if (tmp2 > tmp_bytes) tmp_bytes = tmp2;
  PB_CHECK(hipMalloc(&tmp, tmp_bytes));
  PB_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys_s, ids, p->contrib, (int)n, 0,
                                              end_bit, stream));
  hipFree(keys); keys = nullptr;
  hipFree(ids); ids = nullptr;
  PB_CHECK(hipMalloc(&flags, n * sizeof(int32_t)));
  PB_CHECK(hipMalloc(&scan, n * sizeof(int32_t)));
  head_flags<<<grid_for(n), 256, 0, stream>>>(keys_s, n, flags);
  PB_CHECK(hipGetLastError());
  PB_CHECK(hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, flags, scan, (int)n, stream));
  int32_t nb = 0;
  PB_CHECK(hipMemcpyAsync(&nb, scan + (n - 1), sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  PB_CHECK(hipStreamSynchronize(stream));
  p->nblocks = nb;
  PB_CHECK(hipMalloc(&p->seg, (nb + 1) * sizeof(int32_t)));
  PB_CHECK(hipMalloc(&p->colidx, nb * sizeof(int32_t)));
  PB_CHECK(hipMalloc(&p->brow, nb * sizeof(int32_t)));
  PB_CHECK(hipMalloc(&p->trans, nb * sizeof(int32_t)));
  PB_CHECK(hipMalloc(&p->diag, nnode * sizeof(int32_t)));
  PB_CHECK(hipMalloc(&p->rowptr, (nnode + 1) * sizeof(int32_t)));
  PB_CHECK(hipMalloc(&rowcount, (nnode + 1) * sizeof(int32_t)));
  PB_CHECK(hipMemsetAsync(rowcount, 0, (nnode + 1) * sizeof(int32_t), stream));
  PB_CHECK(hipMemsetAsync(p->diag, 0xff, nnode * sizeof(int32_t), stream));
  fill_blocks<<<grid_for(n), 256, 0, stream>>>(keys_s, scan, n, nnode, p->seg, p->colidx, p->brow, rowcount,
                                                 p->diag);
  PB_CHECK(hipGetLastError());
  const int32_t nn32 = (int32_t)n;
  PB_CHECK(hipMemcpyAsync(p->seg + nb, &nn32, sizeof(int32_t), hipMemcpyHostToDevice, stream));
  PB_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, rowcount, p->rowptr, (int)(nnode + 1), stream));
  if (tmp2 > tmp_bytes) {
    hipFree(tmp); tmp = nullptr;
    tmp_bytes = tmp2;
    PB_CHECK(hipMalloc(&tmp, tmp_bytes));
  }
  PB_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, rowcount, p->rowptr, (int)(nnode + 1), stream));
  transpose_map<<<grid_for(nb), 256, 0, stream>>>(p->rowptr, p->colidx, p->brow, nb, p->trans);
  PB_CHECK(hipGetLastError());
  {
    // longest block row (selects the SpMV variant); rowcount is no longer needed
    int32_t* mx = rowcount + nnode;
    PB_CHECK(hipMemsetAsync(mx, 0, sizeof(int32_t), stream));
    row_max<<<grid_for(nnode), 256, 0, stream>>>(p->rowptr, nnode, mx);
    PB_CHECK(hipGetLastError());
    PB_CHECK(hipMemcpyAsync(&p->max_row_blocks, mx, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
  }
  {
    // sliced-ELL layout: rows ordered by decreasing length within windows of 8 slices
    const int64_t ns = p->nslices;
    {
      const int64_t window = (int64_t)p->R * 8;
      uint64_t *pk = nullptr, *pk_s = nullptr;
      int32_t *pid = nullptr, *order = nullptr;
      PB_CHECK(hipMalloc(&pk, nnode * sizeof(uint64_t)));
      PB_CHECK(hipMalloc(&pk_s, nnode * sizeof(uint64_t)));
      PB_CHECK(hipMalloc(&pid, nnode * sizeof(int32_t)));
      PB_CHECK(hipMalloc(&order, nnode * sizeof(int32_t)));
      perm_keys<<<grid_for(nnode), 256, 0, stream>>>(p->rowptr, nnode, window, pk, pid);
      PB_CHECK(hipGetLastError());
      int pbits = 20;
      while (pbits < 63 && ((uint64_t)1 << (pbits - 20)) <= (uint64_t)(nnode / window)) ++pbits;
      size_t need = 0;
      PB_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, need, pk, pk_s, pid, order, (int)nnode, 0, pbits,
                                                  stream));
      if (need > tmp_bytes) {
        hipFree(tmp); tmp = nullptr;
        tmp_bytes = need;
        PB_CHECK(hipMalloc(&tmp, tmp_bytes));
      }
      PB_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, pk, pk_s, pid, order, (int)nnode, 0, pbits,
                                                  stream));
      PB_CHECK(hipMalloc(&p->sell_row, ns * p->R * sizeof(int32_t)));
      PB_CHECK(hipMalloc(&p->sell_pos, nnode * sizeof(int32_t)));
      sell_row_fill<<<grid_for(ns * p->R), 256, 0, stream>>>(order, nnode, ns * p->R, p->sell_row, p->sell_pos);
      PB_CHECK(hipGetLastError());
      PB_CHECK(hipStreamSynchronize(stream));
      hipFree(pk); hipFree(pk_s); hipFree(pid); hipFree(order);
    }
    int32_t* W = nullptr;
    PB_CHECK(hipMalloc(&W, (ns + 1) * sizeof(int32_t)));
    PB_CHECK(hipMalloc(&p->slice_ptr, (ns + 1) * sizeof(int32_t)));
    slice_width<<<grid_for(ns + 1), 256, 0, stream>>>(p->rowptr, p->sell_row, p->R, ns, W);
    PB_CHECK(hipGetLastError());
    size_t need = 0;
    PB_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, W, p->slice_ptr, (int)(ns + 1), stream));
    if (need > tmp_bytes) {
      hipFree(tmp); tmp = nullptr;
      tmp_bytes = need;
      PB_CHECK(hipMalloc(&tmp, tmp_bytes));
    }
    PB_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, W, p->slice_ptr, (int)(ns + 1), stream));
    int32_t nslots = 0;
    PB_CHECK(hipMemcpyAsync(&nslots, p->slice_ptr + ns, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    PB_CHECK(hipStreamSynchronize(stream));
    hipFree(W);
    p->nslots = nslots;
    const int64_t cells_sell = (int64_t)nslots * p->R;
    HF_REQUIRE(cells_sell * vec * vec < (int64_t)INT32_MAX, "hf_pattern_build: SELL values exceed int32 indexing");
    PB_CHECK(hipMalloc(&p->sell_col, cells_sell * sizeof(int32_t)));
    PB_CHECK(hipMalloc(&p->sell_tpos, cells_sell * sizeof(int32_t)));
    PB_CHECK(hipMalloc(&p->blkpos, nb * sizeof(int32_t)));
    PB_CHECK(hipMalloc(&p->dpos, nnode * sizeof(int32_t)));
    sell_fill<<<grid_for(ns * p->R), 256, 0, stream>>>(p->rowptr, p->colidx, p->slice_ptr, p->sell_row, p->R, vec,
                                                      ns, p->sell_col, p->blkpos);
    PB_CHECK(hipGetLastError());
    sell_tpos_fill<<<grid_for(ns * p->R), 256, 0, stream>>>(p->rowptr, p->trans, p->blkpos, p->slice_ptr,
                                                           p->sell_row, p->R, ns, p->sell_tpos);
    PB_CHECK(hipGetLastError());
    dpos_fill<<<grid_for(nnode), 256, 0, stream>>>(p->diag, p->blkpos, nnode, p->dpos);
    PB_CHECK(hipGetLastError());
    if (nn == 8) {
      PB_CHECK(hipMalloc(&p->cslot, ncell * 64));
      cslot_fill<<<grid_for(ncell * 8), 256, 0, stream>>>(cells, ncell, p->rowptr, p->colidx, p->cslot);
      PB_CHECK(hipGetLastError());
    }
  }
  PB_CHECK(hipStreamSynchronize(stream));
  hipFree(keys_s); hipFree(flags); hipFree(scan); hipFree(rowcount); hipFree(tmp);
#undef PB_CHECK
  *out = p;
  return 0;
}

extern "C" int hf_pattern_sizes(const hf_pattern* p, int64_t* nblocks, int64_t* nnz, int64_t* ncontrib) {
  HF_REQUIRE(p != nullptr, "null pattern");
  if (nblocks) *nblocks = p->nblocks;
  if (nnz) *nnz = p->nblocks * p->vec * p->vec;
  if (ncontrib) *ncontrib = p->ncontrib;
  return 0;
}

extern "C" int32_t hf_pattern_max_row_blocks(const hf_pattern* p) { return p ? p->max_row_blocks : -1; }

extern "C" int64_t hf_matrix_values(const hf_pattern* p) {
  return p ? p->nslots * (int64_t)p->R * p->vec * p->vec : -1;
}

extern "C" int hf_pattern_arrays(const hf_pattern* p, const int32_t** rowptr, const int32_t** colidx,
                                 const int32_t** diag, const int32_t** trans) {
  HF_REQUIRE(p != nullptr, "null pattern");
  if (rowptr) *rowptr = p->rowptr;
  if (colidx) *colidx = p->colidx;
  if (diag) *diag = p->diag;
  if (trans) *trans = p->trans;
  return 0;
}

extern "C" int hf_pattern_csr(const hf_pattern* p, int32_t* rowptr, int32_t* colind, hf_stream_t s) {
  HF_REQUIRE(p && rowptr && colind, "hf_pattern_csr: null argument");
  HF_REQUIRE(p->nblocks * p->vec * p->vec < (int64_t)INT32_MAX, "hf_pattern_csr: nnz exceeds int32");
  csr_view<<<grid_for(p->nnode * p->vec + 1), 256, 0, (hipStream_t)s>>>(p->rowptr, p->colidx, p->nnode, p->vec,
                                                                          rowptr, colind);
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_pattern_destroy(hf_pattern* p) {
  if (!p) return 0;
  hipFree(p->rowptr); hipFree(p->colidx); hipFree(p->seg); hipFree(p->contrib);
  hipFree(p->diag); hipFree(p->trans); hipFree(p->brow);
  hipFree(p->slice_ptr); hipFree(p->sell_col); hipFree(p->sell_tpos); hipFree(p->blkpos); hipFree(p->dpos);
  hipFree(p->sell_row); hipFree(p->sell_pos);
  if (p->cslot) hipFree(p->cslot);
  delete p;
  return 0;
}

extern "C" int hf_matrix_to_csr_f64(const hf_pattern* p, const double* vals, double* csr, hf_stream_t s) {
  HF_REQUIRE(p && vals && csr, "hf_matrix_to_csr_f64: null argument");
  sell_to_csr_kernel<double><<<grid_for(p->nblocks), 256, 0, (hipStream_t)s>>>(p->rowptr, p->brow, p->blkpos,
                                                                                 p->nblocks, p->vec, p->R, vals, csr);
  HF_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int assemble_impl(const hf_pattern* p, const T* elem, T* vals, hipStream_t s) {
  HF_REQUIRE(p && elem && vals, "hf_assemble: null argument");
  const int64_t n = p->nslices * p->R;
#define ASM_LAUNCH(V)                                                                                        \
  assemble_kernel<T, V><<<grid_for(n * kAsmLanes), 256, 0, s>>>(p->rowptr, p->seg, p->contrib, p->slice_ptr,   \
                                                                 p->sell_row, p->R, p->nslices, elem, vals)
  if (p->vec == 3)
    ASM_LAUNCH(3);
  else if (p->vec == 2)
    ASM_LAUNCH(2);
  else if (p->vec == 1)
    ASM_LAUNCH(1);
  else
    HF_REQUIRE(false, "hf_assemble: vec=%d not supported", p->vec);
#undef ASM_LAUNCH
  HF_LAUNCH_CHECK();
  return 0;
}
extern "C" int hf_assemble_f64(const hf_pattern* p, const double* e, double* v, hf_stream_t s) {
  return assemble_impl<double>(p, e, v, (hipStream_t)s);
}
extern "C" int hf_assemble_f32(const hf_pattern* p, const float* e, float* v, hf_stream_t s) {
  return assemble_impl<float>(p, e, v, (hipStream_t)s);
}

extern "C" int hf_gather_residual_f64(const hf_pattern* p, const double* elem, const double* fext, double* R,
                                      hf_stream_t s) {
  HF_REQUIRE(p && elem && R, "hf_gather_residual_f64: null argument");
#define GATHER_LAUNCH(V)                                                        \
  gather_kernel<V><<<grid_for(p->nnode * kGatherLanes), 256, 0, (hipStream_t)s>>>( \
      p->diag, p->seg, p->contrib, p->nnode, p->nn, elem, fext, R)
  if (p->vec == 3)
    GATHER_LAUNCH(3);
  else if (p->vec == 2)
    GATHER_LAUNCH(2);
  else if (p->vec == 1)
    GATHER_LAUNCH(1);
  else
    HF_REQUIRE(false, "vec=%d not supported", p->vec);
#undef GATHER_LAUNCH
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_gather_residual_box_f64(const hf_pattern* p, int32_t nx, int32_t ny, int32_t nz,
                                          const int32_t* corner, const double* elem, const double* fext, double* R,
                                          hf_stream_t s) {
  HF_REQUIRE(p && corner && elem && R, "hf_gather_residual_box_f64: null argument");
  HF_REQUIRE(p->vec == 3 && p->nn == 8, "hf_gather_residual_box_f64: needs a HEX8 vec-3 pattern");
  HF_REQUIRE(nx >= 2 && ny >= 2 && nz >= 2 && (int64_t)nx * ny * nz == p->nnode &&
                 (int64_t)(nx - 1) * (ny - 1) * (nz - 1) == p->ncell,
             "hf_gather_residual_box_f64: %dx%dx%d is not the pattern's lattice", nx, ny, nz);
  GatherCorners ct;
  bool seen[8] = {false, false, false, false, false, false, false, false};
  for (int q = 0; q < 8; ++q) {
    HF_REQUIRE(corner[q] >= 0 && corner[q] < 8 && !seen[corner[q]],
               "hf_gather_residual_box_f64: corner table is not a permutation of 0..7");
    seen[corner[q]] = true;
    ct.idx[q] = (int8_t)corner[q];
  }
  gather_box_kernel<<<grid_for(p->nnode), 256, 0, (hipStream_t)s>>>(nx, ny, nz, ct, elem, fext, R);
  HF_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int bc_rows_impl(const hf_pattern* p, const int32_t* dofs, int64_t nbc, T* vals, hipStream_t s) {
  HF_REQUIRE(p && vals, "hf_bc_rows: null argument");
  if (nbc == 0) return 0;
  bc_rows_kernel<T><<<grid_for(nbc), 256, 0, s>>>(p->rowptr, p->slice_ptr, p->sell_pos, p->dpos, dofs, nbc, p->vec,
                                                   p->R, vals);
  HF_LAUNCH_CHECK();
  return 0;
}
extern "C" int hf_bc_rows_f64(const hf_pattern* p, const int32_t* d, int64_t n, double* b, hf_stream_t s) {
  return bc_rows_impl<double>(p, d, n, b, (hipStream_t)s);
}
extern "C" int hf_bc_rows_f32(const hf_pattern* p, const int32_t* d, int64_t n, float* b, hf_stream_t s) {
  return bc_rows_impl<float>(p, d, n, b, (hipStream_t)s);
}

extern "C" int hf_bc_residual_f64(const int32_t* dofs, const double* vals, int64_t nbc, double scale,
                                  const double* u, double* R, hf_stream_t s) {
  if (nbc == 0) return 0;
  bc_residual_kernel<<<grid_for(nbc), 256, 0, (hipStream_t)s>>>(dofs, vals, nbc, scale, u, R);
  HF_LAUNCH_CHECK();
  return 0;
}

extern "C" int hf_bc_initial_guess_f64(const int32_t* dofs, const double* vals, int64_t nbc, const double* u,
                                       double* x0, int64_t N, hf_stream_t s) {
  HF_CHECK_HIP(hipMemsetAsync(x0, 0, N * sizeof(double), (hipStream_t)s));
  if (nbc == 0) return 0;
  bc_x0_kernel<<<grid_for(nbc), 256, 0, (hipStream_t)s>>>(dofs, vals, nbc, u, x0);
  HF_LAUNCH_CHECK();
  return 0;
}

template <typename T>
static int jacobi_impl(const hf_pattern* p, const T* bsr, T* dinv, hipStream_t s) {
  HF_REQUIRE(p && bsr && dinv, "hf_jacobi: null argument");
  jacobi_kernel<T><<<grid_for(p->nnode * p->vec), 256, 0, s>>>(p->dpos, p->nnode, p->vec, p->R, bsr, dinv);
  HF_LAUNCH_CHECK();
  return 0;
}
extern "C" int hf_jacobi_f64(const hf_pattern* p, const double* b, double* d, hf_stream_t s) {
  return jacobi_impl<double>(p, b, d, (hipStream_t)s);
}
extern "C" int hf_jacobi_f32(const hf_pattern* p, const float* b, float* d, hf_stream_t s) {
  return jacobi_impl<float>(p, b, d, (hipStream_t)s);
}

// internal accessors for the Krylov module
namespace hf {
const int32_t* pattern_rowptr(const hf_pattern* p) { return p->rowptr; }
const int32_t* pattern_colidx(const hf_pattern* p) { return p->colidx; }
const int32_t* pattern_trans(const hf_pattern* p) { return p->trans; }
const int32_t* pattern_blkpos(const hf_pattern* p) { return p->blkpos; }
const int32_t* pattern_sell_pos(const hf_pattern* p) { return p->sell_pos; }
const int32_t* pattern_dpos(const hf_pattern* p) { return p->dpos; }
const uint8_t* pattern_cslot(const hf_pattern* p) { return p->cslot; }
int64_t pattern_nnode(const hf_pattern* p) { return p->nnode; }
int64_t pattern_nblocks(const hf_pattern* p) { return p->nblocks; }
int pattern_R(const hf_pattern* p) { return p->R; }
int64_t pattern_nslices(const hf_pattern* p) { return p->nslices; }
const int32_t* pattern_slice_ptr(const hf_pattern* p) { return p->slice_ptr; }
const int32_t* pattern_sell_col(const hf_pattern* p) { return p->sell_col; }
const int32_t* pattern_sell_tpos(const hf_pattern* p) { return p->sell_tpos; }
const int32_t* pattern_sell_row(const hf_pattern* p) { return p->sell_row; }
const int32_t* pattern_seg(const hf_pattern* p) { return p->seg; }
const int32_t* pattern_contrib(const hf_pattern* p) { return p->contrib; }
const int32_t* pattern_diag(const hf_pattern* p) { return p->diag; }
int pattern_nn(const hf_pattern* p) { return p->nn; }
int pattern_max_row_blocks(const hf_pattern* p) { return p->max_row_blocks; }
int pattern_vec(const hf_pattern* p) { return p->vec; }
}  // namespace hf
