/* libhipfem — MI355X (gfx950) kernels for JAX-FEM's implicit solve path.
 *
 * C ABI: plain pointers and sizes.  Every array argument is a caller-owned
 * DEVICE pointer unless stated otherwise; `stream` is a hipStream_t (the
 * caller's current stream).  Every function returns 0 on success, -1 on an
 * argument error and -2 on a HIP error; hf_last_error() gives the message.
 * Calls allocate no device memory except hf_pattern_build / hf_pattern_destroy
 * (set-up time).  Calls on one stream are ordered; calls sharing workspace
 * must not run concurrently on different streams.
 *
 * DOF numbering is the reference's interleaved 3*node+comp (problem.py:86-92).
 * Element matrices, the node-pair block matrix and its CSR view follow the
 * reference's COO layout jac[r,c] = dR_r/du_c (problem.py:262-266, 453).
 */
#ifndef HIPFEM_H
#define HIPFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* hf_stream_t; /* hipStream_t */

/* ---- constitutive laws (the Problem's get_tensor_map, problem.py:189-214) ---- */
enum { HF_MAT_POISSON = 0, HF_MAT_LINEAR_ELASTIC = 1, HF_MAT_NEO_HOOKEAN = 2 };
/* how the per-quadrature parameter p (Problem.internal_vars[0], problem.py:493-556)
 * enters the moduli */
enum { HF_PARAM_NONE = 0, HF_PARAM_SCALE_E = 1, HF_PARAM_SIMP = 2, HF_PARAM_MU = 3,
       HF_PARAM_SCALE_K = 4 };

typedef struct hf_material {
  int32_t kind;       /* HF_MAT_* */
  int32_t param_mode; /* HF_PARAM_* */
  double E, nu;       /* Young's modulus, Poisson ratio */
  double E0, Emin, penal; /* SIMP: E = Emin + (E0-Emin) p^penal */
  double kappa;       /* bulk modulus when mu = p (HF_PARAM_MU) */
  double k;           /* Poisson coefficient (flux = k grad u) */
  int32_t quad1d;     /* Gauss points per direction of the volume rule (quadrature_order
                         -> (order+2)//2, basis.py:56,64); 0 = HEX8's 2.  Built: HEX8 2,
                         HEX27 3 (quadrature_order 4 or 5). */
  int32_t pad_;
} hf_material;

/* Message of the last failed call on this thread; returns its length. */
int hf_last_error(char* buf, size_t n);
/* Library version string (host pointer, static). */
const char* hf_version(void);

/* ---- sparsity pattern: replaces _PetscTangentCache.__init__'s
 *      setPreallocationCOO(I, J) (solver.py:469-475) over the COO (I, J) of
 *      Problem.__post_init__ (problem.py:86-107).  Built on the device by a
 *      radix sort of node-pair keys; the scalar CSR view is bit-identical to
 *      the sorted-unique (I, J) pattern. ---------------------------------- */
typedef struct hf_pattern hf_pattern;
/* vec: components per node, 1..3 */
int hf_pattern_build(const int32_t* cells, int64_t ncell, int32_t nodes_per_cell, int64_t nnode,
                     int32_t vec, hf_stream_t stream, hf_pattern** out);
/* host outputs: number of blocks (node pairs), scalar nnz, contributions */
int hf_pattern_sizes(const hf_pattern* p, int64_t* nblocks, int64_t* nnz, int64_t* ncontrib);
/* device pointers owned by the pattern: block rowptr (nnode+1), block colidx (nblocks),
 * diag block index (nnode), transposed-block index (nblocks) */
int hf_pattern_arrays(const hf_pattern* p, const int32_t** rowptr, const int32_t** colidx,
                      const int32_t** diag, const int32_t** trans);
/* scalar CSR view: rowptr (N+1), colind (nnz) into caller buffers (solver.py:65, getValuesCSR) */
int hf_pattern_csr(const hf_pattern* p, int32_t* rowptr, int32_t* colind, hf_stream_t stream);
int hf_pattern_destroy(hf_pattern* p);

/* ---- cell kernels: replace the jit(vmap(kernel)) / jit(vmap(kernel_jac)) of
 *      pre_jit_fns (problem.py:261-356) and split_and_compute_cell (:359-396).
 *      points (nnode,3) f64, cells (ncell,nn) i32, u (nnode*vec) f64,
 *      params (ncell,nq) f64 or NULL.  elem_res (ncell, nn*vec); elem_jac
 *      block-major (ncell, nn, nn, vec, vec).  ele: 8 (HEX8, 2x2x2 Gauss)
 *      or 27 (HEX27, 3x3x3 Gauss). ----------------------------------------- */
int hf_cell_residual_f64(int32_t ele, const hf_material* mat, const double* points,
                         const int32_t* cells, int64_t ncell, const double* u,
                         const double* params, double* elem_res, hf_stream_t stream);
/* HEX8 box lattices of identical cells (every cell a translate of cell 0, one
 * corner order: a box_mesh): the per-point physical gradients and weights are
 * formed once from cell 0 into geo (hf_hex8_box_geometry_bytes() bytes, device),
 * and hf_cell_residual_box_f64 evaluates the element residuals as
 * hf_cell_residual_f64 does with them (no per-cell geometry).  Same role as
 * hf_cell_residual_f64 (problem.py:261-356). */
size_t hf_hex8_box_geometry_bytes(void);
int hf_hex8_box_geometry(const double* points, const int32_t* cells, int64_t ncell, double* geo,
                         hf_stream_t stream);
int hf_cell_residual_box_f64(const hf_material* mat, const double* geo, const int32_t* cells, int64_t ncell,
                             const double* u, const double* params, double* elem_res, hf_stream_t stream);
int hf_cell_residual_f32(int32_t ele, const hf_material* mat, const double* points,
                         const int32_t* cells, int64_t ncell, const double* u,
                         const double* params, float* elem_res, hf_stream_t stream);
int hf_cell_jacobian_f64(int32_t ele, const hf_material* mat, const double* points,
                         const int32_t* cells, int64_t ncell, const double* u,
                         const double* params, double* elem_jac, hf_stream_t stream);
int hf_cell_jacobian_f32(int32_t ele, const hf_material* mat, const double* points,
                         const int32_t* cells, int64_t ncell, const double* u,
                         const double* params, float* elem_jac, hf_stream_t stream);
/* per-cell parameter VJP: grad[c,q] = -sum lam . dr/dp[c,q]; replaces the
 * jax.vjp over compute_residual in implicit_vjp (solver.py:1348-1378).
 * lam (nnode*vec) f64 must be zero on Dirichlet rows. grad (ncell, nq) f64. */
int hf_cell_param_vjp_f64(int32_t ele, const hf_material* mat, const double* points,
                          const int32_t* cells, int64_t ncell, const double* u,
                          const double* params, const double* lam, double* grad,
                          hf_stream_t stream);

/* ---- element-agnostic cell kernels for a constitutive map known only pointwise:
 *      the contractions of the reference's laplace and mass kernels and their
 *      jacfwd element Jacobians (problem.py:189-236, 262-266) around a law the
 *      caller evaluates at the quadrature points.  Geometry per (cell, point):
 *      shape_grads (ncell, nq, nn, dim) physical gradients, jxw (ncell, nq)
 *      (fe.py:112-141); shape_vals (nq, nn) reference values.  vec 1..3, dim 2
 *      or 3, nn <= 64.
 * u_grad (ncell, nq, vec, dim) = sum_a u[cells[c,a]] (x) dN_a/dx; u_quad (ncell, nq, vec)
 * (optional) = sum_a u N_a. */
int hf_generic_ugrad_f64(const int32_t* cells, int64_t ncell, int32_t nn, int32_t nq, int32_t vec, int32_t dim,
                         const double* u, const double* shape_grads, const double* shape_vals, double* u_grad,
                         double* u_quad, hf_stream_t stream);
/* elem_res (ncell, nn, vec) = sum_q JxW (P : grad N_a + m N_a); P (ncell,nq,vec,dim) and
 * m (ncell,nq,vec) each optional */
int hf_generic_residual_f64(int64_t ncell, int32_t nn, int32_t nq, int32_t vec, int32_t dim,
                            const double* shape_grads, const double* shape_vals, const double* jxw,
                            const double* P, const double* m, double* elem_res, hf_stream_t stream);
/* elem_jac block-major (ncell, nn, nn, vec, vec) = sum_q JxW (grad N_a . C . grad N_b + N_a Mt N_b);
 * C (ncell,nq,vec,dim,vec,dim) = dP/du_grad and Mt (ncell,nq,vec,vec) = dm/du, each optional */
int hf_generic_jacobian_f64(int64_t ncell, int32_t nn, int32_t nq, int32_t vec, int32_t dim,
                            const double* shape_grads, const double* shape_vals, const double* jxw,
                            const double* C, const double* Mt, double* elem_jac, hf_stream_t stream);
int hf_generic_jacobian_f32(int64_t ncell, int32_t nn, int32_t nq, int32_t vec, int32_t dim,
                            const double* shape_grads, const double* shape_vals, const double* jxw,
                            const double* C, const double* Mt, float* elem_jac, hf_stream_t stream);

/* ---- assembly: replaces compute_residual_vars_helper's scatter-add
 *      (problem.py:426-437) and PETSc setValuesCOO + assemble (solver.py:485-488)
 *      by deterministic segmented reductions over the pattern. -------------- */
/* R[node] = sum over cells of elem_res in cell order (+ f_ext if not NULL) */
int hf_gather_residual_f64(const hf_pattern* p, const double* elem_res, const double* f_ext,
                           double* R, hf_stream_t stream);
/* hf_gather_residual_f64 on a HEX8 vec-3 box lattice of nx*ny*nz nodes: the
 * contributions follow from the coordinates (corner: host, 8 ints, local index of
 * corner (dx, dy, dz) at 4 dx + 2 dy + dz, as hf_mg_assemble_hex8_f32's; the caller
 * vouches that every cell is the lattice's), added in the same order: bit-identical.
 * No reference counterpart beyond the scatter-add it shares with
 * hf_gather_residual_f64 (problem.py:426-437). */
int hf_gather_residual_box_f64(const hf_pattern* p, int32_t nx, int32_t ny, int32_t nz, const int32_t* corner,
                               const double* elem, const double* fext, double* R, hf_stream_t stream);
/* Matrix values live in a sliced-ELL layout over the pattern's node-pair blocks
 * (slices of 64/vec node rows: 21 for vec 3, 32 for vec 2, 64 for vec 1; each slice padded to its
 * longest row): hf_matrix_values() gives the element count of a value buffer.
 * assemble: every block = sum of its element blocks in cell order. */
int64_t hf_matrix_values(const hf_pattern* p);
/* blocks in the longest block row (node neighbours, itself included) */
int32_t hf_pattern_max_row_blocks(const hf_pattern* p);
int hf_assemble_f64(const hf_pattern* p, const double* elem_jac, double* vals, hf_stream_t stream);
int hf_assemble_f32(const hf_pattern* p, const float* elem_jac, float* vals, hf_stream_t stream);
/* HEX8 patterns (vec 3): hf_assemble_* with the rows summed per node over its
 * cells (8 lanes per row, the row's blocks in LDS) instead of per block over
 * contribution segments; same summation order and bits as hf_assemble_*
 * (get_A over the jacfwd element matrices, problem.py:262-266,
 * solver.py:466-515).  Rows must fit the LDS budget (max_row_blocks <= 42 in
 * FP64, 85 in FP32), else the call fails and hf_assemble_* applies. */
int hf_assemble_hex8_f64(const hf_pattern* p, const double* elem, double* vals, hf_stream_t stream);
int hf_assemble_hex8_f32(const hf_pattern* p, const float* elem, float* vals, hf_stream_t stream);
/* scalar CSR values (the order of hf_pattern_csr) from matrix values */
int hf_matrix_to_csr_f64(const hf_pattern* p, const double* vals, double* csr, hf_stream_t stream);

/* ---- Dirichlet row elimination (solver.py:290-363, 489-490). bc_dofs are
 *      unique DOF indices (later specs already applied), bc_vals their values. */
/* zeroRows(rows, diag=1) on the matrix values */
int hf_bc_rows_f64(const hf_pattern* p, const int32_t* bc_dofs, int64_t nbc, double* vals, hf_stream_t stream);
int hf_bc_rows_f32(const hf_pattern* p, const int32_t* bc_dofs, int64_t nbc, float* vals, hf_stream_t stream);
/* apply_bc_vec: R[d] = u[d] - scale*g[d] */
int hf_bc_residual_f64(const int32_t* bc_dofs, const double* bc_vals, int64_t nbc, double scale,
                       const double* u, double* R, hf_stream_t stream);
/* x0 = assign_bc(0) - copy_bc(u): x0[d] = g[d]-u[d] on B, 0 elsewhere (x0 of length N) */
int hf_bc_initial_guess_f64(const int32_t* bc_dofs, const double* bc_vals, int64_t nbc,
                            const double* u, double* x0, int64_t N, hf_stream_t stream);

/* ---- sparse matrix-vector products (the BCOO matvec of jax_solve, solver.py:67-84) */
int hf_spmv_f64(const hf_pattern* p, const double* vals, const double* x, double* y, hf_stream_t stream);
int hf_spmv_f32(const hf_pattern* p, const float* vals, const float* x, float* y, hf_stream_t stream);
/* y = A^T x through the transposed-block map (PETSc A.transpose, solver.py:1369-1370) */
int hf_spmv_t_f64(const hf_pattern* p, const double* vals, const double* x, double* y, hf_stream_t stream);
int hf_spmv_t_f32(const hf_pattern* p, const float* vals, const float* x, float* y, hf_stream_t stream);
/* inverse of the scalar diagonal (Jacobi M = 1/diag(A), solver.py:68-69); a zero
 * diagonal entry (the pressure block of a saddle-point system) gives 1 (unscaled),
 * where the reference's 1/0 would make the preconditioner infinite */
int hf_jacobi_f64(const hf_pattern* p, const double* vals, double* dinv, hf_stream_t stream);
int hf_jacobi_f32(const hf_pattern* p, const float* vals, float* dinv, hf_stream_t stream);

/* ---- device-resident Krylov solves: replace jax.scipy.sparse.linalg.bicgstab
 *      inside jax_solve (solver.py:78-84) and the linear_solver plugin slot
 *      (solver.py:261-284).  Stop rule ||r||^2 <= max(rtol^2 ||b||^2, atol^2)
 *      with r the unpreconditioned recursive residual (HF_KRYLOV_REL_R0: relative
 *      to the initial residual ||b - A x0|| instead of ||b||, the inexact-Newton
 *      form); dot products accumulate in FP64.  The host polls a device flag
 *      every `poll` iterations.  HF_KRYLOV_TRANSPOSE solves A^T x = b.
 *      workspace: hf_krylov_workspace_size. */
typedef struct hf_krylov_info {
  int32_t iterations;
  int32_t status;      /* 0 converged, 1 max iterations, 2 breakdown, 3 non-finite */
  double res_norm;     /* final recursive ||r||_2 */
  double b_norm;       /* ||b||_2 */
} hf_krylov_info;

enum { HF_KRYLOV_PCG = 0, HF_KRYLOV_BICGSTAB = 1 };
enum { HF_KRYLOV_TRANSPOSE = 1, HF_KRYLOV_REL_R0 = 2 };
size_t hf_krylov_workspace_size(int32_t method, int64_t N, int32_t elem_bytes);
int hf_krylov_solve_f64(int32_t method, const hf_pattern* p, const double* vals, const double* dinv,
                        const double* b, double* x, int32_t flags, int32_t maxiter,
                        double rtol, double atol, int32_t poll, void* workspace,
                        hf_krylov_info* info, hf_stream_t stream);
int hf_krylov_solve_f32(int32_t method, const hf_pattern* p, const float* vals, const float* dinv,
                        const float* b, float* x, int32_t flags, int32_t maxiter,
                        double rtol, double atol, int32_t poll, void* workspace,
                        hf_krylov_info* info, hf_stream_t stream);

/* ---- allocation-free Krylov solve.  No counterpart in the reference: it is
 *      for callers that cannot allocate device memory (code under HIP graph
 *      capture, tight serving loops, harnesses that forbid hipMalloc).
 * hf_init(): the library's set-up call for the current device.  Once per
 * device, with that device current: resolves the device address of the
 * library's static workspace array (64 MiB, held by every GPU the library runs
 * on until the process ends) and creates that device's poll state (two pinned
 * scalar slots, two events).  Idempotent; not thread-safe against itself.
 * hf_krylov_solve_noalloc_f64(): hf_krylov_solve_f64 (same methods, flags,
 * stop rule, status and bits) with that static array as the workspace; it calls
 * no hipMalloc / hipFree / allocating routine.  Fails (-1) unless hf_init() has
 * run on the current device, or when hf_krylov_workspace_size(method, N, 8)
 * exceeds 64 MiB (PCG: N <= ~1.67 M DOF; BiCGSTAB: N <= ~1.04 M DOF).  Every
 * call on one GPU shares that GPU's array and its poll state: calls that
 * overlap on one GPU (other streams or host threads) race. */
int hf_init(void);
int hf_krylov_solve_noalloc_f64(int32_t method, const hf_pattern* p, const double* vals, const double* dinv,
                                const double* b, double* x, int32_t flags, int32_t maxiter, double rtol,
                                double atol, int32_t poll, hf_krylov_info* info, hf_stream_t stream);

/* ---- learned preconditioner (no reference counterpart; the reference only has
 *      Jacobi, solver.py:68-69): a 3x3x3 convolutional network on the structured
 *      node grid (nx, ny, nz nodes, box_mesh order n = (i*ny + j)*nz + k,
 *      generate_mesh.py:170-187), channels-last tensors (batch, node, channel),
 *      weights W[tap][cin][cout] with tap = ((di+1)*3 + (dj+1))*3 + (dk+1),
 *      zero padding.  Built layers: 4->16, 16->16, 16->3.  act: 0 none, 1 tanh. */
int hf_cnn_conv3d_fwd_f32(const float* in, int32_t cin, const float* w, const float* b, int32_t cout,
                          int32_t act, int32_t nx, int32_t ny, int32_t nz, int32_t batch, float* out,
                          hf_stream_t stream);
/* gin = conv^T(gpre): gradient w.r.t. the layer input */
int hf_cnn_conv3d_bwd_data_f32(const float* gpre, int32_t cout, const float* w, int32_t cin, int32_t nx,
                               int32_t ny, int32_t nz, int32_t batch, float* gin, hf_stream_t stream);
/* gw, gb: weight and bias gradients, deterministic two-stage reduction */
size_t hf_cnn_bwd_weight_workspace(int32_t cin, int32_t cout, int32_t nx, int32_t ny, int32_t nz, int32_t batch);
int hf_cnn_conv3d_bwd_weight_f32(const float* in, int32_t cin, const float* gpre, int32_t cout, int32_t nx,
                                 int32_t ny, int32_t nz, int32_t batch, void* workspace, float* gw, float* gb,
                                 hf_stream_t stream);
/* gpre = gout * (1 - y^2) */
int hf_cnn_tanh_bwd_f32(const float* y, const float* gout, int64_t n, float* gpre, hf_stream_t stream);
/* preconditioner input f = [D^-1/2 r / s, g] and output z = D^-1 r + s D^-1/2 mask y
 * (r, z: (batch, 3*nnode); dinv, mask: 3*nnode; g: nnode; scale: s per batch entry) */
int hf_lpc_features_f32(const float* r, const float* dinv, const float* gfield, int64_t nnode, int32_t batch,
                        const float* scale, float* f, hf_stream_t stream);
int hf_lpc_output_f32(const float* r, const float* dinv, const float* mask, const float* y, int64_t nnode,
                      int32_t batch, const float* scale, float* z, hf_stream_t stream);
/* the preconditioner applied once, z = M(r), on the inference path the solve
 * uses (MFMA f16 x f16 -> f32 convolutions; s computed on the device).  The
 * FP32 layer calls above are the training path it is tested against. */
size_t hf_lpc_apply_workspace_size(int64_t nnode);
int hf_lpc_apply_f32(const float* r, const float* dinv, const float* gfield, const float* mask,
                     const float* const* weights, int32_t nx, int32_t ny, int32_t nz, void* workspace, float* z,
                     hf_stream_t stream);
/* truncated flexible CG (each direction A-orthogonal to the last `trunc`,
 * 1..4) with the learned preconditioner (weights: w1,b1,w2,b2,w3,b3 device
 * pointers); the network correction is scaled down where needed so that
 * r.z >= 0.1 r.D^-1.r.  status 0 converged, 1 max iterations, 2 p.Ap <= 0,
 * 3 non-finite, 4 stagnation (after ||r|| first fell 10 % below its start: no new
 * 10 % best in max(stall_window, iterations / 2) iterations; before: 10 windows);
 * on 2-4 the caller continues with Jacobi-PCG. */
size_t hf_lpc_workspace_size(int64_t nnode);
int hf_lpc_solve_f32(const hf_pattern* p, const float* vals, const float* dinv, const float* b, float* x,
                     const float* const* weights, const float* gfield, const float* mask, int32_t nx,
                     int32_t ny, int32_t nz, int32_t maxiter, double rtol, double atol, int32_t poll,
                     int32_t flags, int32_t stall_window, int32_t trunc, void* workspace,
                     hf_krylov_info* info, hf_stream_t stream);

/* ---- geometric multigrid on the structured node lattice (no reference
 *      counterpart; the reference only has Jacobi, solver.py:68-69).  Box mesh
 *      node order n = (i*ny + j)*nz + k (generate_mesh.py:170-187), vec = 3,
 *      FP32.  Level 0 is the tangent with its Dirichlet DOFs masked out of rows
 *      and columns; each coarser level halves the intervals per dimension and
 *      holds the Galerkin product P^T A P of trilinear interpolation as a
 *      27-point node-block stencil, stored slot-major
 *      S[(o*9 + 3i + j)*nnode + n], o = (dx+1)*9 + (dy+1)*3 + (dz+1); the
 *      coarsest (<= 88 DOF) is inverted densely in FP64.  V-cycles with a
 *      polynomial smoother in Dinv A (Dinv: inverse 3x3 diagonal blocks) of
 *      `degree` steps on each side: Chebyshev on [lam/ratio, lam] with lam the
 *      Gershgorin bound of Dinv A (default: degree 2, ratio 8), or damped
 *      block Jacobi (param = omega). */
typedef struct hf_mg hf_mg;
enum { HF_MG_JACOBI = 0, HF_MG_CHEBYSHEV = 1 };
/* The fine-level stencil of a symmetric operator is stored half: the diagonal and
 * the 13 upper offsets, block (n, o) of a lower offset read as the transpose of
 * the neighbour's (half the bytes of the cycle's largest stream).
 * HF_MG_FINE_FULL stores all 27 slots (a non-symmetric operator).
 * HF_MG_FINE_SELL: the fine level is the tangent's SELL matrix itself (elements
 * whose nodes couple two lattice steps apart: HEX27 on its (2n+1)^3 node lattice).
 * The smoother then runs SELL products, and level 1 = P^T A P is formed from the
 * SELL rows; with P the interpolation from every second node it is the 27-point
 * HEX8 operator of the same cells.  Couplings whose coarse parents are more than
 * one coarse step apart are a set-up error.
 * HF_MG_FINE_F16: the smoothing passes (and the restricted residuals) of every
 * half-stored level the grid kernels run read an FP16 copy of its stencil, made at
 * each set-up: each block's 9 entries in 5 half2 words, scaled by the power of two
 * that puts the level's largest magnitude in [2^14, 2^15) (280 bytes a node
 * instead of 504).  The outer flexible CG's products keep the FP32 stencil. */
enum { HF_MG_FINE_FULL = 1, HF_MG_FINE_SELL = 2, HF_MG_FINE_F16 = 4 };
int hf_mg_create(int32_t nx, int32_t ny, int32_t nz, int32_t flags, hf_mg** out);   /* nodes per dimension */
int hf_mg_destroy(hf_mg* mg);
/* number of levels; dims (3 per level, up to max_levels) receives the node counts */
int hf_mg_levels(const hf_mg* mg, int32_t* nlev, int32_t* dims, int32_t max_levels);
/* after a set-up: recomputes the coefficients on the default stream and
 * synchronises the device */
int hf_mg_set_smoother(hf_mg* mg, int32_t kind, int32_t degree, float param);
/* post-smoothing steps on the levels run by grid kernels, 1..degree (0: the
 * smoother's degree, the default): the first steps of the same recurrence, i.e.
 * the lower-degree Chebyshev polynomial on the same interval.  The flexible CG
 * accepts the then non-symmetric cycle. */
int hf_mg_set_post_degree(hf_mg* mg, int32_t degree);
/* Matrix-free fine level: for a box lattice whose cells share one geometry and one
 * linear law up to a per-cell factor, the operator is A = sum_c cell_scale[c] K,
 * K the 24x24 cell matrix at unit scale.  k (host, 576 floats): 3x3 block (q, q')
 * of K, row-major, at k[(8 q + q') * 9], corners q = 4 dx + 2 dy + dz (dx, dy, dz in
 * {0, 1}); cell_scale (device, (nx-1)(ny-1)(nz-1) in box cell order
 * (cx (ny-1) + cy)(nz-1) + cz) is kept by reference.  Once set, the fine level's
 * smoothing passes, the flexible CG's products and hf_mg_stencil_matvec_f32 with a
 * null stencil compute each row from the node's 8 cells, applied to neighbour
 * differences (504 FMAs) instead of reading a stencil.  K must be symmetric under
 * the cell's mid-plane reflections (block (a, b) = R_a block (0, a^b) R_a) and its
 * block rows must sum to zero (translations carry no force); both are checked.
 * The set-up still reads the assembled stencil (block inverses, bound, coarse
 * levels), and the caller vouches that both are the same operator.  A new k
 * synchronises the device once; cell_scale NULL switches it off.  No reference
 * counterpart (the reference assembles every tangent, solver.py:466-515). */
int hf_mg_set_cell_operator(hf_mg* mg, const float* k, const float* cell_scale);
/* the half-stored level-0 stencil (hf_mg_stencil_size floats) of the cell operator,
 * for hf_mg_setup_stencil_f32: replaces the element Jacobians + hf_mg_assemble_hex8_f32
 * when the tangent has that form.  A set-up with a cell operator set makes no FP16
 * copy (HF_MG_FINE_F16): switching it off then needs a new set-up. */
int hf_mg_assemble_cellop_f32(const hf_mg* mg, float* S, hf_stream_t stream);
/* The learned network (the layers of hf_lpc_apply_f32: weights w1,b1,w2,b2,w3,b3 as
 * device pointers, gfield the nnode input field, kept by reference) as the fine
 * level's post-smoother: after the coarse correction one step
 *   e += D^-1 res + s D^-1/2 mask N([D^-1/2 res / s, g]),  res = M (r - A e),
 *   s = rms(D^-1/2 res), D = diag of the fine operator,
 * replaces the Chebyshev post-smoothing (mode 1) or follows it (mode 2); mode 0
 * switches it off.  The weights are packed on `stream` into the hierarchy's own
 * buffer (allocated on the first call); a lattice-stencil fine level (not
 * HF_MG_FINE_SELL) and at least two levels are required.  No reference
 * counterpart (the reference preconditions with Jacobi, solver.py:68-69). */
int hf_mg_set_learned(hf_mg* mg, int32_t mode, const float* const* weights, const float* gfield,
                      hf_stream_t stream);
/* lam of every level but the coarsest from the last set-up (synchronises the device) */
int hf_mg_bounds(const hf_mg* mg, float* lam, int32_t max_levels);
/* builds every level from the FP32 tangent (pattern + SELL values) and the
 * level-0 DOF mask (1 free, 0 Dirichlet; kept by reference until the next
 * set-up); *dropped (optional) = coarsest DOFs without support.  Fails if the
 * pattern couples nodes that are not lattice neighbours.  Synchronises the
 * stream on a pattern's first set-up, with a SELL fine level, or when dropped is
 * given (the checks read device flags); otherwise it only enqueues, and the
 * level data and coefficients are ready for work queued behind it on `stream`. */
int hf_mg_setup_f32(hf_mg* mg, const hf_pattern* p, const float* vals, const float* mask, int32_t* dropped,
                    hf_stream_t stream);
/* z = V(r), one cycle; HF_MG_SKIP_FINE_POST stops after the fine level's coarse
 * correction (no fine post-smoothing: the residual r - A z is then what a
 * fine-level post-smoother sees) */
enum { HF_MG_SKIP_FINE_POST = 1 };
int hf_mg_apply_f32(hf_mg* mg, const float* r, float* z, int32_t flags, hf_stream_t stream);
/* copy level `level`'s stencil (243 * nnode floats, all 27 slots; a half-stored level expanded) into dst */
int hf_mg_copy_stencil(const hf_mg* mg, int32_t level, float* dst, hf_stream_t stream);
/* flexible CG (one kept direction) preconditioned by V-cycles; the stop rule,
 * flags (HF_KRYLOV_REL_R0; no transpose) and info of hf_krylov_solve.
 * status 2: p.Ap <= 0.  workspace: hf_mg_workspace_size.
 * A poll batch (poll iterations) is captured once as a HIP graph and replayed;
 * it is captured again when the hierarchy is set up again or the pattern,
 * values pointer, workspace or poll change.  HF_MG_SOLVE_EAGER enqueues every
 * launch instead (the same kernels in the same order: identical results).
 * The direction scalars are summed inside the update kernel while the operator
 * grid has at most 2048 workgroups, by a one-workgroup kernel above that;
 * HF_MG_SOLVE_SPLIT_SCALARS takes the second form at every size (same
 * iterates up to the summation order of those scalars).
 * Calls that share one hf_mg must not overlap.  info is final on return; when
 * the solve converged, the copy of the iterate into x may still be queued on
 * `stream` (work enqueued behind it on that stream sees x). */
enum { HF_MG_SOLVE_EAGER = 4, HF_MG_SOLVE_SPLIT_SCALARS = 8 };
size_t hf_mg_workspace_size(const hf_mg* mg);
int hf_mg_solve_f32(hf_mg* mg, const hf_pattern* p, const float* vals, const float* b, float* x, int32_t flags,
                    int32_t maxiter, double rtol, double atol, int32_t poll, void* workspace,
                    hf_krylov_info* info, hf_stream_t stream);

/* ---- the multigrid's own form of a HEX8 lattice tangent, without the SELL
 *      matrix.  No reference counterpart as a format: it replaces get_A's
 *      COO -> CSR (solver.py:466-515) plus the row elimination of
 *      apply_bc (solver.py:290-363) for the multigrid path on box lattices.
 * The stencil holds, per node, the diagonal block and the blocks toward its 13
 * upper lattice neighbours (offset o = (dx+1)*9 + (dy+1)*3 + (dz+1) > 13),
 * slot-major: S[((o - 13) * 9 + 3i + j) * nnode + n] couples DOF i of node n to
 * DOF j of the neighbour.  Dirichlet rows are kept; the operator is the
 * row-eliminated A: (A v)_d = v_d on Dirichlet DOFs d, with the mask of the
 * hierarchy's last set-up.  hf_mg_stencil_size: floats of one stencil. */
size_t hf_mg_stencil_size(const hf_mg* mg);
/* HEX8 element Jacobians (as hf_cell_jacobian_f32) restricted to the blocks
 * (a, b) whose column node id is >= the row node's, cell-innermost:
 * out[((a * 8 + b) * 9 + 3i + k) * ncell + c] (9 * 64 * ncell floats; the other
 * blocks are left unwritten) */
int hf_hex8_jacobian_upper_f32(const hf_material* mat, const double* points, const int32_t* cells, int64_t ncell,
                               const double* u, const double* params, float* out, hf_stream_t stream);
/* assemble S from hf_hex8_jacobian_upper_f32's blocks (HEX8, vec 3; cells: the
 * pattern's cells, ncell x 8): each block sums its cell contributions in
 * ascending cell order, as hf_assemble_hex8_f32 does.  corner (host, 8 ints, or
 * NULL): the cells are the lattice's, cell (i, j, k) = (i * (ny-1) + j) * (nz-1) + k
 * with its corner (dx, dy, dz) at local index corner[4 dx + 2 dy + dz] (a
 * box_mesh); the contributions then follow from the coordinates (the caller
 * vouches for the layout: a wrong table gives a wrong stencil).  NULL: the
 * cells are read through the pattern's contribution lists. */
int hf_mg_assemble_hex8_f32(const hf_mg* mg, const hf_pattern* p, const int32_t* cells, int64_t ncell,
                            const int32_t* corner, const float* elem, float* S, hf_stream_t stream);
/* hf_mg_setup_f32 with level 0 taken from S (a device copy); the pattern is
 * still checked to be the lattice stencil */
int hf_mg_setup_stencil_f32(hf_mg* mg, const hf_pattern* p, const float* S, const float* mask, int32_t* dropped,
                            hf_stream_t stream);
/* hf_mg_solve_f32 with S as the operator */
int hf_mg_solve_stencil_f32(hf_mg* mg, const float* S, const float* b, float* x, int32_t flags, int32_t maxiter,
                            double rtol, double atol, int32_t poll, void* workspace, hf_krylov_info* info,
                            hf_stream_t stream);
/* y = A x with S (the row-eliminated operator), or with the cell operator of
 * hf_mg_set_cell_operator when S is NULL */
int hf_mg_stencil_matvec_f32(const hf_mg* mg, const float* S, const float* x, float* y, hf_stream_t stream);

/* ---- measurement: HIP-event timing of the Krylov SpMV kernel (the first SpMV
 *      of every poll batch, on the solve's stream).  No reference counterpart:
 *      replaces the reference's wall-clock buckets (solver.py:22-57) at kernel
 *      granularity.  hf_profile_enable also resets the counters. */
int hf_profile_enable(int32_t on);
int hf_profile_read(int64_t* launches, double* total_ms);

/* ---- topology optimisation: filters and the MMA subproblem (jax_fem/mma.py) ----
 * Filters replace applySensitivityFilter / applyDensityFilter (mma.py:58-66)
 * over compute_filter_kd_tree's H (mma.py:27-56), held as a k-wide neighbour
 * table: row i of H has entries w[i*k+t] at columns nbr[i*k+t]; hs = row sums.
 *   sensitivity: out[c*n+i] = sum_t w[i*k+t] rho[j] g[c*n+j] / max(1e-3, rho[j]) / hs[j],
 *                j = nbr[i*k+t], for the ncols gradients g (objective, constraints)
 *   density:     out[i] = sum_t w[i*k+t] rho[nbr[i*k+t]] / hs[i]                    */
int hf_mma_sensitivity_filter_f64(const int32_t* nbr, const double* w, const double* hs, int64_t n, int32_t k,
                                  const double* rho, const double* g, int32_t ncols, double* out,
                                  hf_stream_t stream);
int hf_mma_density_filter_f64(const int32_t* nbr, const double* w, const double* hs, int64_t n, int32_t k,
                              const double* rho, double* out, hf_stream_t stream);
/* One MMA subproblem: replaces MMA.mmasub + subsolv (mma.py:114-410, the
 * m < n branch).  n design variables, m = 1..4 constraints.  Device arrays
 * (FP64): xval, xold1, xold2, xmin, xmax, df0dx (n); fval (m); dfdx (m, n)
 * row-major; low, upp (n) in/out (the asymptotes, read when iter > 2);
 * xmma (n) out.  Host arrays: a, c, d (m).  iter counts from 1 as MMA.epoch.
 * Synchronises the stream (the backtracking line search reads residual norms). */
typedef struct hf_mma hf_mma;
typedef struct {
  int32_t newton_steps;   /* primal-dual Newton steps over all epsi levels */
  int32_t trials;         /* residual evaluations in the backtracking */
  double res_max;         /* max-norm of the KKT residual at the returned point */
  double res_norm;
} hf_mma_info;
int hf_mma_create(int64_t n, int32_t m, hf_mma** out);
int hf_mma_destroy(hf_mma* mma);
int hf_mma_update_f64(hf_mma* mma, int32_t iter, const double* xval, const double* xold1, const double* xold2,
                      const double* xmin, const double* xmax, double* low, double* upp, const double* df0dx,
                      const double* fval, const double* dfdx, double a0, const double* a, const double* c,
                      const double* d, double move, double* xmma, hf_mma_info* info, hf_stream_t stream);

/* ---- small vector helpers (FP64 accumulate) ---- */
size_t hf_dot_workspace_size(void);
/* *result (host) = sum x*y accumulated in FP64, deterministic order */
int hf_dot_f64(const double* x, const double* y, int64_t n, void* workspace, double* result, hf_stream_t stream);
int hf_dot_f32(const float* x, const float* y, int64_t n, void* workspace, double* result, hf_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* HIPFEM_H */
