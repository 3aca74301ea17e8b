"""Problem assembly, row elimination, Krylov, Newton and adjoint (oracle, FP64).

Restates:
  * COO pattern (problem.py:86-107): DOF = vec*node + comp; I[e,r,c] = dof[e,r],
    J[e,r,c] = dof[e,c] row-major per cell, cells first, then each surface's faces.
  * residual scatter (problem.py:426-437) and ``newton_update`` values V (:447-460).
  * ``apply_bc_vec`` / ``assign_bc`` / ``copy_bc`` (solver.py:290-363):
    R[B] = u[B] - g; x0 = g - u on B, 0 elsewhere; later specs win on overlap.
  * ``_PetscTangentCache`` + ``get_A`` (solver.py:466-515): pattern = sorted unique
    (I,J) with explicit zeros kept, duplicates summed in FP64, BC rows zeroed with
    1 on the diagonal, columns kept.  scipy's coo->csr with summed duplicates and
    sorted indices gives the same pattern as PETSc's COO preallocation.
  * ``jax_solve`` (solver.py:63-92): Jacobi M = 1/diag(A), BiCGSTAB with tol=atol=1e-10,
    maxiter=10000, post-check ||Ax-b|| < 0.1.  The BiCGSTAB iteration restates
    jax.scipy.sparse.linalg.bicgstab (JAX, unpinned by the reference; van der
    Vorst 1992 with jax's stop rule ||r||^2 <= max(tol^2 ||b||^2, atol^2), r the
    unpreconditioned residual, breakdown when rho, alpha or omega is 0).
  * ``solver`` Newton (solver.py:1074-1318): loop while ||R||/||R0|| > rel_tol
    (1e-8) and ||R|| > tol (1e-6); u += A^-1 (-R).
  * ``implicit_vjp`` (solver.py:1324-1380): re-linearise at u*, A^T lam = v,
    grad = -(dc/dp)^T lam with c the BC-applied residual.

Pinned by: scipy's coo->csr (library routine; pattern bit-exact), the FEniCSx
fixtures (poisson u = x, cube, cylinder, hyperelastic cylinder + traction),
scipy.sparse.linalg.spsolve on the same systems, the compute_gradients notebook
value dscale_d = 431.59223609853564, and central finite differences of the
adjoint gradient.
"""
import numpy as np
import scipy.sparse as sp

from . import basis, fe, kernels


class OracleProblem:
    """NumPy counterpart of the reference ``Problem`` for one variable."""

    def __init__(self, points, cells, vec, material, ele_type="HEX8", quadrature_order=None,
                 dirichlet_bc_info=None, location_fns=None, surface_maps=None, mass_map=None,
                 params=None, mass_params=None):
        self.points = np.asarray(points, dtype=np.float64)
        self.cells = np.asarray(cells, dtype=np.int64)
        self.vec, self.mat, self.ele_type = vec, material, ele_type
        self.num_nodes = len(self.points)
        self.N = self.num_nodes * vec
        sv, sgr, w = basis.get_shape_vals_and_grads(ele_type, quadrature_order)
        fsv, fsg, fw, fnorm, finds = basis.get_face_shape_vals_and_grads(ele_type, quadrature_order)
        self.shape_vals, self.quad_weights = sv, w
        self.grads, self.JxW = fe.shape_grads(self.points, self.cells, sgr, w)
        self.quad_points = fe.physical_quad_points(self.points, self.cells, sv)
        self.dirichlet_bc_info = dirichlet_bc_info
        self.update_bcs(dirichlet_bc_info)
        self.boundary_inds_list = fe.boundary_inds(self.points, self.cells, location_fns, finds)
        self.surface_maps = surface_maps or []
        assert len(self.surface_maps) == len(self.boundary_inds_list)
        self.mass_map = mass_map
        self.face_data = []
        for bi in self.boundary_inds_list:
            _, nanson = fe.face_shape_grads(self.points, self.cells, bi, fsg, fnorm, fw)
            xs = fe.physical_surface_quad_points(self.points, self.cells, bi, fsv)
            self.face_data.append((bi, fsv[bi[:, 1]], nanson, xs))
        self.params = params
        self.mass_params = mass_params      # internal variable of the mass map, (ncell, nq)
        self._csr = None

    # -- Dirichlet data (fe.py:220-269) --------------------------------------
    def update_bcs(self, dirichlet_bc_info):
        self.node_inds_list, self.vec_inds_list, self.vals_list = fe.dirichlet_bcs(
            self.points, dirichlet_bc_info)

    def bc_rows(self):
        return [n * self.vec + v for n, v in zip(self.node_inds_list, self.vec_inds_list)]

    # -- COO pattern (problem.py:86-107) -------------------------------------
    def coo_pattern(self):
        def dofs(c):
            return (self.vec * c[:, :, None] + np.arange(self.vec)[None, None, :]).reshape(len(c), -1)
        inds = dofs(self.cells)
        I = [np.repeat(inds[:, :, None], inds.shape[1], axis=2).reshape(-1)]
        J = [np.repeat(inds[:, None, :], inds.shape[1], axis=1).reshape(-1)]
        for bi, *_ in self.face_data:
            fi = dofs(self.cells[bi[:, 0]])
            I.append(np.repeat(fi[:, :, None], fi.shape[1], axis=2).reshape(-1))
            J.append(np.repeat(fi[:, None, :], fi.shape[1], axis=1).reshape(-1))
        return np.concatenate(I), np.concatenate(J)

    # -- residual and Jacobian (problem.py:439-491) -------------------------
    def cell_u(self, u):
        return u.reshape(self.num_nodes, self.vec)[self.cells]

    def external(self):
        """Element vectors of the u-independent surface and mass maps, scattered."""
        f = np.zeros((self.num_nodes, self.vec))
        if self.mass_map is not None:
            r = kernels.mass_residual(self.mass_map, self.quad_points, self.shape_vals, self.JxW,
                                      self.mass_params)
            np.add.at(f, self.cells.reshape(-1), r.reshape(-1, self.vec))
        for (bi, fvals, nanson, xs), smap in zip(self.face_data, self.surface_maps):
            r = kernels.surface_residual(smap, xs, fvals, nanson)
            np.add.at(f, self.cells[bi[:, 0]].reshape(-1), r.reshape(-1, self.vec))
        return f.reshape(-1)

    def compute_residual(self, u):
        r = kernels.cell_residual(self.mat, self.cell_u(u), self.grads, self.JxW, self.params)
        R = np.zeros((self.num_nodes, self.vec))
        np.add.at(R, self.cells.reshape(-1), r.reshape(-1, self.vec))
        return R.reshape(-1) + self.external()

    def element_jacobians(self, u):
        return kernels.cell_jacobian(self.mat, self.cell_u(u), self.grads, self.JxW, self.params)

    def assemble(self, u, bc=True):
        """get_A: CSR with the reference pattern, FP64 sums, BC rows -> identity rows."""
        K = self.element_jacobians(u)
        I, J = self.coo_pattern()
        nface = len(I) - K.size
        V = np.concatenate([K.reshape(-1), np.zeros(nface)])
        A = sp.coo_matrix((V, (I, J)), shape=(self.N, self.N)).tocsr()
        A.sum_duplicates()
        A.sort_indices()
        if bc:
            zero_rows_identity(A, self.bc_rows())
        return A

    # -- BC vectors (solver.py:290-363) --------------------------------------
    def apply_bc_vec(self, res, u, scale=1.0):
        res = res.copy()
        for rows, vals in zip(self.bc_rows(), self.vals_list):
            res[rows] = u[rows]
            res[rows] -= vals * scale
        return res

    def assign_bc(self, u):
        u = u.copy()
        for rows, vals in zip(self.bc_rows(), self.vals_list):
            u[rows] = vals
        return u

    def copy_bc(self, u):
        out = np.zeros_like(u)
        for rows in self.bc_rows():
            out[rows] = u[rows]
        return out


def zero_rows_identity(A, rows_list):
    """PETSc zeroRows(rows, diag=1) on a CSR matrix in place (solver.py:489-490)."""
    for rows in rows_list:
        for r in np.asarray(rows):
            s, e = A.indptr[r], A.indptr[r + 1]
            A.data[s:e] = 0.0
            d = s + np.searchsorted(A.indices[s:e], r)
            A.data[d] = 1.0


def bicgstab(A, b, x0=None, M=None, tol=1e-10, atol=1e-10, maxiter=10000):
    """jax.scipy.sparse.linalg.bicgstab semantics (solver.py:78-84). Returns (x, iters, info)."""
    Mf = (lambda v: v) if M is None else M
    x = np.zeros_like(b) if x0 is None else x0.copy()
    bs = float(b @ b)
    atol2 = max(tol * tol * bs, atol * atol)
    r = b - A @ x
    rhat = r.copy()
    alpha = omega = rho = 1.0
    p = r.copy()
    q = r.copy()
    k = 0
    while float(r @ r) > atol2 and k < maxiter and k >= 0:
        rho_ = float(rhat @ r)
        beta = rho_ / rho * alpha / omega
        p = r + beta * (p - omega * q)
        phat = Mf(p)
        q = A @ phat
        alpha = rho_ / float(rhat @ q)
        s = r - alpha * q
        if float(s @ s) < atol2:
            x = x + alpha * phat
            r = s
            omega_ = omega
        else:
            shat = Mf(s)
            t = A @ shat
            omega_ = float(t @ s) / float(t @ t)
            x = x + alpha * phat + omega_ * shat
            r = s - omega_ * t
        k = -11 if (omega_ == 0 or alpha == 0) else k + 1
        k = -10 if rho_ == 0 else k
        rho, omega = rho_, omega_
    return x, k, 0 if k >= 0 else k


def pcg(A, b, x0, M, tol=1e-10, atol=1e-10, maxiter=10000):
    """Preconditioned CG with the same stop rule as ``bicgstab`` (used for cross-checks)."""
    x = x0.copy()
    r = b - A @ x
    atol2 = max(tol * tol * float(b @ b), atol * atol)
    z = M(r)
    p = z.copy()
    rz = float(r @ z)
    k = 0
    while float(r @ r) > atol2 and k < maxiter:
        q = A @ p
        alpha = rz / float(p @ q)
        x += alpha * p
        r -= alpha * q
        z = M(r)
        rz_new = float(r @ z)
        p = z + (rz_new / rz) * p
        rz = rz_new
        k += 1
    return x, k


def linear_solve(A, b, x0, method="bicgstab", tol=1e-10, atol=1e-10, maxiter=10000):
    """jax_solve (solver.py:63-92) with the Jacobi preconditioner."""
    d = A.diagonal()
    M = lambda v: v / d
    if method == "bicgstab":
        x, k, _ = bicgstab(A, b, x0, M, tol, atol, maxiter)
    elif method == "cg":
        x, k = pcg(A, b, np.zeros_like(b) if x0 is None else x0, M, tol, atol, maxiter)
    else:
        import scipy.sparse.linalg as spla
        x, k = spla.spsolve(A.tocsc(), b), 1
    err = np.linalg.norm(A @ x - b)
    assert err < 0.1, "linear solver failed to converge with err = %g" % err
    return x, k


def line_search(prob, u, inc):
    """line_search (solver.py:424-464): from alpha = 1, halve up to 3 times while
    the BC-applied residual norm keeps decreasing; returns (alpha, new u)."""
    def norm_at(alpha):
        v = u + alpha * inc
        return float(np.linalg.norm(prob.apply_bc_vec(prob.compute_residual(v), v)))
    alpha = 1.0
    res = norm_at(alpha)
    for _ in range(3):
        alpha *= 0.5
        half = norm_at(alpha)
        if half > res:
            alpha *= 2.0
            break
        res = half
    return alpha, u + alpha * inc


def newton_solve(prob, u0=None, rel_tol=1e-8, tol=1e-6, method="bicgstab", lin_tol=1e-10,
                 max_newton=50, log=None, line_search_flag=False):
    """solver() Newton loop (solver.py:1247-1318), with the optional line search of
    newton_step (solver.py:416-417). Returns (u, info dict)."""
    u = np.zeros(prob.N) if u0 is None else np.asarray(u0, dtype=np.float64).copy()
    R = prob.apply_bc_vec(prob.compute_residual(u), u)
    A = prob.assemble(u)
    res0 = res = float(np.linalg.norm(R))
    hist = [res]
    lin_iters = []
    alphas = []
    n = 0
    with np.errstate(invalid="ignore", divide="ignore"):
        first = np.float64(res) / np.float64(res0)    # NaN for a zero start: no step, as the reference
    while (first if n == 0 else res / res0) > rel_tol and res > tol and n < max_newton:
        n += 1
        x0 = prob.assign_bc(np.zeros(prob.N)) - prob.copy_bc(u)
        inc, k = linear_solve(A, -R, x0, method, lin_tol, lin_tol)
        lin_iters.append(k)
        if line_search_flag:
            alpha, u = line_search(prob, u, inc)
            alphas.append(alpha)
        else:
            u = u + inc
        R = prob.apply_bc_vec(prob.compute_residual(u), u)
        A = prob.assemble(u)
        res = float(np.linalg.norm(R))
        hist.append(res)
        if log:
            log("newton %d res %.3e rel %.3e lin_iters %d" % (n, res, res / res0, k))
    assert np.isfinite(res) and np.all(np.isfinite(u))
    return u, {"newton_iters": n, "res_hist": hist, "lin_iters": lin_iters, "alphas": alphas}


def implicit_vjp(prob, u, v, method="bicgstab", dirichlet_param_grads=None, wrt="material"):
    """implicit_vjp (solver.py:1324-1380) for per-quad material parameters
    (wrt="material") or the mass map's internal variable (wrt="mass").

    Returns (grad_p (ncell,nq) or None, lam).  ``dirichlet_param_grads`` is an
    optional list of (rows, dg_dp) pairs: the derivative of the Dirichlet values
    of spec rows with respect to a scalar parameter, for which the gradient
    sum(lam_B * dg/dp) is returned in the list ``dir_grads``."""
    A = prob.assemble(u)
    At = A.T.tocsr()
    lam, _ = linear_solve(At, v, None, method)
    lam_I = lam.copy()
    bc_all = np.concatenate(prob.bc_rows()) if prob.bc_rows() else np.zeros(0, dtype=np.int64)
    lam_I[bc_all] = 0.0
    grad_p = None
    if wrt == "mass":
        cl = lam_I.reshape(prob.num_nodes, prob.vec)[prob.cells]
        grad_p = -kernels.mass_theta_vjp(prob.mass_map, prob.quad_points, prob.shape_vals, prob.JxW,
                                         prob.mass_params, cl)
    elif prob.params is not None:
        cl = lam_I.reshape(prob.num_nodes, prob.vec)[prob.cells]
        grad_p = -kernels.cell_param_vjp(prob.mat, prob.cell_u(u), prob.grads, prob.JxW,
                                         prob.params, cl)
    dir_grads = []
    for rows, dg in (dirichlet_param_grads or []):
        # c_B = u_B - g(p): d(lam.c)/dp = -sum lam_B dg/dp; grad = -that.
        dir_grads.append(float(np.sum(lam[rows] * dg)))
    return grad_p, lam, dir_grads
