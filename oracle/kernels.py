"""Constitutive maps, cell residuals and element Jacobians (oracle, FP64).

Restates the reference's laplace / mass / surface kernels and their element
Jacobians (problem.py:189-266, 269-301), vectorised over cells instead of
``jax.vmap``, with the constitutive tangents written in closed form instead of
forward-mode AD:

  * Laplace kernel (problem.py:189-214):
      u_grad[q,i,j] = sum_a u[a,i] dN_a/dx_j
      r[a,i]        = sum_q sum_j P(u_grad)[q,i,j] dN_a/dx_j JxW[q]
  * element Jacobian (problem.py:262-266, jacfwd with out_axes=-1):
      K[a*vec+i, b*vec+k] = dr[a,i]/du[b,k]
                          = sum_q sum_{j,l} dN_a/dx_j A[i,j,k,l] dN_b/dx_l JxW[q]
  * mass kernel (problem.py:216-236) and surface kernel (problem.py:238-259)
    for maps that do not depend on u (their Jacobian contribution is zero).

Materials (the test problems' tensor maps):
  * Poisson: P = k u_grad (tests/benchmarks/linear_poisson/test_linear_poisson.py:15-17).
  * Linear elasticity: P = lam tr(eps) I + 2 mu eps
    (tests/benchmarks/linear_elasticity_cube/test_linear_elasticity_cube.py:15-25).
  * Compressible Neo-Hookean, psi = mu/2 (J^-2/3 I1 - 3) + kappa/2 (J-1)^2
    (tests/benchmarks/hyperelasticity/test_hyper_elasticity.py:15-34):
      P = mu J^-2/3 (F - I1/3 F^-T) + kappa J (J-1) F^-T.
  Per-quadrature parameters p (the reference's ``internal_vars``,
  problem.py:493-556) enter through the moduli: E = E*p ("scale_E", the
  compute_gradients notebook), SIMP E = Emin + (E0-Emin) p^penal ("simp"), or
  mu = p with fixed kappa ("mu").

Pinned by: central finite differences of P (tangent) and of r (element
Jacobian), symmetry of K, rigid-body null space, r = dPi/du by FD of the
energy, the textbook linear-elastic cube stiffness, and end-to-end by the
FEniCSx fixtures (cube, cylinder, hyperelastic cylinder).
"""
import numpy as np


class Material:
    """Constitutive law with optional per-quadrature parameter."""

    def __init__(self, kind, E=1.0, nu=0.3, param=None, E0=1.0, Emin=1e-3, penal=3.0,
                 kappa=None, k=1.0):
        assert kind in ("poisson", "linear_elastic", "neo_hookean")
        assert param in (None, "scale_E", "simp", "mu", "scale_k")
        self.kind, self.E, self.nu, self.param = kind, E, nu, param
        self.E0, self.Emin, self.penal, self.kappa, self.k = E0, Emin, penal, kappa, k

    def _from_E(self, E):
        nu = self.nu
        mu = E / (2.0 * (1.0 + nu))
        if self.kind == "linear_elastic":
            return mu, E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        return mu, E / (3.0 * (1.0 - 2.0 * nu))

    def moduli(self, p):
        """-> (m1, m2, dm1/dp, dm2/dp); (mu, lam) or (mu, kappa) or (k, 0)."""
        if self.kind == "poisson":
            if self.param == "scale_k":
                return self.k * p, 0.0 * p, self.k + 0.0 * p, 0.0 * p
            return self.k, 0.0, 0.0, 0.0
        if self.param is None:
            m1, m2 = self._from_E(self.E)
            return m1, m2, 0.0, 0.0
        if self.param == "scale_E":
            m1, m2 = self._from_E(self.E * p)
            d1, d2 = self._from_E(self.E)
            return m1, m2, d1 + 0.0 * p, d2 + 0.0 * p
        if self.param == "simp":
            E = self.Emin + (self.E0 - self.Emin) * p ** self.penal
            dE = (self.E0 - self.Emin) * self.penal * p ** (self.penal - 1.0)
            m1, m2 = self._from_E(E)
            d1, d2 = self._from_E(1.0)
            return m1, m2, d1 * dE, d2 * dE
        if self.param == "mu":
            return p, self.kappa + 0.0 * p, 1.0 + 0.0 * p, 0.0 * p
        raise ValueError(self.param)

    # -- stress split P = m1*P1 + m2*P2 ---------------------------------------
    def stress_parts(self, ug):
        if self.kind == "poisson":
            return ug, np.zeros_like(ug)
        I = np.eye(3)
        if self.kind == "linear_elastic":
            eps = 0.5 * (ug + np.swapaxes(ug, -1, -2))
            tr = np.trace(eps, axis1=-2, axis2=-1)[..., None, None]
            return 2.0 * eps, tr * I
        F = ug + I
        J = np.linalg.det(F)
        H = np.swapaxes(np.linalg.inv(F), -1, -2)
        I1 = np.sum(F * F, axis=(-2, -1))
        c = J ** (-2.0 / 3.0)
        P1 = c[..., None, None] * (F - (I1 / 3.0)[..., None, None] * H)
        P2 = (J * (J - 1.0))[..., None, None] * H
        return P1, P2

    def stress(self, ug, m1, m2):
        P1, P2 = self.stress_parts(ug)
        return _bc(m1, ug) * P1 + _bc(m2, ug) * P2

    def tangent(self, ug, m1, m2):
        """A[..., i,j,k,l] = dP_ij / d ug_kl."""
        vec = ug.shape[-2]
        shp = ug.shape[:-2]
        if self.kind == "poisson":
            d = ug.shape[-1]
            A = np.zeros(shp + (1, d, 1, d))
            A[..., 0, :, 0, :] = np.eye(d)
            return _bc4(m1, A) * A
        I = np.eye(3)
        if self.kind == "linear_elastic":
            dd = np.einsum("ij,kl->ijkl", I, I)
            sym = np.einsum("ik,jl->ijkl", I, I) + np.einsum("il,jk->ijkl", I, I)
            A = _bc4(m2, sym) * dd + _bc4(m1, sym) * sym
            return np.broadcast_to(A, shp + (3, 3, 3, 3)).copy()
        F = ug + I
        J = np.linalg.det(F)
        H = np.swapaxes(np.linalg.inv(F), -1, -2)
        I1 = np.sum(F * F, axis=(-2, -1))
        c = J ** (-2.0 / 3.0)
        HH = np.einsum("...ij,...kl->...ijkl", H, H)             # H_ij H_kl
        HxH = np.einsum("...il,...kj->...ijkl", H, H)            # H_il H_kj
        FH = np.einsum("...ij,...kl->...ijkl", F, H)             # F_ij H_kl
        HF = np.einsum("...ij,...kl->...ijkl", H, F)             # H_ij F_kl
        Id = np.einsum("ik,jl->ijkl", I, I)
        e = lambda a: a[..., None, None, None, None]
        dev = (-2.0 / 3.0) * FH + Id + (2.0 / 9.0) * e(I1) * HH - (2.0 / 3.0) * HF + e(I1) / 3.0 * HxH
        A_mu = e(c) * dev
        A_k = e((2.0 * J - 1.0) * J) * HH - e(J * (J - 1.0)) * HxH
        return e(np.broadcast_to(m1, shp)) * A_mu + e(np.broadcast_to(m2, shp)) * A_k

    def energy(self, ug, m1, m2):
        """Stored energy density (used only to FD-check r = dPi/du)."""
        if self.kind == "poisson":
            return 0.5 * m1 * np.sum(ug * ug, axis=(-2, -1))
        if self.kind == "linear_elastic":
            eps = 0.5 * (ug + np.swapaxes(ug, -1, -2))
            tr = np.trace(eps, axis1=-2, axis2=-1)
            return 0.5 * m2 * tr ** 2 + m1 * np.sum(eps * eps, axis=(-2, -1))
        F = ug + np.eye(3)
        J = np.linalg.det(F)
        I1 = np.sum(F * F, axis=(-2, -1))
        return 0.5 * m1 * (J ** (-2.0 / 3.0) * I1 - 3.0) + 0.5 * m2 * (J - 1.0) ** 2


def _bc(m, ug):
    m = np.asarray(m, dtype=np.float64)
    return m[..., None, None] if m.ndim else m


def _bc4(m, A):
    m = np.asarray(m, dtype=np.float64)
    return m[..., None, None, None, None] if m.ndim else m


def u_grads(cell_u, grads):
    """cell_u (c,nn,vec), grads (c,q,nn,3) -> (c,q,vec,3) (problem.py:204-206)."""
    return np.einsum("cni,cqnj->cqij", cell_u, grads)


def cell_residual(mat, cell_u, grads, JxW, p=None, chunk=65536):
    """Laplace-kernel element residuals (c, nn*vec) (problem.py:189-214)."""
    nc = cell_u.shape[0]
    out = np.empty((nc, cell_u.shape[1] * cell_u.shape[2]))
    for s in range(0, nc, chunk):
        e = slice(s, min(nc, s + chunk))
        m1, m2, _, _ = mat.moduli(None if p is None else p[e])
        ug = u_grads(cell_u[e], grads[e])
        P = mat.stress(ug, m1, m2)
        r = np.einsum("cqij,cqnj,cq->cni", P, grads[e], JxW[e])
        out[e] = r.reshape(r.shape[0], -1)
    return out


def cell_jacobian(mat, cell_u, grads, JxW, p=None, chunk=16384):
    """Element Jacobians (c, nd, nd), row-major jac[r,c] = dR_r/du_c (problem.py:262-266)."""
    nc, nn, vec = cell_u.shape
    nd = nn * vec
    out = np.empty((nc, nd, nd))
    for s in range(0, nc, chunk):
        e = slice(s, min(nc, s + chunk))
        m1, m2, _, _ = mat.moduli(None if p is None else p[e])
        ug = u_grads(cell_u[e], grads[e])
        A = mat.tangent(ug, m1, m2)                                   # (c,q,i,j,k,l)
        T = np.einsum("cqnj,cqijkl,cq->cqnikl", grads[e], A, JxW[e])
        K = np.einsum("cqnikl,cqml->cnimk", T, grads[e])
        out[e] = K.reshape(-1, nd, nd)
    return out


def cell_param_vjp(mat, cell_u, grads, JxW, p, cell_lam):
    """d/dp of lam . r for per-quad parameters: (c,q) (the material part of implicit_vjp)."""
    _, _, d1, d2 = mat.moduli(p)
    ug = u_grads(cell_u, grads)
    P1, P2 = mat.stress_parts(ug)
    dP = _bc(d1, ug) * P1 + _bc(d2, ug) * P2
    # lam . r = sum_q JxW sum_{a,i,j} lam[a,i] dP[q,i,j] dN_a/dx_j
    lg = np.einsum("cni,cqnj->cqij", cell_lam, grads)
    return np.einsum("cqij,cqij,cq->cq", lg, dP, JxW)


def energy_total(mat, cell_u, grads, JxW, p=None):
    m1, m2, _, _ = mat.moduli(p)
    ug = u_grads(cell_u, grads)
    return float(np.sum(mat.energy(ug, m1, m2) * JxW))


def surface_residual(surface_fn, x_face, face_vals, nanson):
    """Surface kernel (problem.py:238-259) for a map t(x) independent of u.

    x_face (s,fq,3), face_vals (s,fq,nn), nanson (s,fq) -> (s, nn*vec)."""
    t = np.asarray(surface_fn(x_face), dtype=np.float64)            # (s,fq,vec)
    t = np.broadcast_to(t, x_face.shape[:2] + (t.shape[-1],))
    r = np.einsum("sqi,sqn,sq->sni", t, face_vals, nanson)
    return r.reshape(r.shape[0], -1)


def mass_residual(mass_fn, x_quad, shape_vals, JxW, theta=None):
    """Mass kernel (problem.py:216-236) for a map m(x) independent of u, or
    m(x, theta) with a per-quadrature internal variable theta (c,q) (the
    reference passes cell_internal_vars to the mass map, problem.py:225-229)."""
    m = mass_fn(x_quad) if theta is None else mass_fn(x_quad, theta)
    m = np.asarray(m)
    m = np.broadcast_to(m, x_quad.shape[:2] + (m.shape[-1],))
    r = np.einsum("cqi,qn,cq->cni", m, shape_vals, JxW)
    return r.reshape(r.shape[0], -1)


def mass_theta_vjp(mass_fn, x_quad, shape_vals, JxW, theta, cell_lam):
    """d(lam . r_mass)/d theta (c,q): sum_a,i lam[c,a,i] N_a(q) JxW dm_i/dtheta,
    with dm/dtheta by the complex step (exact to rounding for a map analytic in theta)."""
    h = 1e-30
    dm = np.asarray(mass_fn(x_quad, theta + 1j * h)).imag / h
    dm = np.broadcast_to(dm, x_quad.shape[:2] + (dm.shape[-1],))
    lam_q = np.einsum("cni,qn->cqi", cell_lam, shape_vals)
    return np.einsum("cqi,cqi,cq->cq", lam_q, dm, JxW)
