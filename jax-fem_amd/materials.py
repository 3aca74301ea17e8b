"""Constitutive laws with HIP kernels behind them.

In the reference, ``Problem.get_tensor_map()`` returns a traced JAX function
(problem.py:189-214, e.g. ``jax.grad(psi)`` in
tests/benchmarks/hyperelasticity/test_hyper_elasticity.py:15-34) that jacfwd
differentiates per cell.  Here ``get_tensor_map()`` returns one of the objects
below: each is still a callable ``P = f(u_grad, *internal_vars_q)`` (NumPy,
for post-processing such as traction integrals), and carries the id of the
libhipfem kernel that evaluates it and its closed-form tangent on the GPU.

Per-quadrature parameters (``Problem.internal_vars[0]``, shape (ncell, nq);
problem.py:493-556) enter through ``param``:
  "scale_E": E = E * p               (compute_gradients notebook, rho)
  "simp":    E = Emin + (E0-Emin) p^penal   (topology optimisation)
  "mu":      mu = p, kappa fixed     (per-cell shear modulus)
  "scale_k": Poisson coefficient k * p
"""
import numpy as np

from . import _lib


class Material:
    kind = None
    vec = 3
    linear = False   # stress linear in the displacement gradient: the tangent does not depend on u

    def __init__(self, E=1.0, nu=0.3, param=None, E0=1.0, Emin=1e-3, penal=3.0, kappa=None, k=1.0):
        modes = {None: _lib.HF_PARAM_NONE, "scale_E": _lib.HF_PARAM_SCALE_E, "simp": _lib.HF_PARAM_SIMP,
                 "mu": _lib.HF_PARAM_MU, "scale_k": _lib.HF_PARAM_SCALE_K}
        if param not in modes:
            raise ValueError("unknown parameter mode %r" % (param,))
        if param == "mu" and kappa is None:
            kappa = E / (3.0 * (1.0 - 2.0 * nu))
        self.E, self.nu, self.param, self.E0, self.Emin, self.penal = E, nu, param, E0, Emin, penal
        self.kappa = 0.0 if kappa is None else kappa
        self.k = k
        self._mode = modes[param]

    def hf(self):
        return _lib.HFMaterial(self.kind, self._mode, self.E, self.nu, self.E0, self.Emin, self.penal,
                               self.kappa, self.k)

    # -- NumPy evaluation (post-processing; the solve path runs in libhipfem) --
    def _E(self, p):
        if self.param is None or p is None:
            return self.E
        if self.param == "scale_E":
            return self.E * p
        if self.param == "simp":
            return self.Emin + (self.E0 - self.Emin) * p ** self.penal
        raise ValueError(self.param)

    def moduli(self, p=None):
        nu = self.nu
        if self.param == "mu":
            return np.asarray(p, dtype=np.float64), self.kappa
        E = self._E(p)
        mu = E / (2 * (1 + nu))
        if self.kind == _lib.HF_MAT_LINEAR_ELASTIC:
            return mu, E * nu / ((1 + nu) * (1 - 2 * nu))
        return mu, E / (3 * (1 - 2 * nu))

    def __call__(self, u_grad, *iv):
        raise NotImplementedError


class LinearElastic(Material):
    """sigma = lam tr(eps) I + 2 mu eps (test_linear_elasticity_cube.py:15-25)."""
    kind = _lib.HF_MAT_LINEAR_ELASTIC
    linear = True

    def __call__(self, u_grad, *iv):
        mu, lam = self.moduli(iv[0] if iv else None)
        eps = 0.5 * (u_grad + np.swapaxes(u_grad, -1, -2))
        tr = np.trace(eps, axis1=-2, axis2=-1)[..., None, None]
        mu = np.asarray(mu)[..., None, None] if np.ndim(mu) else mu
        lam = np.asarray(lam)[..., None, None] if np.ndim(lam) else lam
        return lam * tr * np.eye(3) + 2 * mu * eps


class NeoHookean(Material):
    """psi = mu/2 (J^-2/3 I1 - 3) + kappa/2 (J-1)^2, P = dpsi/dF (test_hyper_elasticity.py:15-34)."""
    kind = _lib.HF_MAT_NEO_HOOKEAN

    def __call__(self, u_grad, *iv):
        mu, kappa = self.moduli(iv[0] if iv else None)
        F = u_grad + np.eye(3)
        J = np.linalg.det(F)[..., None, None]
        H = np.swapaxes(np.linalg.inv(F), -1, -2)
        I1 = np.sum(F * F, axis=(-2, -1))[..., None, None]
        mu = np.asarray(mu)[..., None, None] if np.ndim(mu) else mu
        kappa = np.asarray(kappa)[..., None, None] if np.ndim(kappa) else kappa
        return mu * J ** (-2.0 / 3.0) * (F - I1 / 3.0 * H) + kappa * J * (J - 1.0) * H


class Poisson(Material):
    """flux = k grad u, vec = 1 (tests/benchmarks/linear_poisson/test_linear_poisson.py:15-17)."""
    kind = _lib.HF_MAT_POISSON
    vec = 1
    linear = True

    def __call__(self, u_grad, *iv):
        k = self.k * (iv[0] if (iv and self.param == "scale_k") else 1.0)
        return k * u_grad
