"""Constitutive laws with HIP kernels behind them.

In the reference, ``Problem.get_tensor_map()`` returns a traced JAX function
(problem.py:189-214, e.g. ``jax.grad(psi)`` in
tests/benchmarks/hyperelasticity/test_hyper_elasticity.py:15-34) that jacfwd
differentiates per cell.  Here ``get_tensor_map()`` returns one of the objects
below: each is still a callable ``P = f(u_grad, *internal_vars_q)`` (NumPy,
for post-processing such as traction integrals), and carries the id of the
libhipfem kernel that evaluates it and its closed-form tangent on the GPU.  The
call is written with jax_fem_amd.numpy, so a law also runs on the pointwise
path (torch.func) where the built kernels do not apply (other element types
or quadrature rules).

Per-quadrature parameters (``Problem.internal_vars[0]``, shape (ncell, nq);
problem.py:493-556) enter through ``param``:
  "scale_E": E = E * p               (compute_gradients notebook, rho)
  "simp":    E = Emin + (E0-Emin) p^penal   (topology optimisation)
  "mu":      mu = p, kappa fixed     (per-cell shear modulus)
  "scale_k": Poisson coefficient k * p
"""
import numpy as np

from . import _lib
from . import numpy as jnp


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
            return (p if hasattr(p, "detach") else np.asarray(p, dtype=np.float64)), self.kappa
        E = self._E(p)
        mu = E / (2 * (1 + nu))
        if self.kind == _lib.HF_MAT_LINEAR_ELASTIC:
            return mu, E * nu / ((1 + nu) * (1 - 2 * nu))
        return mu, E / (3 * (1 - 2 * nu))

    def __call__(self, u_grad, *iv):
        raise NotImplementedError


def _bcast(m):
    """A modulus (scalar or per point) broadcast against (..., 3, 3)."""
    if hasattr(m, "detach"):
        return m[..., None, None] if m.dim() else m
    return np.asarray(m)[..., None, None] if np.ndim(m) else m


class LinearElastic(Material):
    """sigma = lam tr(eps) I + 2 mu eps (test_linear_elasticity_cube.py:15-25)."""
    kind = _lib.HF_MAT_LINEAR_ELASTIC
    linear = True

    def __call__(self, u_grad, *iv):
        mu, lam = self.moduli(iv[0] if iv else None)
        eps = 0.5 * (u_grad + jnp.swapaxes(u_grad, -1, -2))
        tr = jnp.trace(eps, axis1=-2, axis2=-1)[..., None, None]
        return _bcast(lam) * tr * jnp.eye(3) + 2 * _bcast(mu) * eps


class NeoHookean(Material):
    """psi = mu/2 (J^-2/3 I1 - 3) + kappa/2 (J-1)^2, P = dpsi/dF (test_hyper_elasticity.py:15-34)."""
    kind = _lib.HF_MAT_NEO_HOOKEAN

    def __call__(self, u_grad, *iv):
        mu, kappa = self.moduli(iv[0] if iv else None)
        F = u_grad + jnp.eye(3)
        J = jnp.linalg.det(F)[..., None, None]
        H = jnp.swapaxes(jnp.linalg.inv(F), -1, -2)
        I1 = jnp.sum(F * F, axis=(-2, -1))[..., None, None]
        return _bcast(mu) * J ** (-2.0 / 3.0) * (F - I1 / 3.0 * H) + _bcast(kappa) * J * (J - 1.0) * H


class Poisson(Material):
    """flux = k grad u, vec = 1 (tests/benchmarks/linear_poisson/test_linear_poisson.py:15-17)."""
    kind = _lib.HF_MAT_POISSON
    vec = 1
    linear = True

    def __call__(self, u_grad, *iv):
        k = self.k * (iv[0] if (iv and self.param == "scale_k") else 1.0)
        return k * u_grad


def _probe_eval(fn, g, ivs):
    out = np.asarray(fn(g, *ivs), dtype=np.float64)
    if not np.all(np.isfinite(out)):
        raise FloatingPointError
    return out


def _close(a, b, tol=1e-10):
    return np.allclose(a, b, rtol=tol, atol=tol * max(1e-300, float(np.abs(b).max())))


def identify(fn, vec, dim, scalar_param=False):
    """The built law equal to the pointwise map ``fn(u_grad, *internal_vars_q)``, or None.

    The map is evaluated with NumPy arrays at a few points: moduli are fitted on a
    uniaxial strain and the fit is accepted only if the map equals the law to 1e-10
    (relative) at random displacement gradients.  With ``scalar_param`` the map takes
    one scalar per point and must be linear in it (P(g, p) = p P(g, 1): the modulus
    scaled per point, param "scale_E", or "scale_k" for Poisson)."""
    if dim != 3:
        return None
    rng = np.random.default_rng(12345)
    samples = [0.05 * rng.standard_normal((vec, dim)) for _ in range(3)]
    one = (1.0,) if scalar_param else ()
    try:
        if vec == 1:
            g = np.array([[0.3, -0.2, 0.5]])
            k = float(_probe_eval(fn, g, one).reshape(-1)[0] / g[0, 0])
            cand = Poisson(k=k, param="scale_k" if scalar_param else None)
            law = lambda gg: k * gg
        elif vec == 3:
            cand, law = None, None
            eps = 1e-3
            g = np.zeros((3, 3))
            g[0, 0] = eps
            P = _probe_eval(fn, g, one)
            lam = P[1, 1] / eps
            mu = (P[0, 0] - lam * eps) / (2 * eps)
            if mu > 0 and lam + mu > 0:
                E = mu * (3 * lam + 2 * mu) / (lam + mu)
                nu = lam / (2 * (lam + mu))
                lin = LinearElastic(E=E, nu=nu)
                if all(_close(_probe_eval(fn, s, one), lin(s)) for s in samples):
                    cand, law = LinearElastic(E=E, nu=nu, param="scale_E" if scalar_param else None), lin
            if cand is None:
                e = 0.1
                F = np.diag([1 + e, 1.0, 1.0])
                J, I1 = 1 + e, (1 + e) ** 2 + 2
                Finv_T = np.diag([1 / (1 + e), 1.0, 1.0])
                a = J ** (-2 / 3) * (F - I1 / 3 * Finv_T)
                b = J * (J - 1) * Finv_T
                P = _probe_eval(fn, F - np.eye(3), one)
                M = np.array([[a[0, 0], b[0, 0]], [a[1, 1], b[1, 1]]])
                mu, kappa = np.linalg.solve(M, [P[0, 0], P[1, 1]])
                if mu > 0 and kappa > 0:
                    E = 9 * kappa * mu / (3 * kappa + mu)
                    nu = (3 * kappa - 2 * mu) / (2 * (3 * kappa + mu))
                    nh = NeoHookean(E=E, nu=nu)
                    if all(_close(_probe_eval(fn, s, one), nh(s)) for s in samples):
                        cand, law = NeoHookean(E=E, nu=nu, param="scale_E" if scalar_param else None), nh
            if cand is None:
                return None
        else:
            return None
        if scalar_param:
            for p in (0.7, 1.9):
                for s in samples:
                    if not _close(_probe_eval(fn, s, (p,)), p * law(s)):
                        return None
        return cand
    except Exception:
        return None
