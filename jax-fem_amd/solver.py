"""Newton / Krylov / adjoint drivers on one GPU (drop-in for jax_fem/solver.py).

Entry points with the reference's names and meaning:
  * ``solver(problem, solver_options)`` (solver.py:1074-1318): Newton on the
    row-eliminated system, loop while ||R||/||R0|| > rel_tol and ||R|| > tol.
  * ``linear_solver(A, b, x0, linear_options)`` (solver.py:261-284).
  * ``get_A(problem)`` (solver.py:502-515), ``apply_bc_vec`` / ``assign_bc`` /
    ``copy_bc`` (solver.py:290-363).
  * ``implicit_vjp`` / ``ad_wrapper`` (solver.py:1324-1417), the latter as a
    ``torch.autograd.Function``.

``solver_options`` keeps the reference schema (one ``newton`` method block or
a legacy flat dict, solver.py:1014-1068).  Linear keys:
  * ``jax_solver`` {precond}: the default, FP64 Jacobi-BiCGSTAB, tol = atol
    = 1e-10, maxiter 10000 (solver.py:63-92), on the device.
  * ``pcg`` / ``bicgstab`` {precond: 'jacobi', tol, atol, maxiter}: the same
    device Krylov methods with explicit settings.
  * ``learned`` {model: LearnedPC, tol, atol, maxiter, stall}: flexible CG with
    the learned CNN preconditioner on the FP32 tangent (precision 'mixed'),
    Jacobi-PCG as its fallback; the Newton residual stays the FP64 judge.
  * ``mg`` {tol, atol, maxiter, smoother, degree, post_degree, param, rebuild, fine_full,
    fine_f16, matrix_free, poll, model, learned_mode}: flexible CG preconditioned by geometric-multigrid V-cycles
    on the FP32 tangent (precision 'mixed'; box_mesh HEX8 lattices, multigrid.py), Jacobi-PCG
    as its fallback; fine_f16 (default True) smooths the fine level with an FP16 copy of
    its stencil; matrix_free (default True) computes the fine level's products from one
    cell matrix and per-cell moduli when the tangent has that form (a built linear-elastic
    law on a box lattice, Problem.cell_operator); smoother 'learned' with model=LearnedPC makes the learned
    network the fine level's post-smoother;
    rebuild 'solve' (default: the hierarchy of a Newton solve's first tangent
    preconditions its later steps), 'always' or 'never'.
  * ``custom_solver``: callable(A, b, x0, linear_options) -> x (solver.py:278-280);
    A is the device ``TangentMatrix``.
  * ``spsolve_solver`` (a host direct solve in the reference, solver.py:93-106)
    is a dense LU on the device up to 32,768 DOFs and, above, FP64 Jacobi-BiCGSTAB
    to a relative residual of 1e-13 (the accuracy a direct solve delivers in FP64);
  * ``petsc_solver`` {ksp_type, pc_type} (solver.py:109-138): with pc_type 'lu'
    the same dense LU up to 32,768 DOFs; otherwise FP64 Jacobi-BiCGSTAB on the
    device to PETSc's default relative tolerance 1e-5, with the reference's
    ||Ax - b|| < 0.1 check; ``amgx_solver`` (NVIDIA AmgX)
    raises ``NotImplementedError``.
Added Newton keys:
  * ``precision``: 'fp64' (default: everything FP64, as the reference) or
    'mixed' (FP32 element Jacobian, FP32 SpMV/Krylov with FP64 dot products;
    the Newton residual and update stay FP64, i.e. FP64 iterative refinement).
  * ``forcing``: 'fixed' (default for fp64: the linear tolerance as given) or
    'ew' (Eisenstat-Walker inexact Newton; default for mixed).  With 'ew' the
    first Newton step of a solve uses ``eta_first`` (default 1e-3): it carries
    the Dirichlet increment of a load step, and a loose solve would leave it
    in a thin boundary layer whose tangent can be indefinite.  A step that
    raises ||R|| is halved up to 8 times (``backtrack``, default on with 'ew'),
    and a PCG breakdown (p.Ap <= 0) re-solves that step with BiCGSTAB.
  * ``lag_ratio`` (default 10 with 'ew', 0 otherwise): when ||R||/||R0|| is
    within this factor of ``rel_tol`` and the last step formed a fresh tangent,
    the next step reuses that tangent (a chord step): what is left is a small
    reduction that the lagged tangent gives as well, without a new element
    Jacobian, assembly and preconditioner set-up.  Convergence is still judged
    on the FP64 residual, and a chord step is never followed by another.  With
    ``lag_ratio`` > 0 and a nonlinear law the last tangent of a solve is kept on
    the problem: the first step of the next solve of the same problem reuses it
    when that solve starts from the last solution (a load step warm-started from
    the previous one: the tangent is the one at, or one small correction from,
    that state) and would form a tangent of the same kind; a LatticeTangent is
    reused only while no newer stencil was assembled.  It holds the tangent's
    memory until the next solve.
Unknown keys are ignored, as in the reference.
"""
import ctypes
import logging
import time

import numpy as np
import torch

from . import _lib

logger = logging.getLogger("jax_fem_amd")

_METHOD_KEYS = ("newton", "arc_length", "dynamic_relax")
_LINEAR_KEYS = ("jax_solver", "amgx_solver", "spsolve_solver", "petsc_solver", "custom_solver", "pcg", "bicgstab",
                "learned", "mg")
_NEWTON_KEYS = ("tol", "rel_tol", "line_search_flag", "initial_guess", "precision", "forcing", "max_iter",
                "eta_max", "eta_min", "eta_first", "backtrack", "lag_ratio")


def _resolve_solver_options(solver_options):
    """(method, cfg); legacy flat dicts become Newton (solver.py:1050-1068)."""
    opts = solver_options or {}
    methods = [m for m in _METHOD_KEYS if m in opts]
    if not methods:
        cfg = {k: opts[k] for k in _NEWTON_KEYS if k in opts}
        lin = {k: opts[k] for k in _LINEAR_KEYS if k in opts}
        if lin:
            cfg["linear"] = lin
        return "newton", cfg
    if len(methods) > 1:
        raise ValueError("Pick one nonlinear method, got %s." % methods)
    m = methods[0]
    if not isinstance(opts[m], dict):
        raise ValueError("solver_options['%s'] must be a dict." % m)
    return m, opts[m]


# ---------------------------------------------------------------------------
class TangentMatrix:
    """Row-eliminated tangent on the device: node-pair block values over the problem's pattern.

    The object ``get_A`` returns (the reference returns a PETSc Mat)."""

    def __init__(self, problem, vals):
        self.problem = problem
        self.pattern = problem.pattern
        self.vals = vals
        self.dtype = vals.dtype
        self.shape = (problem.N, problem.N)
        self._dinv = None

    @property
    def tag(self):
        return "64" if self.dtype == torch.float64 else "32"

    def matvec(self, x, transpose=False):
        x = x.to(self.dtype).contiguous()
        y = torch.empty_like(x)
        name = "hf_spmv_t_f" if transpose else "hf_spmv_f"
        _lib.call(name + self.tag, self.pattern.handle, _lib.ptr(self.vals), _lib.ptr(x), _lib.ptr(y), _lib.stream())
        return y

    def __matmul__(self, x):
        return self.matvec(x)

    def dinv(self):
        if self._dinv is None:
            d = torch.empty(self.problem.N, dtype=self.dtype, device=self.vals.device)
            _lib.call("hf_jacobi_f" + self.tag, self.pattern.handle, _lib.ptr(self.vals), _lib.ptr(d), _lib.stream())
            self._dinv = d
        return self._dinv

    def diagonal(self):
        return 1.0 / self.dinv()

    def csr(self):
        """(rowptr, colind, values) of the scalar CSR view (FP64), device tensors."""
        rp, ci = self.pattern.csr()
        v = torch.empty(self.pattern.nnz, dtype=torch.float64, device=self.vals.device)
        vals = self.vals.to(torch.float64).contiguous()
        _lib.call("hf_matrix_to_csr_f64", self.pattern.handle, _lib.ptr(vals), _lib.ptr(v), _lib.stream())
        return rp, ci, v

    def to_scipy(self):
        import scipy.sparse as sp
        rp, ci, v = self.csr()
        return sp.csr_matrix((v.cpu().numpy(), ci.cpu().numpy(), rp.cpu().numpy()), shape=self.shape)

    def getValuesCSR(self):
        rp, ci, v = self.csr()
        return rp.cpu().numpy(), ci.cpu().numpy(), v.cpu().numpy()

    def getSize(self):
        return self.shape


def get_A(problem, dtype=torch.float64):
    """Assembled, row-eliminated tangent at the last ``newton_update`` state (solver.py:502-515)."""
    u = getattr(problem, "_newton_u", None)
    if u is None:
        raise RuntimeError("call problem.newton_update(sol_list) first")
    cache = getattr(problem, "_A_cache", None)
    if cache is not None and cache.dtype == dtype:
        return cache
    A = TangentMatrix(problem, problem.jacobian_bsr(u, dtype))
    problem._A_cache = A
    return A


# ---------------------------------------------------------------------------
def apply_bc_vec(res_vec, dofs, problem, scale=1.0):
    """R[B] = u[B] - scale*g (solver.py:290-304); returns a new vector."""
    d, v = problem.bc_arrays()
    R = res_vec.to(torch.float64).clone().contiguous()
    dofs = dofs.to(torch.float64).contiguous()
    _lib.call("hf_bc_residual_f64", _lib.ptr(d), _lib.ptr(v), d.numel(), float(scale), _lib.ptr(dofs), _lib.ptr(R),
              _lib.stream())
    return R


def assign_bc(dofs, problem):
    d, v = problem.bc_arrays()
    out = dofs.to(torch.float64).clone()
    out[d.long()] = v
    return out


def copy_bc(dofs, problem):
    d, _ = problem.bc_arrays()
    out = torch.zeros_like(dofs, dtype=torch.float64)
    out[d.long()] = dofs.to(torch.float64)[d.long()]
    return out


def _initial_guess(problem, u):
    """x0 = assign_bc(0) - copy_bc(u) (solver.py:402-409)."""
    d, v = problem.bc_arrays()
    x0 = torch.empty(problem.N, dtype=torch.float64, device=u.device)
    _lib.call("hf_bc_initial_guess_f64", _lib.ptr(d), _lib.ptr(v), d.numel(), _lib.ptr(u), _lib.ptr(x0),
              problem.N, _lib.stream())
    return x0


class KrylovStats:
    def __init__(self):
        self.calls = []
        self.token = self    # identifies the solve sequence for a preconditioner's rebuild policy

    def add(self, method, iters, status, res, bnorm, seconds):
        self.calls.append({"method": method, "iterations": iters, "status": status, "res": res,
                           "b_norm": bnorm, "seconds": seconds})

    @property
    def iterations(self):
        return sum(c["iterations"] for c in self.calls)


def krylov(A, b, x0=None, method="pcg", rtol=1e-10, atol=1e-10, maxiter=10000, precond=True,
           transpose=False, poll=8, stats=None, dinv=None, rel_r0=False):
    """Device Krylov solve on a TangentMatrix; returns (x, info) with x in A's dtype.

    Stops at ||r||^2 <= max(rtol^2 ref, atol^2) with ref = ||b||^2 (the reference's
    jax_solve rule) or, with ``rel_r0``, ref = ||b - A x0||^2 (inexact Newton)."""
    dt = A.dtype
    N = A.shape[0]
    b = b.to(dt).contiguous()
    x = (torch.zeros(N, dtype=dt, device=b.device) if x0 is None else x0.to(dt).clone()).contiguous()
    if dinv is None:
        dinv = A.dinv() if precond else torch.ones(N, dtype=dt, device=b.device)
    mcode = _lib.HF_KRYLOV_PCG if method == "pcg" else _lib.HF_KRYLOV_BICGSTAB
    esz = 8 if dt == torch.float64 else 4
    wsz = _lib.load().hf_krylov_workspace_size(mcode, N, esz)
    ws = _workspace(b.device, wsz)
    info = _lib.HFKrylovInfo()
    t0 = time.perf_counter()
    _lib.call("hf_krylov_solve_f" + A.tag, mcode, A.pattern.handle, _lib.ptr(A.vals), _lib.ptr(dinv), _lib.ptr(b),
              _lib.ptr(x), (1 if transpose else 0) | (2 if rel_r0 else 0), int(maxiter), float(rtol), float(atol),
              int(poll),
              _lib.ptr(ws), ctypes.byref(info), _lib.stream())
    if stats is not None:
        stats.add(method, info.iterations, info.status, info.res_norm, info.b_norm, time.perf_counter() - t0)
    if info.status == 3:
        raise FloatingPointError("Krylov solve produced non-finite values")
    return x, info


_WS = {}


def _workspace(device, nbytes):
    w = _WS.get(device)
    if w is None or w.numel() < nbytes:
        w = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
        _WS[device] = w
    return w


def linear_solver(A, b, x0, linear_options, stats=None, rtol=None):
    """Dispatch on linear_options (solver.py:261-284).

    ``rtol`` (set by inexact Newton) overrides the configured tolerance and makes
    it relative to the initial linear residual ||b - A x0||."""
    opts = linear_options if linear_options is not None else {}
    if not any(k in opts for k in _LINEAR_KEYS):
        opts = dict(opts)
        opts["jax_solver"] = {}
    if "jax_solver" in opts:
        precond = opts["jax_solver"].get("precond", True)
        x, info = krylov(A, b, x0, "bicgstab", 1e-10 if rtol is None else rtol, 1e-10, 10000, precond,
                         stats=stats, rel_r0=rtol is not None)
        if rtol is None:
            err = torch.linalg.norm(A.matvec(x).to(torch.float64) - b.to(torch.float64)).item()
            assert err < 0.1, "linear solver failed to converge with err = %g" % err
        return x
    for key in ("pcg", "bicgstab"):
        if key in opts:
            o = opts[key] or {}
            pc = o.get("precond", "jacobi")
            if pc not in ("jacobi", True, None, False, "none"):
                raise NotImplementedError("preconditioner %r" % (pc,))
            x, info = krylov(A, b, x0, key, o.get("tol", 1e-10) if rtol is None else rtol,
                             o.get("atol", 1e-10) if rtol is None else 0.0,
                             o.get("maxiter", 10000), pc not in (False, "none"), stats=stats,
                             poll=o.get("poll", 8), rel_r0=rtol is not None)
            return x
    if "learned" in opts:
        o = opts["learned"] or {}
        model = o["model"]
        x, info, fb = model.solve(A, b, x0, A.problem, o.get("tol", 1e-6) if rtol is None else rtol,
                                  o.get("atol", 0.0) if rtol is not None else o.get("atol", 1e-10),
                                  o.get("maxiter", 10000), rel_r0=rtol is not None, poll=o.get("poll", 8),
                                  stall=o.get("stall", 60), stats=stats)
        if stats is not None:
            stats.calls[-1]["fallback"] = fb
        return x
    if "mg" in opts:
        from .multigrid import solve_with_fallback
        o = opts["mg"] or {}
        if A.dtype != torch.float32:
            raise ValueError("the multigrid preconditioner runs on the FP32 tangent: use precision='mixed'")
        x, info, fb = solve_with_fallback(A, b, x0, A.problem, o.get("tol", 1e-8) if rtol is None else rtol,
                                          o.get("atol", 1e-10) if rtol is None else 0.0, o.get("maxiter", 1000),
                                          rel_r0=rtol is not None, poll=o.get("poll", 1),
                                          smoother=o.get("smoother", "chebyshev"), degree=o.get("degree", 2),
                                          param=o.get("param"), rebuild=o.get("rebuild", "solve"),
                                          fine_full=o.get("fine_full", False), stats=stats, model=o.get("model"),
                                          fine_f16=o.get("fine_f16", True), post_degree=o.get("post_degree"),
                                          learned_mode=o.get("learned_mode", "replace"))
        if stats is not None:
            stats.calls[-1]["fallback"] = fb
        return x
    if "custom_solver" in opts:
        return opts["custom_solver"](A, b, x0, opts)
    if _direct(opts) and A.shape[0] <= _DENSE_MAX and A.dtype == torch.float64:
        return dense_solve(A, b)
    for key, tol in (("spsolve_solver", 1e-13), ("petsc_solver", 1e-5)):
        if key in opts:
            x, info = krylov(A, b, x0, "bicgstab", tol if rtol is None else rtol, 0.0, 50000, True, stats=stats,
                             rel_r0=rtol is not None)
            if info.status != 0:
                logger.warning("%s: device BiCGSTAB ended with status %d at relative residual %.3e", key,
                               info.status, info.res_norm / max(info.b_norm, 1e-300))
            if key == "petsc_solver" and rtol is None:
                err = torch.linalg.norm(A.matvec(x).to(torch.float64) - b.to(torch.float64)).item()
                assert err < 0.1, "PETSc linear solver failed to converge, err = %g" % err
            return x
    if "amgx_solver" in opts:
        raise NotImplementedError("amgx_solver is NVIDIA AmgX; use jax_solver, pcg, bicgstab or mg")
    raise NotImplementedError("Unknown linear solver.")


_DENSE_MAX = 32768     # direct solves: a dense LU on the device up to this many DOFs (8 GiB)


def _direct(opts):
    """A direct solve was asked for: spsolve_solver, or petsc_solver with an LU
    preconditioner (one application of which solves the system)."""
    if "spsolve_solver" in opts:
        return True
    o = opts.get("petsc_solver")
    return o is not None and (o or {}).get("pc_type") == "lu"


def dense_solve(A, b, transpose=False):
    """x = A^-1 b (or A^-T b) by LU on the device (rocSOLVER through torch), A formed
    densely from its CSR view: the direct solve of the reference's spsolve_solver
    (solver.py:93-106) for systems that fit (N <= 32768: <= 8 GiB)."""
    rp, ci, v = A.csr()
    n = A.shape[0]
    rows = torch.repeat_interleave(torch.arange(n, device=v.device), (rp[1:] - rp[:-1]).long())
    D = torch.zeros(n, n, dtype=torch.float64, device=v.device)
    D[rows, ci.long()] = v.to(torch.float64)
    if transpose:
        D = D.T.contiguous()
    return torch.linalg.solve(D, b.to(torch.float64).reshape(-1, 1)).reshape(-1)


# ---------------------------------------------------------------------------
def _flat_initial_guess(problem, cfg):
    if "initial_guess" in cfg and cfg["initial_guess"] is not None:
        g = cfg["initial_guess"]
        if isinstance(g, (list, tuple)):
            g = g if problem.num_vars > 1 else g[0]
        return problem.flatten_sol_list([t.detach() if torch.is_tensor(t) else t for t in g]
                                        if isinstance(g, (list, tuple)) else
                                        (g.detach() if torch.is_tensor(g) else g)).clone()
    return torch.zeros(problem.N, dtype=torch.float64, device=problem.device)


def _stencil_multigrid(problem, lin_opts, adt):
    """The Multigrid whose lattice-stencil form the Newton tangent takes, or None:
    the 'mg' linear method on the FP32 tangent of a HEX8 vec-3 box lattice with the
    half-stored fine level ('stencil': False keeps the SELL tangent)."""
    if "mg" not in lin_opts or adt != torch.float32:
        return None
    o = lin_opts["mg"] or {}
    if not o.get("stencil", True) or o.get("fine_full", False):
        return None
    if problem.fe.ele_type != "HEX8" or problem.vec != 3:
        return None
    from .multigrid import Multigrid, lattice
    try:
        lattice(problem)
    except NotImplementedError:
        return None
    return Multigrid.for_problem(problem, o.get("smoother", "chebyshev"), o.get("degree", 2), o.get("param"), False,
                                 o.get("model"), o.get("learned_mode", "replace"), o.get("fine_f16", True),
                                 o.get("post_degree"))


def solver(problem, solver_options={}):
    """Newton solve (solver.py:1074-1318); returns sol_list = [u (nnode, vec) f64 device]."""
    method, cfg = _resolve_solver_options(solver_options)
    if method != "newton":
        raise NotImplementedError("%s is not part of the GPU solve path" % method)
    precision = cfg.get("precision", "fp64")
    if precision not in ("fp64", "mixed"):
        raise ValueError("precision must be 'fp64' or 'mixed'")
    forcing = cfg.get("forcing", "ew" if precision == "mixed" else "fixed")
    rel_tol = cfg.get("rel_tol", 1e-8)
    tol = cfg.get("tol", 1e-6)
    max_iter = cfg.get("max_iter", 100)
    lin_opts = dict(cfg.get("linear", {}))
    if precision == "mixed" and not any(k in lin_opts for k in _LINEAR_KEYS):
        lin_opts["pcg"] = {}
    for key in ("learned", "mg"):
        if key in lin_opts and precision != "mixed":
            raise ValueError("the %s preconditioner runs on the FP32 tangent: use precision='mixed'" % key)
    adt = torch.float32 if precision == "mixed" else torch.float64
    problem.hold_params(True)
    try:
        return _newton(problem, cfg, precision, forcing, rel_tol, tol, max_iter, lin_opts, adt)
    finally:
        problem.hold_params(False)


def _newton(problem, cfg, precision, forcing, rel_tol, tol, max_iter, lin_opts, adt):
    eta_max = cfg.get("eta_max", 0.1)
    eta_min = cfg.get("eta_min", 1e-6 if precision == "mixed" else 1e-10)
    eta_first = cfg.get("eta_first", 1e-3)
    backtrack = cfg.get("backtrack", forcing == "ew")
    lag_ratio = float(cfg.get("lag_ratio", 10.0 if forcing == "ew" else 0.0))
    timing = {"local_assembly": 0.0, "global_matrix": 0.0, "linear": 0.0}
    stats = KrylovStats()
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    u = _flat_initial_guess(problem, cfg)

    def residual(u):
        t0 = time.perf_counter()
        R = apply_bc_vec(problem.residual_vec(u), u, problem)
        res = float(torch.linalg.norm(R).item())
        timing["local_assembly"] += time.perf_counter() - t0
        return R, res

    asm_events, lin_events = [], []
    smg = _stencil_multigrid(problem, lin_opts, adt)
    if smg is not None:
        from .multigrid import LatticeTangent
    R, res = residual(u)
    res0 = res
    hist = [res]
    n = 0
    eta_prev = None
    A = None
    linear_law = bool(getattr(problem.material, "linear", False))
    fresh = False     # the last step formed its own tangent
    chords = 0
    lag_key = lag_refs = None
    if lag_ratio > 0.0 and not linear_law:
        from .problem import _fingerprint
        # the internal variables by content (arrays) or by storage and version (tensors); the
        # kept tangent holds references to them, so no new tensor can take a kept one's storage
        lag_refs = (getattr(problem, "internal_vars", None), getattr(problem, "internal_vars_surfaces", None))
        lag_key = (smg, adt, (lin_opts.get("mg") or {}).get("matrix_free", True) if smg is not None else None,
                   _fingerprint(lag_refs[0]), _fingerprint(lag_refs[1]))
    lag = getattr(problem, "_lag_tangent", None)
    problem._lag_tangent = None
    if lag is not None:
        lagA, key, lag_u = lag[:3]
        if (lag_ratio <= 0.0 or linear_law or key != lag_key or lag_u.shape != u.shape
                or (isinstance(lagA, LatticeTangent if smg is not None else ()) and
                    getattr(problem, "_stencil_version", None) != lagA._version)
                or not (_same_tensors(cfg.get("initial_guess"), lag[5]) or torch.equal(lag_u, u))):
            lag = None
    def rel_of(r):
        # ||R||/||R0||; a zero initial residual gives NaN as the reference's NumPy
        # division does, so the loop below does not start (solver.py:1283-1295)
        return r / res0 if res0 > 0.0 else float("nan")

    while (rel_of(res) > rel_tol) and (res > tol) and n < max_iter:
        n += 1
        # the tangent's time is read from events after the solve: no stall between
        # the assembly and the linear solve the host enqueues behind it
        ev0 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        if A is not None and linear_law:
            pass   # a linear law's tangent does not depend on u: the first step's serves the solve
        elif A is None and lag is not None:
            # chord step on the last solve's tangent; a multigrid keeps the hierarchy
            # it had for that solve (its own staleness check still rebuilds it)
            A, stats.token, lag = lag[0], lag[3], None
            fresh = False
            chords += 1
        elif A is not None and fresh and rel_of(res) <= lag_ratio * rel_tol:
            fresh = False   # chord step on the last tangent
            chords += 1
        elif smg is not None:
            A = LatticeTangent(problem, u, smg, matrix_free=(lin_opts["mg"] or {}).get("matrix_free", True))
            fresh = True
        else:
            A = TangentMatrix(problem, problem.jacobian_bsr(u, adt))
            fresh = True
        ev1 = torch.cuda.Event(enable_timing=True)
        ev1.record()
        asm_events.append((ev0, ev1))
        x0 = _initial_guess(problem, u)
        if forcing == "ew":
            if eta_prev is None:
                eta = eta_first
            else:
                eta = 0.9 * (hist[-1] / hist[-2]) ** 2
                if 0.9 * eta_prev ** 2 > 0.1:
                    eta = max(eta, 0.9 * eta_prev ** 2)
            eta = max(eta, 0.5 * rel_tol * res0 / res)     # do not oversolve past the target
            eta = min(max(eta, eta_min), eta_max)
            eta_prev = eta
        else:
            eta = None
        # the solve's time comes from events as well: the update and the next
        # residual queue behind the solve's last copies without a stall
        ev2 = torch.cuda.Event(enable_timing=True)
        ev2.record()
        inc = linear_solver(A, -R, x0, lin_opts, stats=stats, rtol=eta)
        if stats.calls and stats.calls[-1]["status"] == 2 and "pcg" in lin_opts:
            # p.Ap <= 0: the tangent is not positive definite here; BiCGSTAB does not need it
            logger.debug("newton %d: PCG breakdown, re-solving with BiCGSTAB", n + 1)
            o = dict(lin_opts["pcg"] or {})
            inc = linear_solver(A, -R, x0, {"bicgstab": o}, stats=stats, rtol=eta)
        ev3 = torch.cuda.Event(enable_timing=True)
        ev3.record()
        lin_events.append((ev2, ev3))
        inc = inc.to(torch.float64)
        if cfg.get("line_search_flag", False):
            u, R, res = _line_search(problem, u, inc, residual)
        else:
            u_new = u + inc
            R_new, res_new = residual(u_new)
            if backtrack:
                # keep the Dirichlet part of the step exact: only the interior part is scaled
                d, _ = problem.bc_arrays()
                alpha = 1.0
                for _ in range(8):
                    if np.isfinite(res_new) and res_new <= res:
                        break
                    alpha *= 0.5
                    step = inc * alpha
                    step[d.long()] = inc[d.long()]
                    u_new = u + step
                    R_new, res_new = residual(u_new)
                if alpha < 1.0:
                    logger.debug("newton %d: backtracked to alpha %.3g", n + 1, alpha)
            u, R, res = u_new, R_new, res_new
        hist.append(res)
        last = stats.calls[-1] if stats.calls else {}
        logger.debug("newton %d: res %.3e rel %.3e | eta %s krylov %s status %s", n, res, rel_of(res),
                     "%.2e" % eta if eta is not None else "-", last.get("iterations"), last.get("status"))
    if not (np.isfinite(res) and bool(torch.isfinite(u).all())):
        raise FloatingPointError("Newton produced non-finite values")
    torch.cuda.synchronize()
    timing["global_matrix"] = sum(a.elapsed_time(b) for a, b in asm_events) * 1e-3
    timing["linear"] = sum(a.elapsed_time(b) for a, b in lin_events) * 1e-3
    problem.solve_info = {"newton_iters": n, "res_hist": hist, "res0": res0, "rel": res / res0 if res0 > 0.0 else 0.0,
                          "krylov": stats.calls, "krylov_iters": stats.iterations,
                          "fallbacks": sum(1 for c in stats.calls if c.get("fallback")), "chord_steps": chords,
                          "timing": timing, "wall": time.perf_counter() - wall0, "precision": precision}
    problem._newton_u = u
    problem._A_cache = None
    sol = problem.unflatten_fn_sol_list(u)
    if A is not None and lag_ratio > 0.0 and not linear_law:
        # the returned tensors (views of u, which the kept tuple holds) by storage, version
        # and shape: a next solve started from them is known to start from u without a sync
        problem._lag_tangent = (A, lag_key, u, stats.token, lag_refs, _tensor_sig(sol))
    elif lag is not None:
        problem._lag_tangent = lag[:2] + (u,) + lag[3:5] + (_tensor_sig(sol),)   # no step taken: still valid
    return sol


def _tensor_sig(ts):
    return tuple((t.data_ptr(), t._version, tuple(t.shape)) for t in ts)


def _same_tensors(g, sig):
    """Whether the initial guess g is the tensors a solve returned (signature sig),
    unmodified since."""
    if g is None:
        return False
    g = list(g) if isinstance(g, (list, tuple)) else [g]
    return len(g) == len(sig) and all(torch.is_tensor(t) and (t.data_ptr(), t._version, tuple(t.shape)) == q
                                      for t, q in zip(g, sig))


def _line_search(problem, u, inc, residual):
    """Up to 3 step halvings while the residual norm keeps decreasing (solver.py:424-460)."""
    alpha = 1.0
    R, res = residual(u + alpha * inc)
    for _ in range(3):
        R_half, res_half = residual(u + 0.5 * alpha * inc)
        if res_half > res:
            break
        alpha *= 0.5
        R, res = R_half, res_half
    return u + alpha * inc, R, res


# ---------------------------------------------------------------------------
def _flatten(params):
    if isinstance(params, (list, tuple)):
        out, specs = [], []
        for p in params:
            f, s = _flatten(p)
            out += f
            specs.append(s)
        return out, (type(params), specs)
    return [params], None


def _unflatten(flat, spec):
    if spec is None:
        return flat.pop(0)
    typ, specs = spec
    return typ(_unflatten(flat, s) for s in specs)


def adjoint_solve(problem, u, v, adjoint_solver_options=None):
    """lam with A^T lam = v at u (solver.py:1358-1371); returns (lam, lam_interior).

    Options: 'bicgstab' / 'jax_solver' (default: FP64 Jacobi-BiCGSTAB on A^T, as the
    reference), 'pcg' (FP64 Jacobi-PCG on the symmetric A_II), 'mg' {tol, inner_tol,
    max_refine, maxiter} (FP64 refinement around the FP32 multigrid flexible CG on
    A_II; box_mesh HEX8 lattices)."""
    opts = adjoint_solver_options or {}
    A = TangentMatrix(problem, problem.jacobian_bsr(u, torch.float64))
    v = v.to(torch.float64).contiguous()
    d, _ = problem.bc_arrays()
    dl = d.long()
    if "pcg" in opts:
        # A_II is symmetric for these laws: solve A_II lam_I = v_I, then
        # lam_B = v_B - (A_IB^T lam_I), the transposed product of the kept columns.
        o = opts["pcg"] or {}
        vI = v.clone()
        vI[dl] = 0.0
        lamI, _ = krylov(A, vI, None, "pcg", o.get("tol", 1e-10), o.get("atol", 1e-10), o.get("maxiter", 10000))
        t = A.matvec(lamI, transpose=True)
        lam = lamI.clone()
        lam[dl] = v[dl] - t[dl]
    elif "mg" in opts:
        # A_II symmetric: A_II lam_I = v_I by FP64 iterative refinement around the FP32
        # multigrid flexible CG (each correction solved to `inner_tol` relative to its
        # right-hand side); stop when the FP64 residual ||v_I - A_II lam_I|| <= tol ||v_I||.
        from .multigrid import solve_with_fallback
        o = opts["mg"] or {}
        A32 = TangentMatrix(problem, A.vals.to(torch.float32))
        vI = v.clone()
        vI[dl] = 0.0
        nv = float(torch.linalg.norm(vI))
        lamI = torch.zeros_like(vI)
        r = vI.clone()
        stats = KrylovStats()
        for _ in range(o.get("max_refine", 20)):
            if float(torch.linalg.norm(r)) <= o.get("tol", 1e-10) * nv:
                break
            c, _, _ = solve_with_fallback(A32, r, None, problem, o.get("inner_tol", 1e-4), 0.0, o.get("maxiter", 1000),
                                          rel_r0=True, stats=stats)
            lamI += c.to(torch.float64)
            lamI[dl] = 0.0
            r = vI - A.matvec(lamI)
            r[dl] = 0.0
        problem.adjoint_info = {"refinements": len(stats.calls), "krylov": [c["iterations"] for c in stats.calls],
                                "rel": float(torch.linalg.norm(r)) / max(nv, 1e-300)}
        t = A.matvec(lamI, transpose=True)
        lam = lamI.clone()
        lam[dl] = v[dl] - t[dl]
    else:
        o = opts.get("bicgstab", opts.get("jax_solver", {})) or {}
        tol, atol, maxiter = o.get("tol", 1e-10), o.get("atol", 1e-10), o.get("maxiter", 10000)
        if _direct(opts) and A.shape[0] <= _DENSE_MAX:     # a direct solve, as linear_solver
            lam = dense_solve(A, v, transpose=True)
            lamI = lam.clone()
            lamI[dl] = 0.0
            return lam, lamI
        if "spsolve_solver" in opts:         # direct-solve accuracy (see linear_solver)
            tol, atol, maxiter = 1e-13, 0.0, 50000
        elif "petsc_solver" in opts:
            tol, atol, maxiter = 1e-5, 0.0, 50000
        lam, _ = krylov(A, v, None, "bicgstab", tol, atol, maxiter, o.get("precond", True) not in (False, "none"),
                        transpose=True)
    lamI = lam.clone()
    lamI[dl] = 0.0
    return lam, lamI


def param_vjp(problem, u, lamI):
    """-(d R / d p)^T lam per quadrature point, (ncell, nq) f64 (the material part of implicit_vjp)."""
    pts, cells = problem._dev()
    prm = problem._params()
    if prm is None:
        return None
    grad = torch.empty_like(prm)
    mat = problem._hfmat()
    u, lamI = u.contiguous(), lamI.contiguous()
    _lib.call("hf_cell_param_vjp_f64", problem.nn, ctypes.byref(mat), _lib.ptr(pts), _lib.ptr(cells),
              problem.num_cells, _lib.ptr(u), _lib.ptr(prm), _lib.ptr(lamI), _lib.ptr(grad), _lib.stream())
    return grad


def param_vjp_pointwise(problem, u, lamI, which):
    """-(dR/dq)^T lam for the internal variables q = internal_vars[k], k in `which`,
    of maps evaluated pointwise (generic.PointwiseTerms): with lam's gradient and
    values at the quadrature points (the same contraction kernels as u's) as the
    cotangents, torch.func.vjp of the tensor map and the mass map with respect to
    their per-point internal-variable arguments, and of the universal kernel with
    respect to its per-cell ones.  Returns a list of tensors shaped like
    cell_internal_vars()[k] (a per-cell value repeated over the points receives the
    sum over its points)."""
    terms = problem._pointwise()
    plan = problem._plan()
    raw = problem.cell_internal_vars()
    total = [torch.zeros_like(a) for a in raw]

    from . import numpy as jnp

    if terms.tensor_fn is not None or terms.mass_fn is not None:
        ivs = problem.volume_internal_vars()
        Q = problem.num_cells * problem.fe.num_quads
        flat = [a.reshape((Q,) + tuple(a.shape[2:])) for a in ivs]
        need_q = plan["mass"] == "cell"
        ug, uq = terms.point_data(u.contiguous(), need_q)
        lg, lq = terms.point_data(lamI.contiguous(), need_q)
        w = terms.jxw.reshape(-1)

        def add(fn, fixed, cot):
            nf = len(fixed)

            def one(*a):
                fx, c, rest = a[:nf], a[nf], a[nf + 1:]
                _, pull = torch.func.vjp(lambda *q: fn(*fx, *q), *rest)
                return pull(c)
            with jnp.on_device(u.device):
                g = torch.func.vmap(one)(*fixed, cot, *flat)
            for k in which:
                gk = g[k].reshape(ivs[k].shape)
                if raw[k].dim() == 1:
                    gk = gk.sum(dim=1)
                total[k] += gk
        if terms.tensor_fn is not None:
            add(terms.tensor_fn, [ug], lg * w[:, None, None])
        if terms.mass_fn is not None:
            add(terms.mass_fn, [uq, terms.xq.reshape(Q, -1)], lq * w[:, None])
    if terms.universal is not None:
        cu = terms._cell_sol_flat(u)
        cl = terms._cell_sol_flat(lamI)
        sgc, jxw_all, vgj = terms._all_vars_geometry()
        args = [cu, terms.xq, sgc, jxw_all, vgj]
        k5 = len(args)
        kern = terms.universal

        def one_cell(*a):
            fx, c, rest = a[:k5], a[k5], a[k5 + 1:]
            _, pull = torch.func.vjp(lambda *q: kern(*fx, *q), *rest)
            return pull(c)
        with jnp.on_device(u.device):
            g = torch.func.vmap(one_cell)(*args, cl, *raw)
        for k in which:
            total[k] += g[k].reshape(total[k].shape)
    return [-t for t in total]


def implicit_vjp(problem, sol_list, params, v_list, adjoint_solver_options={}):
    """Gradient of <v, u*(params)> w.r.t. params (solver.py:1324-1380).

    Differentiates through ``problem.set_params`` with torch autograd: the
    per-quadrature field ``internal_vars[0]`` receives -(dR/dp)^T lam, and
    Dirichlet values given as torch tensors (``fe.vals_list``) receive lam_B."""
    flat, spec = _flatten(params)
    leaves = [torch.as_tensor(p, dtype=torch.float64).detach().clone().requires_grad_(True)
              if (torch.is_tensor(p) or np.ndim(p) >= 0) else p for p in flat]
    with torch.enable_grad():
        problem.set_params(_unflatten(list(leaves), spec))
        ivs_req = list(problem.internal_vars or ())
        bc_vals = [sp[4] for sp in problem.bc_specs()]
    u = problem.flatten_sol_list(sol_list)
    v = problem.flatten_sol_list(v_list)
    problem._newton_u = u
    lam, lamI = adjoint_solve(problem, u, v, adjoint_solver_options)
    outs, gouts = [], []
    which = [k for k, a in enumerate(ivs_req) if torch.is_tensor(a) and a.requires_grad]
    if which and problem._plan()["pure_fast"]:
        iv = ivs_req[0]
        g = param_vjp(problem, u, lamI)
        outs.append(iv)
        gouts.append(g.reshape(iv.shape).to(iv.device, iv.dtype))
    elif which:
        if problem._plan()["law"] == "fast":
            raise NotImplementedError("parameter gradients of a built law combined with pointwise terms")
        gs = param_vjp_pointwise(problem, u, lamI, which)
        for k in which:
            gk = gs[k]
            outs.append(ivs_req[k])
            gouts.append(gk.reshape(ivs_req[k].shape).to(ivs_req[k].device, ivs_req[k].dtype))
    # Dirichlet values: c_B = u_B - g  =>  d<lam, c>/dg = -lam_B, grad = +lam_B (later specs win)
    specs = problem.bc_specs()
    owner = {}
    for s, (o, vec, nodes, comps, _) in enumerate(specs):
        for j, dof in enumerate((o + np.asarray(nodes) * vec + np.asarray(comps)).tolist()):
            owner[dof] = (s, j)
    lam_h = lam.detach().cpu().numpy()
    for s, vals in enumerate(bc_vals):
        if torch.is_tensor(vals) and vals.requires_grad:
            o, vec, nodes, comps, _ = specs[s]
            dofs = (o + np.asarray(nodes) * vec + np.asarray(comps)).tolist()
            gv = np.array([lam_h[d] if owner[d] == (s, j) else 0.0 for j, d in enumerate(dofs)])
            outs.append(vals)
            gouts.append(torch.as_tensor(gv, dtype=vals.dtype, device=vals.device))
    req = [l for l in leaves if torch.is_tensor(l) and l.requires_grad]
    if outs:
        grads = torch.autograd.grad(outs, req, grad_outputs=gouts, allow_unused=True)
    else:
        grads = [None] * len(req)
    grads = [torch.zeros_like(l) if g is None else g for g, l in zip(grads, req)]
    problem.set_params(_unflatten([l.detach() for l in leaves], spec))
    return _unflatten(list(grads), spec)


def ad_wrapper(problem, solver_options={}, adjoint_solver_options={}):
    """fwd_pred(params) -> sol_list, differentiable with torch autograd (solver.py:1383-1417)."""

    class _ImplicitSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, spec_holder, *flat):
            params = _unflatten([f.detach() for f in flat], spec_holder[0])
            problem.set_params(params)
            sol = problem.flatten_sol_list(solver(problem, solver_options))   # every variable's DOFs
            ctx.spec = spec_holder[0]
            ctx.save_for_backward(*flat)
            ctx.sol = sol.detach()
            return sol.clone()

        @staticmethod
        def backward(ctx, v):
            flat = ctx.saved_tensors
            params = _unflatten(list(flat), ctx.spec)
            g = implicit_vjp(problem, [ctx.sol], params, [v], adjoint_solver_options)
            gflat, _ = _flatten(g)
            return (None,) + tuple(gflat)

    def fwd_pred(params):
        flat, spec = _flatten(params)
        flat = [f if torch.is_tensor(f) else torch.as_tensor(f, dtype=torch.float64) for f in flat]
        return problem.unflatten_fn_sol_list(_ImplicitSolve.apply([spec], *flat))

    return fwd_pred
