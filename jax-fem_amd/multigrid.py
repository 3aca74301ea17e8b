"""Geometric multigrid preconditioner on the structured node lattice (GPU, libhipfem).

No reference counterpart: the reference's Krylov solves use Jacobi only
(jax_fem/solver.py:68-69).  The hierarchy, the V-cycle and the flexible CG are
csrc/mg.hip; this module owns the device object, builds the Dirichlet mask,
re-runs the set-up when the tangent changes and implements the fallback.

    level 0      the FP32 tangent with Dirichlet DOFs masked out (M A M), as a 27-point
                 stencil (HEX8) or the SELL matrix itself (HEX27, whose nodes couple two
                 lattice steps apart; its level 1 is the HEX8 operator of the same cells)
    level l+1    P^T A_l P, P trilinear interpolation from the lattice with
                 every dimension's interval count halved (rounded up)
    coarsest     <= 88 DOF, dense FP64 inverse
    storage      FP32; the fine-level stencil of a symmetric operator half-stored
                 (diagonal + 13 upper offsets, the lower blocks read as the transposes
                 of the neighbours'; fine_full=True keeps all 27 slots); with
                 fine_f16 the smoothing passes of the half-stored levels read FP16
                 copies (scaled by a power of two; the outer CG keeps FP32)
    cycle        V-cycle, `degree` smoothing steps per side: Chebyshev in Dinv A on
                 [lam/ratio, lam], lam the Gershgorin bound of Dinv A (Dinv: inverse
                 3x3 diagonal blocks), or damped block Jacobi (omega)
    outer        flexible CG (one kept direction), FP64 dot products

Used through ``solver_options={'newton': {'precision': 'mixed', 'linear':
{'mg': {...}}}}`` (solver.linear_solver) or directly:

    mg = Multigrid.for_problem(problem)
    x, info = mg.solve(A, b, x0, rtol=1e-8)
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib


def lattice(problem):
    """Node counts (nx, ny, nz) of a structured box_mesh (HEX8) / box_mesh_hex27 problem with vec = 3."""
    g = getattr(problem.fe.mesh, "node_grid", None)
    if g is None or int(np.prod(g)) != problem.fe.num_total_nodes or problem.vec != 3 or \
            problem.fe.ele_type not in ("HEX8", "HEX27"):
        raise NotImplementedError("the multigrid preconditioner needs a structured box_mesh HEX8 / HEX27 "
                                  "lattice, vec=3")
    return tuple(int(x) for x in g)


def dirichlet_mask(problem):
    """1 on free DOFs, 0 on Dirichlet DOFs (float32 device, N); rebuilt when the
    problem's Dirichlet DOF set (the tensor bc_arrays returns) changes."""
    d, _ = problem.bc_arrays()
    hit = getattr(problem, "_dirichlet_mask", None)
    if hit is not None and hit[0] is d:
        return hit[1]
    m = torch.ones(problem.N, dtype=torch.float32, device=problem.device)
    m[d.long()] = 0.0
    problem._dirichlet_mask = (d, m)
    return m


class LatticeTangent:
    """FP32 tangent of a HEX8 box-lattice problem in the multigrid's stencil form
    (Problem.jacobian_stencil): the operator is the row-eliminated tangent, as the
    TangentMatrix the reference's get_A / apply_bc give (solver.py:466-515,
    290-363), applied from each node's diagonal and upper node blocks.  It is valid
    until the problem assembles its next stencil; to_sell() assembles the SELL
    TangentMatrix at the same state for the paths that need one (the Jacobi-PCG
    fallback)."""

    def __init__(self, problem, u, mg, matrix_free=True):
        self.problem = problem
        self.pattern = problem.pattern
        self.mg = mg
        self.u = u
        self.dtype = torch.float32
        self.shape = (problem.N, problem.N)
        # the operator as sum_c scale[c] K, when it has that form (Problem.cell_operator):
        # the multigrid's fine-level products then run matrix-free, and the stencil its
        # set-up reads comes from the cell matrix instead of the element Jacobians
        self.cell_op = problem.cell_operator() if matrix_free else None
        if self.cell_op is not None:
            mg.set_cell_operator(self.cell_op)
            self.stencil, self._version = problem.cell_stencil(mg)
        else:
            self.stencil, self._version = problem.jacobian_stencil(u, mg)
        self._sell = None

    @property
    def ident(self):
        """Identity of this tangent for the hierarchy's rebuild policy (the stencil buffer is reused)."""
        return self

    def check_current(self):
        if getattr(self.problem, "_stencil_version", None) != self._version:
            raise RuntimeError("LatticeTangent: the problem has assembled a newer stencil over this one")

    def to_sell(self):
        if self._sell is None:
            from .solver import TangentMatrix
            self._sell = TangentMatrix(self.problem, self.problem.jacobian_bsr(self.u, torch.float32))
        return self._sell

    def matvec(self, x):
        """A x with the stencil (needs a set-up of the hierarchy for the Dirichlet mask)."""
        self.check_current()
        x = x.to(torch.float32).contiguous()
        y = torch.empty_like(x)
        _lib.call("hf_mg_stencil_matvec_f32", self.mg.handle, _lib.ptr(self.stencil), _lib.ptr(x), _lib.ptr(y),
                  _lib.stream())
        return y

    def __matmul__(self, x):
        return self.matvec(x)


class Multigrid:
    def __init__(self, grid, smoother="chebyshev", degree=2, param=None, device=None, fine_full=False,
                 fine_sell=False, fine_f16=False):
        """fine_sell: the fine level is the tangent's SELL matrix (HEX27 lattices);
        fine_full: store all 27 fine stencil slots (a non-symmetric operator);
        fine_f16: the smoothing passes of the half-stored levels read FP16 copies of
        their stencils (HF_MG_FINE_F16; with fine_sell the coarser levels only)."""
        _lib.load()
        self.grid = tuple(int(v) for v in grid)
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.fine_full = bool(fine_full)
        self.fine_sell = bool(fine_sell)
        self.fine_f16 = bool(fine_f16) and not self.fine_full
        h = ctypes.c_void_p()
        flags = (_lib.HF_MG_FINE_FULL if self.fine_full else 0) | (_lib.HF_MG_FINE_SELL if self.fine_sell else 0) | \
            (_lib.HF_MG_FINE_F16 if self.fine_f16 else 0)
        _lib.call("hf_mg_create", self.grid[0], self.grid[1], self.grid[2], flags, ctypes.byref(h))
        self.handle = h
        self.smoother_cfg = None
        self.set_smoother(smoother, degree, param)
        self._tangent = None     # the matrix values and mask the hierarchy was built from (kept alive)
        self._mask = None
        self._pattern = None
        self.setup_seconds = 0.0
        self.setups = 0
        self.last = None
        self._token = None       # the solve sequence (Newton solve) the hierarchy was built in
        self._iters_at_setup = None
        self._stale = False

    def __del__(self):
        h = getattr(self, "handle", None)
        lib = getattr(_lib, "_lib", None) if _lib is not None else None
        if h is not None and lib is not None:
            lib.hf_mg_destroy(h)
            self.handle = None

    @classmethod
    def for_problem(cls, problem, smoother="chebyshev", degree=2, param=None, fine_full=False, model=None,
                    learned_mode="replace", fine_f16=True, post_degree=None):
        """The problem's cached hierarchy (one per lattice) with the given smoother.
        smoother 'learned': Chebyshev everywhere, and on the fine level the learned
        network `model` (a LearnedPC) as the post-smoother (learned_mode 'replace') or
        one network step after the Chebyshev post-smoothing ('after')."""
        grid = lattice(problem)
        sell = problem.fe.ele_type == "HEX27"
        poly = "chebyshev" if smoother == "learned" else smoother
        mg = getattr(problem, "_mg", None)
        f16 = bool(fine_f16) and not fine_full
        if mg is None or mg.grid != grid or mg.fine_full != bool(fine_full) or mg.fine_sell != sell or \
                mg.fine_f16 != f16:
            mg = cls(grid, poly, degree, param, device=problem.device, fine_full=fine_full, fine_sell=sell,
                     fine_f16=f16)
            problem._mg = mg
        mg.set_smoother(poly, degree, param, post_degree)
        if smoother == "learned":
            if model is None:
                raise ValueError("the 'learned' smoother needs linear option 'model' (a LearnedPC)")
            mg.set_learned(model, problem, learned_mode)
        else:
            mg.set_learned(None)
        return mg

    def set_learned(self, model, problem=None, mode="replace"):
        """The fine level's post-smoothing by the learned network (hf_mg_set_learned):
        one step e += M(r - A e), M(v) = D^-1 v + s D^-1/2 mask N([D^-1/2 v / s, g]),
        in place of the Chebyshev post-smoothing ('replace') or after it ('after');
        model None switches it off.  The network's input field g is the problem's
        log-modulus at the nodes (LearnedPC.gfield), refreshed when its parameters change."""
        if model is None:
            if getattr(self, "_learned", None) is not None:
                _lib.call("hf_mg_set_learned", self.handle, 0, None, None, _lib.stream())
                self._learned = None
            return
        from .learned_pc import LearnedPC
        modes = {"replace": 1, "after": 2}
        if mode not in modes:
            raise ValueError("learned_mode must be 'replace' or 'after'")
        g = LearnedPC.gfield_device(problem)
        key = (id(model), modes[mode], g.data_ptr(), getattr(model, "version", 0))
        if getattr(self, "_learned", None) is not None and self._learned[0] == key:
            return
        _lib.call("hf_mg_set_learned", self.handle, modes[mode], model.weight_ptrs(), _lib.ptr(g), _lib.stream())
        self._learned = (key, model, g)     # keeps the weights and the field alive

    def set_smoother(self, smoother="chebyshev", degree=2, param=None, post_degree=None):
        """smoother 'chebyshev' (param: interval ratio, default 8) or 'jacobi' (param: omega,
        default 0.5); degree: pre-smoothing steps, post_degree: post-smoothing steps
        (1..degree, default degree; the first steps of the same recurrence)."""
        kinds = {"jacobi": _lib.HF_MG_JACOBI, "chebyshev": _lib.HF_MG_CHEBYSHEV}
        if smoother not in kinds:
            raise ValueError("smoother must be 'chebyshev' or 'jacobi'")
        if param is None:
            param = 8.0 if smoother == "chebyshev" else 0.5
        post = 0 if post_degree is None or int(post_degree) == int(degree) else int(post_degree)
        cfg = (smoother, int(degree), float(param), post)
        if cfg != self.smoother_cfg:
            _lib.call("hf_mg_set_smoother", self.handle, kinds[smoother], int(degree), ctypes.c_float(param))
            _lib.call("hf_mg_set_post_degree", self.handle, post)
            self.smoother_cfg = cfg

    def bounds(self):
        """Gershgorin bound of Dinv A per smoothed level (last set-up)."""
        lam = (ctypes.c_float * 16)()
        _lib.call("hf_mg_bounds", self.handle, lam, 16)
        return [lam[i] for i in range(len(self.levels()) - 1)]

    def levels(self):
        """[(nx, ny, nz), ...] finest first."""
        n = ctypes.c_int32()
        dims = (ctypes.c_int32 * 48)()
        _lib.call("hf_mg_levels", self.handle, ctypes.byref(n), dims, 16)
        return [tuple(dims[3 * i:3 * i + 3]) for i in range(n.value)]

    def stencil_size(self):
        """Floats of one level-0 stencil (LatticeTangent)."""
        return int(_lib.load().hf_mg_stencil_size(self.handle))

    def setup(self, A, mask):
        """Build every level from the FP32 tangent A (a TangentMatrix, or a
        LatticeTangent of this hierarchy's lattice) and the Dirichlet mask."""
        if A.dtype != torch.float32:
            raise ValueError("the multigrid preconditioner runs on the FP32 tangent")
        # one mask buffer per hierarchy: the cycle reads it by pointer, and a captured
        # solve batch stays valid across set-ups while the pointer does not change
        m = mask.to(device=self.device, dtype=torch.float32).reshape(-1)
        if self._mask is None or self._mask.numel() != m.numel():
            self._mask = torch.empty_like(m)
        self._mask.copy_(m)
        t0 = time.perf_counter()
        # no dropped-pivot count requested: after a pattern's first set-up the call
        # only enqueues (the solve's launches queue behind it)
        if isinstance(A, LatticeTangent):
            A.check_current()
            _lib.call("hf_mg_setup_stencil_f32", self.handle, A.pattern.handle, _lib.ptr(A.stencil),
                      _lib.ptr(self._mask), None, _lib.stream())
        else:
            _lib.call("hf_mg_setup_f32", self.handle, A.pattern.handle, _lib.ptr(A.vals), _lib.ptr(self._mask),
                      None, _lib.stream())
        self.setup_seconds += time.perf_counter() - t0
        self._tangent = _ident(A)   # identity of the tangent the levels come from
        self._pattern = A.pattern
        self.setups += 1
        self._iters_at_setup = None
        self._stale = False

    def is_current(self, A):
        return self._tangent is _ident(A)

    def needs_setup(self, A, rebuild, token):
        """rebuild: 'always' (every tangent), 'solve' (once per Newton solve, identified
        by token) or 'never' (the first tangent's hierarchy for every later one).  A
        hierarchy is also rebuilt for a new pattern, or when a solve with it took more
        than max(2 n0, n0 + 10) iterations, n0 the count of the first solve after its set-up."""
        if self._tangent is None or self._pattern is not A.pattern or self._stale:
            return True
        if rebuild == "always":
            return not self.is_current(A)
        if rebuild == "solve":
            return token is not self._token
        if rebuild == "never":
            return False
        raise ValueError("rebuild must be 'always', 'solve' or 'never'")

    def note_iterations(self, n):
        if self._iters_at_setup is None:
            self._iters_at_setup = n
        elif n > max(2 * self._iters_at_setup, self._iters_at_setup + 10):
            self._stale = True

    def set_cell_operator(self, op):
        """op = (K, scale) from Problem.cell_operator: the fine level's smoothing passes
        and the flexible CG's products run matrix-free (hf_mg_set_cell_operator) with
        it; None returns them to the stencil.  The scale tensor is kept alive here."""
        cur = getattr(self, "_cell_op", None)
        if op is None:
            if cur is not None:
                _lib.call("hf_mg_set_cell_operator", self.handle, None, None)
                self._cell_op = None
                self._stale = True    # a set-up with the cell operator made no FP16 copy of the stencil
            return
        K, scale = op
        if cur is not None and cur[1] is scale and cur[0] is K:
            return
        k = np.ascontiguousarray(K, dtype=np.float32)
        _lib.call("hf_mg_set_cell_operator", self.handle, k.ctypes.data_as(ctypes.c_void_p), _lib.ptr(scale))
        self._cell_op = (K, scale, k)

    def cell_matvec(self, x):
        """y = A x with the cell operator set by set_cell_operator (the row-eliminated
        operator, as LatticeTangent.matvec)."""
        x = x.to(torch.float32).contiguous()
        y = torch.empty_like(x)
        _lib.call("hf_mg_stencil_matvec_f32", self.handle, None, _lib.ptr(x), _lib.ptr(y), _lib.stream())
        return y

    def apply(self, r, skip_fine_post=False):
        """z = V(r) (one cycle); skip_fine_post stops after the fine level's coarse correction."""
        r = r.to(torch.float32).contiguous()
        z = torch.empty_like(r)
        _lib.call("hf_mg_apply_f32", self.handle, _lib.ptr(r), _lib.ptr(z),
                  _lib.HF_MG_SKIP_FINE_POST if skip_fine_post else 0, _lib.stream())
        return z

    def stencil(self, level):
        """Level `level`'s stencil as a (27, 3, 3, nnode) float32 tensor: [o, i, j, n] couples
        DOF i of node n to DOF j of its neighbour at offset o = (dx+1)*9 + (dy+1)*3 + (dz+1)."""
        nx, ny, nz = self.levels()[level]
        n = nx * ny * nz
        out = torch.empty(243 * n, dtype=torch.float32, device=self.device)
        _lib.call("hf_mg_copy_stencil", self.handle, int(level), _lib.ptr(out), _lib.stream())
        return out.reshape(27, 3, 3, n)

    def solve(self, A, b, x0=None, rtol=1e-8, atol=0.0, maxiter=500, rel_r0=False, poll=1, stats=None,
              mask=None, problem=None, rebuild="always", token=None, eager=False, split_scalars=False):
        """Flexible CG preconditioned by V-cycles; the operator is always the FP32
        tangent A, the hierarchy is rebuilt from A as `rebuild` says (needs_setup).
        Each batch of `poll` iterations replays a captured HIP graph (eager=True
        launches every kernel instead).  split_scalars: the direction scalars by
        their own one-workgroup kernel at every size (the form large grids take).
        Returns (x float32, info)."""
        self.set_cell_operator(A.cell_op if isinstance(A, LatticeTangent) else None)
        if self.needs_setup(A, rebuild, token):
            self.setup(A, dirichlet_mask(problem or A.problem) if mask is None else mask)
            self._token = token
        N = A.shape[0]
        b = b.to(torch.float32).contiguous()
        x = (torch.zeros(N, dtype=torch.float32, device=b.device) if x0 is None else
             x0.to(torch.float32).clone()).contiguous()
        ws = _workspace(b.device, int(_lib.load().hf_mg_workspace_size(self.handle)))
        info = _lib.HFKrylovInfo()
        t0 = time.perf_counter()
        flags = (2 if rel_r0 else 0) | (_lib.HF_MG_SOLVE_EAGER if eager else 0) | \
            (_lib.HF_MG_SOLVE_SPLIT_SCALARS if split_scalars else 0)
        if isinstance(A, LatticeTangent):
            A.check_current()
            _lib.call("hf_mg_solve_stencil_f32", self.handle, _lib.ptr(A.stencil), _lib.ptr(b), _lib.ptr(x), flags,
                      int(maxiter), float(rtol), float(atol), int(poll), _lib.ptr(ws), ctypes.byref(info),
                      _lib.stream())
        else:
            _lib.call("hf_mg_solve_f32", self.handle, A.pattern.handle, _lib.ptr(A.vals), _lib.ptr(b), _lib.ptr(x),
                      flags, int(maxiter), float(rtol), float(atol), int(poll), _lib.ptr(ws), ctypes.byref(info),
                      _lib.stream())
        self.last = (int(info.iterations), int(info.status))
        self.note_iterations(int(info.iterations))
        if info.status != 0:
            self._stale = True
        if stats is not None:
            stats.add("fcg-mg", info.iterations, info.status, info.res_norm, info.b_norm, time.perf_counter() - t0)
        return x, info


def _ident(A):
    """What identifies a tangent for the rebuild policy: the values tensor of a
    TangentMatrix, the object itself for a LatticeTangent (its buffer is reused)."""
    return A.ident if isinstance(A, LatticeTangent) else A.vals


_WS = {}


def _workspace(device, nbytes):
    w = _WS.get(device)
    if w is None or w.numel() < nbytes:
        w = torch.empty(int(nbytes), dtype=torch.uint8, device=device)
        _WS[device] = w
    return w


def solve_with_fallback(A, b, x0, problem, rtol, atol, maxiter, rel_r0, poll=1, smoother="chebyshev", degree=2,
                        param=None, rebuild="solve", fine_full=False, stats=None, model=None, learned_mode="replace",
                        fine_f16=True, post_degree=None):
    """Multigrid FCG; on breakdown or non-finite values Jacobi-PCG solves to the same
    absolute target from the better of the FCG iterate and x0.  `stats` (the Newton
    solve's KrylovStats) identifies the solve for rebuild='solve'.  Returns (x, info, fallback)."""
    from .solver import krylov
    mg = Multigrid.for_problem(problem, smoother, degree, param, fine_full, model, learned_mode, fine_f16,
                               post_degree)
    x_start = (torch.zeros(A.shape[0], dtype=torch.float32, device=b.device) if x0 is None else
               x0.to(torch.float32).clone())
    x, info = mg.solve(A, b, x_start, rtol, atol, maxiter, rel_r0, poll, stats, problem=problem, rebuild=rebuild,
                       token=getattr(stats, "token", stats) if stats is not None else object())
    if info.status in (0, 1):
        return x, info, False
    if isinstance(A, LatticeTangent):
        A = A.to_sell()
    b32 = b.to(torch.float32)
    r0 = torch.linalg.norm((b32 - A.matvec(x_start)).double()).item()
    target = max(rtol * (r0 if rel_r0 else info.b_norm), atol)
    start = x_start
    if info.status != 3 and bool(torch.isfinite(x).all()):
        if torch.linalg.norm((b32 - A.matvec(x)).double()).item() < r0:
            start = x
    x2, info2 = krylov(A, b32, start, "pcg", 0.0, target, max(maxiter, 10000), True, stats=stats)
    info2.iterations += info.iterations
    return x2, info2, True
