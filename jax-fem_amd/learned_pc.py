"""Learned preconditioner: a 3x3x3 CNN on the structured node grid (GPU, libhipfem).

No reference counterpart (the reference's Krylov solves use Jacobi only,
jax_fem/solver.py:68-69).  The network and the solve are defined in
csrc/lpc.hip; this module holds the weights, builds the input fields, runs
inference / training through the C ABI and implements the fallback:

    M(r) = D^-1 r + s D^-1/2 mask N([D^-1/2 r / s, g]),   s = rms(D^-1/2 r)

N: conv 4->16 (tanh), 16->16 (tanh), 16->3, all 3x3x3 with zero padding; g is
the log Young's modulus averaged to the nodes and divided by 3 (0 for a
homogeneous body).  With the last layer at zero M is Jacobi.

Training minimises, over residual samples r with r^T D^-1 r = 1, the energy
    L(z) = 1/2 z^T A z - z^T r,   z = M(r),
whose minimiser is z = A^-1 r, so dL/dz = A z - r (one SpMV).  The residuals
are the ones Jacobi-PCG and the model's own flexible CG produce on the
training matrices (iterations 0..K).  Forward and backward passes are the FP32
HIP kernels (hf_cnn_conv3d_fwd/bwd_data/bwd_weight, hf_lpc_features /
output); only the Adam update of the ~10k weights is a torch operation.  The
solve applies the network on MFMA with f16 activations (hf_lpc_apply_f32 is
that path for one residual).

The shipped weights (weights/lpc_elastic.npz) come from
    python jax-fem_amd/train_lpc.py --instances 2000 --epochs 1 --domains 1.0,0.25
(32^3 heterogeneous elasticity, seeds 1..2000, fields at two grid spacings,
2000 Adam steps, 73 s on one MI355X).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib

LAYERS = ((4, 16, 1), (16, 16, 1), (16, 3, 0))   # (cin, cout, tanh)


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


class LearnedPC:
    trunc = 2   # directions kept by the flexible CG

    def __init__(self, seed=0, device=None, zero_last=True):
        self.device = device or _dev()
        g = torch.Generator().manual_seed(seed)
        self.W, self.B = [], []
        for li, (ci, co, _) in enumerate(LAYERS):
            std = math.sqrt(1.0 / (27 * ci))
            w = torch.randn(27, ci, co, generator=g) * std
            if zero_last and li == len(LAYERS) - 1:
                w.zero_()
            self.W.append(w.float().to(self.device).contiguous())
            self.B.append(torch.zeros(co, dtype=torch.float32, device=self.device))
        self._ptrs = None

    # -- parameters ------------------------------------------------------------
    def params(self):
        out = []
        for w, b in zip(self.W, self.B):
            out += [w, b]
        return out

    def weight_ptrs(self):
        arr = (ctypes.c_void_p * 6)(*[p.data_ptr() for p in self.params()])
        self._ptrs = arr
        return arr

    def save(self, path):
        np.savez(path, **{"p%d" % i: p.detach().cpu().numpy() for i, p in enumerate(self.params())})

    @classmethod
    def load(cls, path, device=None):
        m = cls(device=device)
        with np.load(path, allow_pickle=False) as d:
            for i, p in enumerate(m.params()):
                p.copy_(torch.as_tensor(d["p%d" % i]))
        return m

    def poisoned(self, mode):
        """A copy with broken weights (tests of the fallback): 'nan' or 'huge'."""
        m = LearnedPC(device=self.device)
        for a, b in zip(m.params(), self.params()):
            a.copy_(b)
        if mode == "nan":
            m.W[-1].fill_(float("nan"))
        else:
            m.W[-1].normal_(0, 1e6)
        return m

    # -- problem fields ----------------------------------------------------------
    @staticmethod
    def grid(problem):
        """The structured node lattice (nx, ny, nz) the network runs on: a box lattice of
        HEX8 cells, or of HEX27 cells (nodes two per cell edge plus one), vec = 3."""
        g = getattr(problem.fe.mesh, "node_grid", None)
        if g is None or int(np.prod(g)) != problem.fe.num_total_nodes or problem.vec != 3 or \
                problem.fe.ele_type not in ("HEX8", "HEX27"):
            raise NotImplementedError("the learned preconditioner needs a structured box_mesh HEX8 or HEX27 node "
                                      "lattice, vec=3")
        return tuple(int(x) for x in g)

    @staticmethod
    def gfield(problem):
        """log E averaged from cells to nodes, / 3 (float32 device, nnode)."""
        mat = problem.material
        iv = problem.internal_vars
        nnode = problem.fe.num_total_nodes
        if not iv or mat.param is None:
            return torch.zeros(nnode, dtype=torch.float32, device=problem.device)
        p = iv[0].detach().cpu().numpy() if torch.is_tensor(iv[0]) else np.asarray(iv[0])
        pc = p.reshape(problem.num_cells, -1).mean(axis=1)
        E = mat._E(pc) if mat.param != "mu" else pc
        logE = np.log(np.maximum(E / max(mat.E if mat.param == "scale_E" else 1.0, 1e-30), 1e-30))
        cells = problem.fe.cells
        acc = np.zeros(nnode)
        cnt = np.zeros(nnode)
        np.add.at(acc, cells.reshape(-1), np.repeat(logE, cells.shape[1]))
        np.add.at(cnt, cells.reshape(-1), 1.0)
        return torch.as_tensor(acc / np.maximum(cnt, 1) / 3.0, dtype=torch.float32, device=problem.device)

    @staticmethod
    def gfield_device(problem):
        """gfield computed on the device from the problem's parameter tensor (the same
        field: log E averaged from cells to nodes, / 3), cached per parameter tensor."""
        mat = problem.material
        nnode = problem.fe.num_total_nodes
        prm = problem._params()
        key = (id(prm), prm._version if prm is not None else None)
        hit = getattr(problem, "_gfield_dev", None)
        if hit is not None and hit[0] == key and hit[1] is prm:
            return hit[2]
        if prm is None or mat.param is None:
            g = torch.zeros(nnode, dtype=torch.float32, device=problem.device)
        else:
            pc = prm.mean(dim=1)
            if mat.param == "scale_E":
                E = mat.E * pc / max(mat.E, 1e-30)
            elif mat.param == "simp":
                E = mat.Emin + (mat.E0 - mat.Emin) * pc ** mat.penal
            else:
                E = pc
            logE = torch.log(torch.clamp(E, min=1e-30))
            _, cells = problem._dev()
            idx = cells.long().reshape(-1)
            acc = torch.zeros(nnode, dtype=torch.float64, device=problem.device)
            cnt = torch.zeros(nnode, dtype=torch.float64, device=problem.device)
            acc.index_add_(0, idx, logE.repeat_interleave(cells.shape[1]))
            cnt.index_add_(0, idx, torch.ones_like(idx, dtype=torch.float64))
            g = (acc / torch.clamp(cnt, min=1.0) / 3.0).float().contiguous()
        problem._gfield_dev = (key, prm, g)
        return g

    @staticmethod
    def mask(problem):
        m = torch.ones(problem.N, dtype=torch.float32, device=problem.device)
        d, _ = problem.bc_arrays()
        m[d.long()] = 0.0
        return m

    # -- inference / training passes ---------------------------------------------
    def forward(self, A, R, gfield, mask, grid, keep=False):
        """z = M(r) for a batch R (B, N) float32; returns z and the cached activations."""
        nx, ny, nz = grid
        B = R.shape[0]
        nnode = nx * ny * nz
        dinv = A.dinv()
        R = R.contiguous()
        s = torch.sqrt(torch.clamp((R * R * dinv).sum(1) / R.shape[1], min=1e-30)).float().contiguous()
        st = _lib.stream()
        f = torch.empty(B, nnode, 4, dtype=torch.float32, device=R.device)
        _lib.call("hf_lpc_features_f32", _lib.ptr(R), _lib.ptr(dinv), _lib.ptr(gfield), nnode, B, _lib.ptr(s),
                  _lib.ptr(f), st)
        acts = [f]
        x = f
        for (ci, co, act), w, b in zip(LAYERS, self.W, self.B):
            y = torch.empty(B, nnode, co, dtype=torch.float32, device=R.device)
            _lib.call("hf_cnn_conv3d_fwd_f32", _lib.ptr(x), ci, _lib.ptr(w), _lib.ptr(b), co, act, nx, ny, nz, B,
                      _lib.ptr(y), st)
            acts.append(y)
            x = y
        z = torch.empty_like(R)
        _lib.call("hf_lpc_output_f32", _lib.ptr(R), _lib.ptr(dinv), _lib.ptr(mask), _lib.ptr(x), nnode, B,
                  _lib.ptr(s), _lib.ptr(z), st)
        cache = {"acts": acts, "s": s, "dinv": dinv, "mask": mask, "grid": grid, "B": B} if keep else None
        return z, cache

    def backward(self, cache, gz):
        """Gradients of <gz, z> w.r.t. (W1, b1, W2, b2, W3, b3) through the HIP backward kernels."""
        nx, ny, nz = cache["grid"]
        B = cache["B"]
        nnode = nx * ny * nz
        st = _lib.stream()
        # z = dinv r + s sqrt(dinv) mask y  =>  dL/dy = s sqrt(dinv) mask gz
        gy = (gz * cache["s"][:, None] * torch.sqrt(cache["dinv"]) * cache["mask"]).reshape(B, nnode, 3).contiguous()
        acts = cache["acts"]
        grads = [None] * 6
        g_out = gy
        for li in reversed(range(len(LAYERS))):
            ci, co, act = LAYERS[li]
            y = acts[li + 1]
            if act:
                gpre = torch.empty_like(g_out)
                _lib.call("hf_cnn_tanh_bwd_f32", _lib.ptr(y), _lib.ptr(g_out), g_out.numel(), _lib.ptr(gpre), st)
            else:
                gpre = g_out
            xin = acts[li]
            gw = torch.empty_like(self.W[li])
            gb = torch.empty_like(self.B[li])
            wsz = _lib.load().hf_cnn_bwd_weight_workspace(ci, co, nx, ny, nz, B)
            ws = torch.empty(int(wsz), dtype=torch.uint8, device=gz.device)
            _lib.call("hf_cnn_conv3d_bwd_weight_f32", _lib.ptr(xin), ci, _lib.ptr(gpre), co, nx, ny, nz, B,
                      _lib.ptr(ws), _lib.ptr(gw), _lib.ptr(gb), st)
            grads[2 * li], grads[2 * li + 1] = gw, gb
            if li > 0:
                gin = torch.empty(B, nnode, ci, dtype=torch.float32, device=gz.device)
                _lib.call("hf_cnn_conv3d_bwd_data_f32", _lib.ptr(gpre), co, _lib.ptr(self.W[li]), ci, nx, ny, nz, B,
                          _lib.ptr(gin), st)
                g_out = gin
        return grads

    def apply(self, A, r, problem):
        """z = M(r) for one residual on the solve's inference path (MFMA, f16 activations)."""
        grid = self.grid(problem)
        r = r.to(torch.float32).contiguous()
        nnode = grid[0] * grid[1] * grid[2]
        ws = torch.empty(int(_lib.load().hf_lpc_apply_workspace_size(nnode)), dtype=torch.uint8, device=r.device)
        z = torch.empty_like(r)
        # every tensor passed by pointer stays referenced until the kernels are queued
        gfield, mask, dinv = self.gfield(problem), self.mask(problem), A.dinv()
        _lib.call("hf_lpc_apply_f32", _lib.ptr(r), _lib.ptr(dinv), _lib.ptr(gfield), _lib.ptr(mask),
                  self.weight_ptrs(), grid[0], grid[1], grid[2], _lib.ptr(ws), _lib.ptr(z), _lib.stream())
        return z

    # -- the solve ---------------------------------------------------------------
    def solve(self, A, b, x0, problem, rtol=1e-6, atol=0.0, maxiter=10000, rel_r0=False, poll=8, stall=60,
              trunc=None, stats=None):
        """Truncated flexible CG with M (each direction A-orthogonal to the last
        `trunc`, default self.trunc); on breakdown / stagnation / non-finite values,
        Jacobi-PCG continues (from the FCG iterate, or from x0 if it is not finite).
        Returns (x, info, fallback_taken)."""
        from .solver import krylov
        if A.dtype != torch.float32:
            raise ValueError("the learned preconditioner runs on the FP32 tangent")
        grid = self.grid(problem)
        gfield = self.gfield(problem)
        mask = self.mask(problem)
        N = A.shape[0]
        b = b.to(torch.float32).contiguous()
        x = (torch.zeros(N, dtype=torch.float32, device=b.device) if x0 is None else
             x0.to(torch.float32).clone()).contiguous()
        x_start = x.clone()
        ws = torch.empty(int(_lib.load().hf_lpc_workspace_size(N // 3)), dtype=torch.uint8, device=b.device)
        info = _lib.HFKrylovInfo()
        ptrs = self.weight_ptrs()
        import time
        t0 = time.perf_counter()
        _lib.call("hf_lpc_solve_f32", A.pattern.handle, _lib.ptr(A.vals), _lib.ptr(A.dinv()), _lib.ptr(b), _lib.ptr(x),
                  ptrs, _lib.ptr(gfield), _lib.ptr(mask), grid[0], grid[1], grid[2], int(maxiter), float(rtol),
                  float(atol), int(poll), 2 if rel_r0 else 0, int(stall), int(trunc or self.trunc), _lib.ptr(ws),
                  ctypes.byref(info), _lib.stream())
        self.last_fcg = (int(info.iterations), int(info.status))
        if stats is not None:
            stats.add("fcg-learned", info.iterations, info.status, info.res_norm, info.b_norm, time.perf_counter() - t0)
        if info.status in (0, 1):
            return x, info, False
        # fallback: Jacobi-PCG to the same absolute target
        if rel_r0:
            r0 = torch.linalg.norm((b - A.matvec(x_start)).double()).item()
            target = max(rtol * r0, atol)
        else:
            target = max(rtol * info.b_norm, atol)
        # continue from the FCG iterate only if it is finite and better than the start
        start = x_start
        if info.status != 3 and bool(torch.isfinite(x).all()):
            rx = torch.linalg.norm((b - A.matvec(x)).double()).item()
            rs = torch.linalg.norm((b - A.matvec(x_start)).double()).item()
            if rx < rs:
                start = x
        x2, info2 = krylov(A, b, start, "pcg", 0.0, target, maxiter, True, stats=stats)
        info2.iterations += info.iterations
        return x2, info2, True


# ---------------------------------------------------------------------------
# training data: heterogeneous linear elasticity with a log-normal modulus
def gaussian_random_field(shape, spacing, sigma, ell, rng):
    """Stationary field with squared-exponential covariance (length ell, std sigma),
    spectral synthesis on a periodic grid padded to twice the size, cropped."""
    pad = tuple(2 * n for n in shape)
    noise = rng.standard_normal(pad)
    ks = [np.fft.fftfreq(n, d=spacing) for n in pad]
    K2 = ks[0][:, None, None] ** 2 + ks[1][None, :, None] ** 2 + ks[2][None, None, :] ** 2
    spec = np.exp(-2.0 * np.pi ** 2 * ell ** 2 * K2)
    f = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.sqrt(spec)))
    f = f[:shape[0], :shape[1], :shape[2]]
    f = (f - f.mean()) / (f.std() + 1e-12) * sigma
    return f


def heterogeneous_modulus(n_cells, seed, sigma=0.5, ell=0.1, lo=0.05, hi=20.0, domain=(1.0, 1.0, 1.0)):
    """E per cell (cell-centroid log-normal field, clipped), box_mesh cell order."""
    rng = np.random.default_rng(seed)
    h = domain[0] / n_cells[0]
    g = gaussian_random_field(tuple(n_cells), h, sigma, ell, rng)
    return np.clip(np.exp(g), lo, hi).reshape(-1)


def elasticity_instance(jfa, n, E_cell, Lz=1.0, ele_type="HEX8"):
    """Linear elasticity on a box with E per cell, u = 0 on x = 0 (the c0/c2 boundary
    conditions); ele_type HEX27 gives the quadratic lattice (3x3x3 Gauss)."""
    nx, ny, nz = (n, n, n) if np.ndim(n) == 0 else n

    class Elastic(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=1.0, nu=0.3, param="scale_E")

    left = lambda p: np.isclose(p[0], 0.0, atol=1e-7)
    bc = [[left] * 3, [0, 1, 2], [lambda p: 0.0] * 3]
    if ele_type == "HEX27":
        prob = Elastic(jfa.box_mesh_hex27(nx, ny, nz, 1.0, ny / nx, Lz * nz / nx), vec=3, dim=3, ele_type="HEX27",
                       quadrature_order=4, dirichlet_bc_info=bc)
    else:
        prob = Elastic(jfa.box_mesh(nx, ny, nz, 1.0, ny / nx, Lz * nz / nx), vec=3, dim=3, dirichlet_bc_info=bc)
    set_modulus(prob, E_cell)
    return prob


def set_modulus(prob, E_cell):
    """The instance's per-cell modulus, held at every quadrature point."""
    prob.internal_vars = [np.repeat(np.asarray(E_cell, dtype=np.float64)[:, None], prob.fe.num_quads, axis=1)]


RHS_KINDS = ("noise", "face", "smooth")


def random_rhs(A, problem, gen, kind):
    """A right-hand side (zero on Dirichlet rows) of one of three kinds:
    'noise'  white noise on every free DOF;
    'face'   a constant load of random direction on the nodes of one random box face
             (the shape of a traction load: large in one layer, zero elsewhere);
    'smooth' A y for a smooth random displacement y (a few low Fourier modes)."""
    mask = LearnedPC.mask(problem)
    dev = A.vals.device
    N = A.shape[0]
    if kind == "noise":
        return torch.randn(N, generator=gen, device=dev, dtype=torch.float32) * mask
    nx, ny, nz = LearnedPC.grid(problem)
    idx = torch.arange(N // 3, device=dev)
    i, j, k = idx // (ny * nz), (idx // nz) % ny, idx % nz
    if kind == "face":
        face = int(torch.randint(0, 6, (1,), generator=gen, device=dev))
        coord, ext = [(i, nx), (j, ny), (k, nz)][face // 2]
        sel = (coord == (0 if face % 2 == 0 else ext - 1)).float()
        d = torch.randn(3, generator=gen, device=dev)
        b = (sel[:, None] * d[None, :]).reshape(-1) * mask
        if not bool((b != 0).any()):   # the face is clamped: no load can act there
            return torch.randn(N, generator=gen, device=dev, dtype=torch.float32) * mask
        return b
    y = torch.zeros(N // 3, 3, device=dev)
    for _ in range(4):
        f = torch.rand(3, generator=gen, device=dev) * 3.0
        ph = torch.rand(3, generator=gen, device=dev) * 6.283
        amp = torch.randn(3, generator=gen, device=dev)
        w = torch.sin(f[0] * 3.1416 * i / nx + ph[0]) * torch.sin(f[1] * 3.1416 * j / ny + ph[1]) * \
            torch.sin(f[2] * 3.1416 * k / nz + ph[2])
        y += w[:, None] * amp[None, :]
    return A.matvec((y.reshape(-1) * mask).contiguous()) * mask


def residual_samples(A, problem, k_list, seed, kind="noise"):
    """Jacobi-PCG residuals r_k = b - A x_k for a random right-hand side of the given
    kind (random_rhs), normalised to r^T D^-1 r = 1.  Returns (len(k_list), N) float32."""
    from .solver import krylov
    gen = torch.Generator(device=A.vals.device).manual_seed(int(seed))
    mask = LearnedPC.mask(problem)
    b = random_rhs(A, problem, gen, kind)
    out = []
    for k in k_list:
        if k == 0:
            r = b.clone()
        else:
            x, _ = krylov(A, b, None, "pcg", 0.0, 0.0, int(k), True)
            r = b - A.matvec(x)
        r = r * mask
        nrm = torch.sqrt((r.double() * r.double() * A.dinv().double()).sum()).float()
        out.append(r / torch.clamp(nrm, min=1e-30))
    return torch.stack(out)


def fcg_residuals(model, A, problem, k_list, seed, gfield=None, kind="noise"):
    """Residuals after k steps of the model's own flexible CG (no fallback) on a
    random right-hand side of the given kind, normalised like residual_samples."""
    gen = torch.Generator(device=A.vals.device).manual_seed(int(seed))
    mask = LearnedPC.mask(problem)
    b = random_rhs(A, problem, gen, kind)
    grid = model.grid(problem)
    gfield = model.gfield(problem) if gfield is None else gfield
    N = A.shape[0]
    ws = torch.empty(int(_lib.load().hf_lpc_workspace_size(N // 3)), dtype=torch.uint8, device=b.device)
    ptrs = model.weight_ptrs()
    out = []
    for k in k_list:
        x = torch.zeros_like(b)
        if k > 0:
            info = _lib.HFKrylovInfo()
            _lib.call("hf_lpc_solve_f32", A.pattern.handle, _lib.ptr(A.vals), _lib.ptr(A.dinv()), _lib.ptr(b),
                      _lib.ptr(x), ptrs, _lib.ptr(gfield), _lib.ptr(mask), grid[0], grid[1], grid[2], int(k), 0.0,
                      0.0, 8, 0, 1 << 30, model.trunc, _lib.ptr(ws), ctypes.byref(info), _lib.stream())
        r = (b - A.matvec(x)) * mask
        if not bool(torch.isfinite(r).all()):
            r = b.clone()
        nrm = torch.sqrt((r.double() * r.double() * A.dinv().double()).sum()).float()
        out.append(r / torch.clamp(nrm, min=1e-30))
    return torch.stack(out)


def mg_smoother_residuals(mg, A, problem, k_list, seed, kind="noise"):
    """Residuals a fine-level multigrid post-smoother sees: for a random right-hand
    side of the given kind, x_k after k multigrid flexible-CG iterations, r_k = b - A x_k,
    then r_k - A e with e one V-cycle of r_k stopped after the fine level's coarse
    correction (multigrid.Multigrid with its hierarchy set up for A).  Normalised
    like residual_samples.  A sampler for train() (sampler=...)."""
    gen = torch.Generator(device=A.vals.device).manual_seed(int(seed))
    mask = LearnedPC.mask(problem)
    b = random_rhs(A, problem, gen, kind)
    out = []
    for k in k_list:
        if k == 0:
            r = b.clone()
        else:
            x, _ = mg.solve(A, b, None, 0.0, 0.0, int(k))
            r = (b - A.matvec(x)) * mask
        e = mg.apply(r, skip_fine_post=True)
        rp = (r - A.matvec(e)) * mask
        if not bool(torch.isfinite(rp).all()):
            rp = r
        nrm = torch.sqrt((rp.double() ** 2 * A.dinv().double()).sum()).float()
        out.append(rp / torch.clamp(nrm, min=1e-30))
    return torch.stack(out)


def train(model, matrices, steps, batch=4, lr=2e-3, k_max=150, seed=0, log=None, log_every=50, self_frac=0.5,
          rhs_kinds=("noise",), sampler=None):
    """self_frac: share of the steps whose residuals come from the model's own
    flexible CG (fcg_residuals) instead of Jacobi-PCG (residual_samples);
    rhs_kinds: right-hand-side kinds (random_rhs) drawn uniformly per step;
    sampler(model, A, problem, k_list, seed, gfield, kind) -> (B, N) residuals
    replaces both sources when given."""
    """Adam on the energy loss over (A, problem) pairs; returns the loss history.

    matrices: callable i -> (TangentMatrix FP32, problem[, key]) or a list of such
    tuples; the input fields are rebuilt whenever the key (default id(problem))
    changes, so a caller that reuses one problem object for several instances
    passes the instance id as the key."""
    opt = torch.optim.Adam(model.params(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=max(steps, 1), eta_min=lr * 0.05)
    rng = np.random.default_rng(seed)
    hist = []
    cache_fields = {}
    for step in range(steps):
        idx = step
        item = matrices(idx) if callable(matrices) else matrices[idx % len(matrices)]
        A, prob = item[0], item[1]
        key = item[2] if len(item) > 2 else id(prob)
        if key not in cache_fields:
            cache_fields = {key: (model.gfield(prob), model.mask(prob), model.grid(prob))}
        g, mask, grid = cache_fields[key]
        ks = sorted(int(k) for k in rng.integers(0, k_max + 1, size=batch))
        sd = rng.integers(1 << 31)
        kind = rhs_kinds[int(rng.integers(len(rhs_kinds)))]
        if sampler is not None:
            R = sampler(model, A, prob, ks, sd, g, kind)
        else:
            R = (fcg_residuals(model, A, prob, ks, sd, gfield=g, kind=kind) if rng.random() < self_frac else
                 residual_samples(A, prob, ks, seed=sd, kind=kind))
        z, cache = model.forward(A, R, g, mask, grid, keep=True)
        Az = torch.stack([A.matvec(zi) for zi in z])
        loss_per = 0.5 * (z * Az).sum(1) - (z * R).sum(1)
        # Jacobi reference energy (the untrained model) for monitoring
        zj = R * A.dinv()
        Azj = torch.stack([A.matvec(zi) for zi in zj])
        ej = 0.5 * (zj * Azj).sum(1) - (zj * R).sum(1)
        gz = (Az - R) / batch
        grads = model.backward(cache, gz)
        if not (bool(torch.isfinite(loss_per).all()) and all(bool(torch.isfinite(gp).all()) for gp in grads)):
            sched.step()   # a degenerate sample: skip the update, keep the schedule
            if log:
                log("step %d: non-finite loss or gradient, update skipped" % step)
            continue
        for p, gp in zip(model.params(), grads):
            p.grad = gp
        opt.step()
        sched.step()
        hist.append((float(loss_per.mean()), float(ej.mean())))
        if log and (step % log_every == 0 or step == steps - 1):
            log("step %d loss %.5f (jacobi %.5f)" % (step, hist[-1][0], hist[-1][1]))
    return hist
