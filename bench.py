"""Benchmark: DOF/s and wall-to-1e-8 FP64 residual on one MI355X.

Workload (default, 1 GPU): heterogeneous linear elasticity, HEX8 128x128x64 on
[0,1]x[0,1]x[0,0.5] (1,048,576 cells, 3,244,995 DOF), E per cell a clipped
log-normal Gaussian random field (seed 0, sigma 0.5, correlation length 0.1,
clipped to [0.05, 20]), nu = 0.3, u = 0 on x = 0, traction (0,0,-10) on x = 1;
solved to ||R||/||R0|| <= 1e-8 (FP64 residual).  `--workload` selects the
others: neohookean-hex8-64 (5 load steps), neohookean-hex27-48 (10 load
steps), simp-hex8-96x48x48 (+ adjoint), elastic-hex8-20.

A "step" is the whole solve (every load step).  Methods (FP32 tangent + FP64
Newton residual/update, i.e. FP64 iterative refinement, Eisenstat-Walker
forcing, unless stated):
  mg        : flexible CG preconditioned by geometric-multigrid V-cycles (Galerkin
              levels, Chebyshev smoothing; box lattices).  This is `value` where
              the mesh is a box lattice.
  mixed     : Jacobi-PCG with FP64 dot products (`value` on other meshes).
  learned   : the multigrid with the learned CNN as the fine level's post-smoother
              (reported as learned_pc); --learned-fcg also times the CNN as the
              whole preconditioner of a flexible CG.
  plain     : the drop-in: everything FP64, Jacobi-PCG to 1e-10 per Newton step
              (the reference's jax_solve tolerances).
  reference : the reference's defaults, FP64 Jacobi-BiCGSTAB tol = atol = 1e-10;
              its counts drive the CPU baseline's extrapolation.
roofline: the dominant kernel of the `value` path (the fine-level Chebyshev step
of the multigrid; the SELL product on HEX27, the Krylov SpMV without multigrid),
timed live with HIP events; traffic is its HBM bytes per launch from rocprofv3
PMC counters committed under profiles/ (null if absent).  Every one of these
kernels has an arithmetic intensity below MI355X's FP32 ridge point, so its
roofline is HBM; the matrix-free step also reports its flop rate.
cpu_baseline: the FP64 oracle on this host, in full for elastic-hex8-20,
extrapolated from a bounded sample otherwise (`--cpu-only full` times a whole
workload on the CPU alone and prints the sampled estimate beside it).
`--calibrate-copy` runs only a 256 MiB device copy (the PMC calibration kernel).
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def load_pkg():
    import importlib.util
    if "jax_fem_amd" in sys.modules:
        return sys.modules["jax_fem_amd"]
    pkg = os.path.join(ROOT, "jax-fem_amd")
    spec = importlib.util.spec_from_file_location("jax_fem_amd", os.path.join(pkg, "__init__.py"),
                                                  submodule_search_locations=[pkg])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["jax_fem_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_problem(jfa, n):
    import numpy as np

    class HyperElasticity(jfa.Problem):
        def get_tensor_map(self):
            return jfa.NeoHookean(E=10.0, nu=0.3)

    left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-5)

    def bc(d):
        return [[left, left, left, right], [0, 1, 2, 0], [lambda p: 0.0] * 3 + [lambda p: d]]

    prob = HyperElasticity(jfa.box_mesh(n, n, n), vec=3, dim=3, dirichlet_bc_info=bc(0.0))
    return prob, bc


def build_hex27(jfa, n):
    """Neo-Hookean HEX27 n^3 (3x3x3 Gauss), per-cell shear modulus mu_e (the
    adjoint parameter, uniform E/(2(1+nu)) here) with kappa fixed; BCs as the
    HEX8 workload, u_x = 0.1 applied in 10 load steps."""
    import numpy as np
    E, nu = 10.0, 0.3
    mat = jfa.NeoHookean(E=E, nu=nu, param="mu", kappa=E / (3 * (1 - 2 * nu)))

    class HyperElasticity27(jfa.Problem):
        def get_tensor_map(self):
            return mat

        def set_params(self, p):
            self.internal_vars = [p[:, None].expand(-1, self.fe.num_quads)]

    left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-5)

    def bc(d):
        return [[left, left, left, right], [0, 1, 2, 0], [lambda p: 0.0] * 3 + [lambda p: d]]

    prob = HyperElasticity27(jfa.box_mesh_hex27(n, n, n), vec=3, dim=3, ele_type="HEX27", quadrature_order=4,
                             dirichlet_bc_info=bc(0.0))
    import torch
    prob.set_params(torch.full((prob.num_cells,), E / (2 * (1 + nu)), dtype=torch.float64, device="cuda"))
    return prob, bc


def build_hetero(jfa, nx=128, ny=128, nz=64, seed=0):
    """Heterogeneous linear elasticity on [0,1]x[0,1]x[0,0.5]: E per cell a clipped
    log-normal Gaussian random field (sigma 0.5, correlation length 0.1), nu 0.3,
    u = 0 on x = 0, traction (0,0,-10) on x = 1 (surface map (0,0,+10))."""
    import numpy as np
    import torch
    from jax_fem_amd.learned_pc import heterogeneous_modulus
    E = heterogeneous_modulus((nx, ny, nz), seed)

    class Hetero(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=1.0, nu=0.3, param="scale_E")

        def get_surface_maps(self):
            return [lambda u, x: np.array([0.0, 0.0, 10.0])]

    left = lambda p: np.isclose(p[0], 0.0, atol=1e-7)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-7)
    prob = Hetero(jfa.box_mesh(nx, ny, nz, 1.0, ny / nx, nz / nx), vec=3, dim=3,
                  dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.0] * 3], location_fns=[right])
    # the modulus field is an input: resident in HBM before the timed region, as the mesh is
    prob.internal_vars = [torch.as_tensor(np.repeat(E[:, None], 8, axis=1), device="cuda")]
    return prob


def build_simp(jfa, nx=96, ny=48, nz=48, seed=0):
    """SIMP cantilever on [0,2]x[0,1]x[0,1]: theta = clip(0.5 + N(0, 0.05^2), 1e-3, 1)
    per cell, E = Emin + (E0 - Emin) theta^3 (E0 = 1, Emin = 1e-3), nu 0.3, fixed
    on x = 0, surface map (0,0,+1) on x = 2, z <= 0.1."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    theta = np.clip(0.5 + 0.05 * rng.standard_normal(nx * ny * nz), 1e-3, 1.0)
    mat = jfa.LinearElastic(nu=0.3, param="simp", E0=1.0, Emin=1e-3, penal=3.0)

    class Topo(jfa.Problem):
        def get_tensor_map(self):
            return mat

        def get_surface_maps(self):
            return [lambda u, x: np.array([0.0, 0.0, 1.0])]

        def set_params(self, p):
            self.internal_vars = [p[:, None].expand(-1, 8) if torch.is_tensor(p) else np.repeat(p[:, None], 8, 1)]

    fixed = lambda p: np.isclose(p[0], 0.0, atol=1e-7)
    load = lambda p: np.isclose(p[0], 2.0, atol=1e-7) & (p[2] <= 0.1 + 1e-7)
    prob = Topo(jfa.box_mesh(nx, ny, nz, 2.0, 1.0, 1.0), vec=3, dim=3,
                dirichlet_bc_info=[[fixed] * 3, [0, 1, 2], [lambda p: 0.0] * 3], location_fns=[load])
    prob.set_params(torch.as_tensor(theta, device="cuda"))   # the density field resident in HBM
    return prob, theta


def build_elastic20(jfa, n=20):
    """Linear elasticity on the unit cube, E = 70e3, nu = 0.3, u = 0 on x = 0,
    traction (0,0,-10) on x = 1 (surface map (0,0,+10))."""
    import numpy as np

    class Elastic(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=70e3, nu=0.3)

        def get_surface_maps(self):
            return [lambda u, x: np.array([0.0, 0.0, 10.0])]

    left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-5)
    return Elastic(jfa.box_mesh(n, n, n), vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.0] * 3],
                   location_fns=[right])


def run_sequence(jfa, prob, bc, opts, nload=5, dmax=0.1):
    import torch
    sol = None
    infos = []
    for k in range(1, nload + 1):
        prob.fe.update_Dirichlet_boundary_conditions(bc(dmax * k / nload))
        o = {"newton": dict(opts["newton"], initial_guess=sol)}
        sol = jfa.solver(prob, o)
        infos.append(dict(prob.solve_info))
    torch.cuda.synchronize()
    return sol, infos


def _plane_x(x1):
    """Location function of the plane x = x1 (one argument: the point)."""
    return lambda p: abs(p[0] - x1) < 1e-9


def _oracle_case(wl, cells=None):
    """The workload's problem as an oracle.solver.OracleProblem (the FP64 NumPy/SciPy
    restatement of the reference algorithm: the CPU baseline), at its full size or on
    a box of `cells` cells with the same cell size.  Returns (problem, load steps:
    a list of Dirichlet specs, or [None] for one solve)."""
    import numpy as np
    from oracle import fe as ofe, kernels as ok, solver as osv
    left = lambda p: abs(p[0]) < 1e-9
    if wl == "hetero-elastic-hex8-128x128x64":
        from jax_fem_amd.learned_pc import heterogeneous_modulus
        full = (128, 128, 64)
        c = cells or full
        L = tuple(n / 128.0 for n in c)
        E = heterogeneous_modulus(c, 0, domain=L)
        pts, cl = ofe.box_mesh(*c, *L)
        right = _plane_x(L[0])
        op = osv.OracleProblem(pts, cl, 3, ok.Material("linear_elastic", E=1.0, nu=0.3, param="scale_E"),
                               dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.0] * 3], location_fns=[right],
                               surface_maps=[lambda x: np.broadcast_to(np.array([0.0, 0.0, 10.0]), x.shape)],
                               params=np.repeat(E[:, None], 8, axis=1))
        return op, [None]
    if wl == "elastic-hex8-20":
        c = cells or (20, 20, 20)
        L = tuple(n / 20.0 for n in c)
        pts, cl = ofe.box_mesh(*c, *L)
        right = _plane_x(L[0])
        op = osv.OracleProblem(pts, cl, 3, ok.Material("linear_elastic", E=70e3, nu=0.3),
                               dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.0] * 3], location_fns=[right],
                               surface_maps=[lambda x: np.broadcast_to(np.array([0.0, 0.0, 10.0]), x.shape)])
        return op, [None]
    if wl == "neohookean-hex8-64":
        c = cells or (64, 64, 64)
        L = tuple(n / 64.0 for n in c)
        pts, cl = ofe.box_mesh(*c, *L)
        right = _plane_x(L[0])

        def bc(d):
            return [[left] * 3 + [right], [0, 1, 2, 0], [lambda p: 0.0] * 3 + [lambda p, d=d: d]]
        op = osv.OracleProblem(pts, cl, 3, ok.Material("neo_hookean", E=10.0, nu=0.3), dirichlet_bc_info=bc(0.0))
        return op, [bc(0.1 * k / 5) for k in range(1, 6)]
    raise ValueError("no CPU baseline for workload %s" % wl)


def _oracle_solve(op, steps, progress=None):
    """The reference's solve of a workload: for each load step a Newton solve with
    Jacobi-BiCGSTAB (tol = atol = 1e-10, the jax_solve defaults, solver.py:63-92),
    warm-started from the last.  Returns (seconds, residual+tangent updates, BiCGSTAB iterations)."""
    import numpy as np
    from oracle import solver as osv
    u = None
    t0 = time.perf_counter()
    upd = its = 0
    for spec in steps:
        if spec is not None:
            op.update_bcs(spec)
        u, info = osv.newton_solve(op, u0=u, method="bicgstab", log=progress)
        upd += info["newton_iters"] + 1
        its += sum(info["lin_iters"])
    return time.perf_counter() - t0, upd, its


def _oracle_unit_times(op):
    """One residual + tangent update (apply_bc_vec(compute_residual) + get_A, what
    newton_update_helper does per Newton iterate, solver.py:1259-1276), one SpMV and
    one Jacobi-BiCGSTAB iteration of the oracle on `op`, in seconds."""
    import numpy as np
    from oracle import solver as osv
    u = 1e-4 * np.random.default_rng(0).standard_normal(op.N)   # small: no element inverts
    op.apply_bc_vec(op.compute_residual(u), u)                    # warm caches / lazily built data
    t0 = time.perf_counter()
    R = op.apply_bc_vec(op.compute_residual(u), u)
    A = op.assemble(u)
    t_upd = time.perf_counter() - t0
    x = np.random.default_rng(1).standard_normal(op.N)
    t0 = time.perf_counter()
    for _ in range(10):
        A @ x
    t_spmv = (time.perf_counter() - t0) / 10
    d = A.diagonal()
    k = 40
    t0 = time.perf_counter()
    _, it, _ = osv.bicgstab(A, -R, np.zeros(op.N), lambda v: v / d, 0.0, 0.0, k)
    t_it = (time.perf_counter() - t0) / max(it, 1)
    return t_upd, t_spmv, t_it, A.nnz


SAMPLE_CELLS = {"hetero-elastic-hex8-128x128x64": (40, 40, 20), "neohookean-hex8-64": (32, 32, 32),
                "elastic-hex8-20": (10, 10, 10)}
FULL_CELLS = {"hetero-elastic-hex8-128x128x64": (128, 128, 64), "neohookean-hex8-64": (64, 64, 64),
              "elastic-hex8-20": (20, 20, 20)}


def cpu_estimate(wl, updates, iters):
    """Oracle wall time of the whole workload, extrapolated from a bounded sample:
    residual + tangent updates scale with the cell count, a BiCGSTAB iteration is
    2 SpMVs (scaled by nnz) plus vector work (scaled by N); `updates` and `iters`
    are the counts of the reference's method (Jacobi-BiCGSTAB, jax stop rule)."""
    import numpy as np
    sub = SAMPLE_CELLS[wl]
    op, _ = _oracle_case(wl, sub)
    t_upd, t_spmv, t_it, nnz_s = _oracle_unit_times(op)
    nx, ny, nz = FULL_CELLS[wl]
    cells_full, cells_s = nx * ny * nz, int(np.prod(sub))
    N_full, N_s = 3 * (nx + 1) * (ny + 1) * (nz + 1), op.N
    nnz_full = 9 * (3 * nx + 1) * (3 * ny + 1) * (3 * nz + 1)
    t_vec = max(t_it - 2 * t_spmv, 0.0)
    wall = updates * t_upd * cells_full / cells_s + iters * (2 * t_spmv * nnz_full / nnz_s + t_vec * N_full / N_s)
    return wall, {"sample_cells": list(sub), "t_update_s": t_upd, "t_spmv_s": t_spmv, "t_bicgstab_iter_s": t_it,
                  "updates": int(updates), "bicgstab_iters": int(iters)}


def cpu_baseline(wl, dof, nload, ref_counts):
    """The reported CPU baseline: the oracle timed on this host.  The 20^3 workload
    runs in full; larger ones are extrapolated from a bounded sample (cpu_estimate)
    with the residual+tangent update and BiCGSTAB iteration counts of the reference's
    method as measured by the FP64 Jacobi-BiCGSTAB run on the GPU (ref_counts)."""
    cores = len(os.sched_getaffinity(0))
    common = {"unit": "DOF/s", "cores": cores, "kind": "port", "cpu_model": _cpu_model(),
              "threads_note": "BLAS-backed NumPy parts may use every core of the process's affinity; SciPy "
                              "sparse products (the Krylov iterations) run on one thread"}
    if wl == "elastic-hex8-20":
        op, steps = _oracle_case(wl)
        _oracle_solve(op, steps)                           # warm-up
        runs = sorted(_oracle_solve(_oracle_case(wl)[0], steps)[0] for _ in range(3))
        wall = runs[1]
        return dict(common, value=dof * nload / wall, wall_s=wall, runs_s=runs,
                    sample="full oracle solve (FP64 NumPy/SciPy restatement: one Newton step, Jacobi-BiCGSTAB "
                           "1e-10), one warm-up, median of 3")
    wall, parts = cpu_estimate(wl, ref_counts["updates"], ref_counts["iters"])
    return dict(common, value=dof * nload / wall, wall_s=wall, parts=parts,
                sample=("oracle timed on a %s-cell box of the same problem (one residual+tangent update %.2f s, "
                        "SpMV %.4f s, BiCGSTAB iteration %.4f s), scaled by cells / nnz / DOF to the full mesh "
                        "and by the GPU FP64 Jacobi-BiCGSTAB run's %d updates and %d iterations -> %.0f s "
                        "estimated wall" % ("x".join(map(str, parts["sample_cells"])), parts["t_update_s"],
                                            parts["t_spmv_s"], parts["t_bicgstab_iter_s"], parts["updates"],
                                            parts["bicgstab_iters"], wall)))


def cpu_only(wl, mode):
    """--cpu-only: the oracle alone (no GPU).  'full' times the whole workload (one
    warm-up unless --no-warmup, median of 3 unless --runs) and the sampled estimate
    for the same counts beside it; 'sample' prints the unit times."""
    nx, ny, nz = FULL_CELLS[wl]
    out = {"workload": wl, "dof": 3 * (nx + 1) * (ny + 1) * (nz + 1), "cpu_model": _cpu_model(),
           "host_cores_available": len(os.sched_getaffinity(0))}
    if mode == "full":
        op, steps = _oracle_case(wl)
        runs = []
        nrun = int(os.environ.get("HF_CPU_RUNS", "3"))
        if os.environ.get("HF_CPU_WARMUP", "1") == "1":
            _oracle_solve(op, steps)
        for _ in range(nrun):
            op, steps = _oracle_case(wl)
            t, upd, its = _oracle_solve(op, steps, progress=log)
            runs.append(t)
            log("full oracle run: %.1f s, %d updates, %d BiCGSTAB iterations" % (t, upd, its))
        runs.sort()
        out.update(full_runs_s=runs, full_median_s=runs[len(runs) // 2], updates=upd, bicgstab_iters=its)
        est, parts = cpu_estimate(wl, upd, its)
        out.update(estimate_s=est, estimate_parts=parts, estimate_over_full=est / out["full_median_s"])
    else:
        est, parts = cpu_estimate(wl, 1, 1)
        out.update(unit_times=parts)
    print(json.dumps(out))


def _pmc_traffic(name):
    """HBM bytes per launch from a committed PMC summary under profiles/, or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def _calibrate_copy():
    import torch
    n = 256 * 1024 * 1024
    x = torch.ones(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read_per_launch": n, "bytes_written_per_launch": n,
                      "launches": 3}))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


WORKLOADS = {
    "neohookean-hex8-64": "neo-hookean hex8 64^3, 5 load steps, u_x=0.1 on x=1",
    "neohookean-hex27-48": "neo-hookean hex27 48^3 (3x3x3 Gauss), per-cell mu, 10 load steps, u_x=0.1 on x=1",
    "hetero-elastic-hex8-128x128x64": "heterogeneous linear elasticity hex8 128x128x64, log-normal E (seed 0), "
                                      "traction (0,0,-10) on x=1",
    "simp-hex8-96x48x48": "SIMP cantilever hex8 96x48x48 on [0,2]x[0,1]x[0,1], theta=0.5+N(0,0.05^2) (seed 0), "
                          "p=3, load (0,0,-1) on x=2, z<=0.1; forward + adjoint of the compliance",
    "elastic-hex8-20": "linear elasticity hex8 20^3, E=70e3, traction (0,0,-10) on x=1",
}


def solve_case(jfa, prob, bc, opts, nload, dmax):
    if bc is None:
        sol = jfa.solver(prob, opts)
        return sol, [dict(prob.solve_info)]
    return run_sequence(jfa, prob, bc, opts, nload, dmax)


def hex27_jacobian_mfma(jfa, prob):
    """HIP-event time of one FP32 HEX27 element-Jacobian launch (MFMA) at u = 0,
    averaged over 5 launches after a warm one."""
    import ctypes
    import torch
    pts, cells = prob._dev()
    u = torch.zeros(prob.N, dtype=torch.float64, device="cuda")
    ej = torch.empty(prob.num_cells * 27 * 27 * 9, dtype=torch.float32, device="cuda")
    mat, prm = prob._hfmat(), prob._params()
    st = jfa._lib.stream()

    def launch():
        jfa._lib.call("hf_cell_jacobian_f32", 27, ctypes.byref(mat), jfa._lib.ptr(pts), jfa._lib.ptr(cells),
                      prob.num_cells, jfa._lib.ptr(u), jfa._lib.ptr(prm), jfa._lib.ptr(ej), st)
    launch()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    reps = 5
    for _ in range(reps):
        launch()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    # the six blocks i <= k of the symmetric tangent (the other three are written
    # transposed): per block the GEMM K^{ik} = G^T V^{ik} and forming V^{ik} = C G
    flop_block = (2 * 3 * 81 * 81 * 27 + 2 * 3 * 81 * 81 * 3) // 9
    flop_cell = 6 * flop_block
    tf = prob.num_cells * flop_cell / (ms * 1e-3) / 1e12
    return {"kernel": "hex27_jacobian_mfma<NEO_HOOKEAN> (wavefront per cell, the 6 blocks i <= k of the "
                      "symmetric tangent as 27x27x84 GEMMs on v_mfma_f32_32x32x2f32, 3 written transposed)",
            "avg_ms": ms, "flop_per_cell": flop_cell,
            "achieved": tf, "peak": 157.3, "unit": "TFLOP/s", "frac": tf / 157.3,
            "achieved_9_block_equivalent": tf * 9 / 6,
            "output_GBps": prob.num_cells * 6561 * 4 / (ms * 1e-3) / 1e9,
            "note": "flop count of the algorithmic factored form for the six blocks formed (per (i,k): V = C G, "
                    "K = G^T V; 1,032,192 flop per cell executed on MFMA with the 27 -> 32 padding); "
                    "achieved_9_block_equivalent counts all nine blocks as the earlier form computed them; "
                    "the dense B^T D B form counts 3,542,940 flop/cell"}


def hex27_adjoint(jfa, prob, u, fd_cells, adj_opts):
    """Gradients of the total strain energy Pi = sum JxW psi(F; mu_e, kappa) and of
    S = sum u^2 at the last load step with respect to the per-cell shear modulus
    mu_e, through the reference's implicit_vjp (solver.py:1324-1380): the transposed
    tangent solve by multigrid flexible CG with FP64 refinement (adj_opts), the
    parameter VJP by the HEX27 cell kernel; Pi's explicit dependence on mu_e added
    by autograd.  Checked by central differences (h = 1e-4) on fd_cells seeded
    cells: mixed-precision Newton re-solves warm-started at u* to a 1e-10 relative
    residual (of the perturbation), Pi and S differenced point by point (sum of
    differences)."""
    import numpy as np
    import torch
    from jax_fem_amd.solver import implicit_vjp
    kappa = float(prob.material.kappa)
    nq = prob.fe.num_quads
    terms = prob._pointwise()
    jxw = terms.jxw.reshape(-1)
    eye = torch.eye(3, dtype=torch.float64, device="cuda")
    mu0 = prob.internal_vars[0][:, 0].detach().clone()
    u = u.reshape(-1).contiguous()

    def pi_points(uf, mu_cells):
        ug, _ = terms.point_data(uf.contiguous(), False)
        F = ug + eye
        J = torch.linalg.det(F)
        I1 = (F * F).sum(dim=(-2, -1))
        m = mu_cells[:, None].expand(-1, nq).reshape(-1)
        return jxw * (0.5 * m * (J ** (-2.0 / 3.0) * I1 - 3.0) + 0.5 * kappa * (J - 1.0) ** 2)

    out = {"objectives": ["total strain energy Pi", "sum u^2"], "adjoint_method": "mg, FP64 refinement"}
    mu_leaf = mu0.clone().requires_grad_(True)
    Pi = pi_points(u, mu_leaf).sum()
    explicit = torch.autograd.grad(Pi, mu_leaf)[0]
    grads, times = [], []
    for v in (prob.residual_vec(u), 2.0 * u):        # dPi/du = internal force; dS/du = 2u
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = implicit_vjp(prob, [u.reshape(-1, 3)], mu0.clone(), [v.reshape(-1, 3)], adj_opts)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        grads.append(g)
        out.setdefault("adjoint_refinement", []).append(getattr(prob, "adjoint_info", None))
    grads[0] = grads[0] + explicit
    prob.set_params(mu0)
    out["adjoint_s"] = times
    out["value"] = [float(Pi.detach()), float((u * u).sum())]
    out["grad_norm"] = [float(torch.linalg.norm(g)) for g in grads]
    if fd_cells:
        rng = np.random.default_rng(0)
        cells = rng.choice(prob.num_cells, fd_cells, replace=False)
        fd_opts = {"newton": {"precision": "mixed", "tol": 0.0, "rel_tol": 1e-10, "max_iter": 10,
                              "initial_guess": [u.reshape(-1, 3)], "linear": {"mg": {}}}}
        h = 1e-4
        errs = [[], []]
        t0 = time.perf_counter()
        for c in cells:
            sols, mus = [], []
            for sgn in (1, -1):
                m = mu0.clone()
                m[c] += sgn * h
                prob.set_params(m)
                sols.append(jfa.solver(prob, fd_opts)[0].reshape(-1))
                mus.append(m)
            d_pi = float((pi_points(sols[0], mus[0]) - pi_points(sols[1], mus[1])).sum()) / (2 * h)
            d_s = float(((sols[0] - sols[1]) * (sols[0] + sols[1])).sum()) / (2 * h)
            for k, fd in enumerate((d_pi, d_s)):
                gk = float(grads[k][c])
                errs[k].append(abs(fd - gk) / max(abs(gk), 1e-300))
            log("fd cell %d: Pi fd %.10e adjoint %.10e | S fd %.10e adjoint %.10e" % (
                c, d_pi, float(grads[0][c]), d_s, float(grads[1][c])))
        prob.set_params(mu0)
        out["fd_cells"] = int(fd_cells)
        out["fd_h"] = h
        out["fd_s"] = time.perf_counter() - t0
        out["fd_max_rel_err"] = [max(e) for e in errs]
    return out


def simp_adjoint(jfa, prob, theta, fd_cells, opts_fwd, adj_opts=None):
    """Compliance C = integral of t.u over the loaded face = -f_ext.u (the surface
    map is s = -t): gradient w.r.t. theta by ad_wrapper (transposed solve +
    parameter VJP), timed, and checked by central differences (h = 1e-4) on
    fd_cells seeded cells with FP64 forward solves to a 1e-12 relative residual."""
    import numpy as np
    import torch
    th = torch.tensor(theta, dtype=torch.float64, device="cuda", requires_grad=True)
    ref_opts = {"pcg": {"tol": 1e-12, "atol": 0.0}}
    adj_opts = adj_opts or ref_opts
    fwd = jfa.ad_wrapper(prob, opts_fwd, adj_opts)
    fext = prob._external()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u = fwd(th)[0]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    obj = -(fext * u.reshape(-1)).sum()
    obj.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g = th.grad.detach().cpu().numpy()
    out = {"objective": "compliance", "value": float(obj.detach()), "forward_s": t1 - t0, "adjoint_s": t2 - t1,
           "grad_norm": float(np.linalg.norm(g)), "adjoint_method": next(iter(adj_opts))}
    if "mg" in adj_opts:
        out["adjoint_refinement"] = getattr(prob, "adjoint_info", None)
        # the same gradient with the FP64 Jacobi-PCG adjoint (tol 1e-12) beside it
        th2 = th.detach().clone().requires_grad_(True)
        fwd2 = jfa.ad_wrapper(prob, opts_fwd, ref_opts)
        u2 = fwd2(th2)[0]
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        (-(fext * u2.reshape(-1)).sum()).backward()
        torch.cuda.synchronize()
        g2 = th2.grad.detach().cpu().numpy()
        out["fp64_pcg_adjoint_s"] = time.perf_counter() - t3
        out["rel_l2_vs_fp64_pcg_adjoint"] = float(np.linalg.norm(g - g2) / np.linalg.norm(g2))
    if fd_cells:
        rng = np.random.default_rng(0)
        cells = rng.choice(len(theta), fd_cells, replace=False)
        # warm start from u*: the relative tolerance then applies to the O(h) residual
        # of the perturbation, i.e. 1e-10 of the change in C, not of C itself
        fd_opts = {"newton": {"precision": "fp64", "tol": 0.0, "rel_tol": 1e-10, "forcing": "fixed", "max_iter": 5,
                              "initial_guess": [u.detach()], "linear": {"pcg": {"tol": 1e-12, "atol": 0.0}}}}
        errs = []
        for c in cells:
            vals = []
            for sgn in (1, -1):
                t = th.detach().clone()
                t[c] += sgn * 1e-4
                prob.set_params(t)
                uu = jfa.solver(prob, fd_opts)[0]
                vals.append(float(-(fext * uu.reshape(-1)).sum()))
            fd = (vals[0] - vals[1]) / 2e-4
            errs.append(abs(fd - g[c]) / max(abs(g[c]), 1e-30))
            log("fd cell %d: fd %.10e adjoint %.10e rel err %.2e" % (c, fd, g[c], errs[-1]))
        prob.set_params(th.detach())
        out["fd_cells"] = int(fd_cells)
        out["fd_max_rel_err"] = float(max(errs))
        out["fd_h"] = 1e-4
    return out


def fine_model_bytes(n0):
    """HBM bytes of one flexible-CG iteration's fine-level work per the kernels'
    algorithmic traffic: the smoothing passes read the FP16 copy of the half-stored
    stencil (14 blocks x 5 half2 words = 280 B per node) and the FP32 block inverse
    (36 B), the outer product the FP32 stencil (126 x 4 B): update 180, pre-smoothing
    step 388, residual 328, restriction read 12, prolongation 36, post-smoothing start
    376 and step 388, outer product 564 bytes per node.  The coarse levels (about a
    sixth more) are not counted."""
    return n0 * FINE_NODE_BYTES


FINE_NODE_BYTES = 180 + 388 + 328 + 12 + 36 + 376 + 388 + 564
# the same with matrix-free fine-level rows (Problem.cell_operator): no stencil bytes, a
# 4-byte cell modulus per node instead: update 180, pre-smoothing step 112, residual 52,
# restriction 12, prolongation 36, post-smoothing start 100 and step 112, outer product 64
FINE_NODE_BYTES_MF = 180 + 112 + 52 + 12 + 36 + 100 + 112 + 64
# flop of one matrix-free smoothing step per node: the row (8 cells x 7 blocks x 9 FMA,
# 8 x 3 scale FMA: 1,056 flop, and 26 x 3 neighbour differences) and the update (33)
MF_STEP_FLOP = 2 * (8 * 7 * 9 + 8 * 3) + 26 * 3 + 33
MF_STEP_BYTES = 112


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="hetero-elastic-hex8-128x128x64", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=None, help="cells per direction (cube workloads)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-ref", action="store_true", help="skip the FP64 Jacobi-BiCGSTAB (reference defaults) leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-fcg", action="store_true",
                    help="also time the learned network as the whole preconditioner of a flexible CG")
    ap.add_argument("--no-mg", action="store_true", help="value path: Jacobi-PCG even on box lattices")
    ap.add_argument("--no-pcg", action="store_true", help="skip the mixed Jacobi-PCG leg beside the multigrid")
    ap.add_argument("--fd-cells", type=int, default=0,
                    help="SIMP and HEX27 workloads: cells whose adjoint gradient is checked by central FD")
    ap.add_argument("--no-adjoint", action="store_true", help="HEX27 workload: skip the adjoint legs")
    ap.add_argument("--jacobian-only", action="store_true",
                    help="HEX27 workload: time the FP32 element-Jacobian kernel alone (for profiling) and exit")
    ap.add_argument("--cpu-only", choices=["full", "sample"], default=None,
                    help="time only the CPU oracle of the workload (no GPU)")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--debug", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the multi-GPU design is documented, not built); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.cpu_only:
        load_pkg()
        cpu_only(args.workload, args.cpu_only)
        return
    if args.calibrate_copy:
        _calibrate_copy()
        return
    import numpy as np
    import torch
    jfa = load_pkg()
    jfa._lib.load()
    if args.debug:
        import logging
        logging.getLogger("jax_fem_amd").setLevel(logging.DEBUG)
    torch.cuda.init()
    wl = args.workload
    bc, nload, dmax, theta = None, 1, 0.0, None
    if wl == "neohookean-hex8-64":
        n = args.n or 64
        prob, bc = build_problem(jfa, n)
        nload, dmax = 5, 0.1
        desc = "neo-hookean hex8 %d^3, 5 load steps, u_x=0.1 on x=1" % n
    elif wl == "neohookean-hex27-48":
        n = args.n or 48
        prob, bc = build_hex27(jfa, n)
        nload, dmax = 10, 0.1
        desc = WORKLOADS[wl].replace("48^3", "%d^3" % n)
    elif wl == "hetero-elastic-hex8-128x128x64":
        prob = build_hetero(jfa)
        desc = WORKLOADS[wl]
    elif wl == "simp-hex8-96x48x48":
        prob, theta = build_simp(jfa)
        desc = WORKLOADS[wl]
    else:
        n = args.n or 20
        prob = build_elastic20(jfa, n)
        desc = WORKLOADS[wl].replace("20^3", "%d^3" % n)
    full_size = args.n is None
    N = prob.N
    log("problem: %s, %d DOF" % (desc, N))
    if args.jacobian_only:
        if wl != "neohookean-hex27-48":
            raise SystemExit("--jacobian-only applies to the HEX27 workload")
        print(json.dumps(hex27_jacobian_mfma(jfa, prob)))
        return
    t0 = time.perf_counter()
    _ = prob.pattern
    torch.cuda.synchronize()
    log("pattern build %.3f s (nnz %d)" % (time.perf_counter() - t0, prob.pattern.nnz))

    # tol=0: stop on the relative FP64 residual 1e-8 only (the reference also stops at ||R|| <= 1e-6)
    mixed = {"newton": {"precision": "mixed", "tol": 0.0, "eta_first": 1e-4, "linear": {"pcg": {}}}}
    mgo = {"newton": {"precision": "mixed", "tol": 0.0, "eta_first": 1e-4, "linear": {"mg": {}}}}
    # plain: the reference's jax_solve tolerances (tol = atol = 1e-10, solver.py:78-84) with
    # Jacobi-PCG.  The SIMP load is small (||R0|| ~ 1e-2), where the absolute 1e-10 would stop
    # the Krylov solves above the 1e-8 relative Newton target: there the tolerance is relative only.
    plain = {"newton": {"precision": "fp64", "tol": 0.0, "forcing": "fixed",
                        "linear": {"pcg": {"tol": 1e-10, "atol": 0.0 if wl == "simp-hex8-96x48x48" else 1e-10}}}}
    # the reference's defaults: FP64 Jacobi-BiCGSTAB, tol = atol = 1e-10, Newton rel_tol 1e-8 / tol 1e-6
    refdef = {"newton": {"linear": {"jax_solver": {}}}}
    if wl == "elastic-hex8-20":
        plain = {"newton": {"tol": 0.0, "linear": {"jax_solver": {}}}}
    from jax_fem_amd.multigrid import lattice
    try:
        lattice(prob)
        use_mg = not args.no_mg
    except NotImplementedError:
        use_mg = False
    main_opts = mgo if use_mg else mixed

    for w in range(args.warmup):
        solve_case(jfa, prob, bc, main_opts, nload, dmax)
        log("warmup %d done" % w)

    import ctypes
    nnz = prob.pattern.nnz
    # the roofline kernel is timed with HIP events in the first timed step (one launch
    # per poll batch that holds it, run eagerly instead of as the captured graph); the
    # later steps run as a caller runs them
    nl, ms = ctypes.c_int64(), ctypes.c_double()
    jfa._lib.call("hf_profile_enable", 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        sol_m, infos_m = solve_case(jfa, prob, bc, main_opts, nload, dmax)
        if s == 0 and args.steps > 1:
            jfa._lib.call("hf_profile_read", ctypes.byref(nl), ctypes.byref(ms))
            jfa._lib.call("hf_profile_enable", 0)
        log("timed step %d done" % s)
    torch.cuda.synchronize()
    wall_m = (time.perf_counter() - t0) / args.steps
    if args.steps <= 1:
        jfa._lib.call("hf_profile_read", ctypes.byref(nl), ctypes.byref(ms))
        jfa._lib.call("hf_profile_enable", 0)
    kern_us = 1e3 * ms.value / max(nl.value, 1)
    tag = {"hetero-elastic-hex8-128x128x64": "hetero_hex8_128x128x64", "neohookean-hex8-64": "hex8_64",
           "neohookean-hex27-48": "hex27_48", "simp-hex8-96x48x48": "simp_hex8_96x48x48",
           "elastic-hex8-20": "elastic_hex8_20"}[wl]
    n0 = prob.N // 3
    kflop = None
    # the fine level ran matrix-free when the hierarchy holds a cell operator
    mf_on = use_mg and getattr(getattr(prob, "_mg", None), "_cell_op", None) is not None
    if use_mg and prob.fe.ele_type == "HEX27":
        # the fine level is the SELL tangent: one smoothing product y = A v from the FP16
        # copy of its values (2 B per value), one column index per node-pair block, v
        # read and y written (4 B per DOF each)
        f16 = getattr(getattr(prob, "_mg", None), "fine_f16", False)
        kbytes = (2 if f16 else 4) * nnz + 4 * prob.pattern.nblocks + 8 * N
        kname = ("mg_sell_spmv16 (fine-level smoothing product y = A v, SELL column-form node blocks, FP16 "
                 "values)" if f16 else "mg_sell_spmv (fine-level smoothing product y = A v, SELL column-form "
                 "node blocks)")
        pmc = ("mg_sell_spmv16_%s_pmc.md" if f16 else "mg_sell_spmv_%s_pmc.md") % tag
    elif use_mg and mf_on:
        # one fine-level Chebyshev smoothing step with matrix-free rows
        # (mg_poly_step<true, MfOp>): per node the 8 cells' 7 off-diagonal blocks of the
        # unit cell matrix (48 values in scalar registers) applied to the 26 neighbour
        # differences, scaled by each cell's modulus; bytes: the vectors (res_in, d_in
        # read, e read and written, res_out, d_out written: 6 x 12 B), the block inverse
        # (36 B) and the cell modulus (4 B)
        kbytes = n0 * MF_STEP_BYTES
        kflop = n0 * MF_STEP_FLOP
        kname = ("mg_poly_step<true, MfOp> (fine-level Chebyshev smoothing step, matrix-free rows: one "
                 "reflection-symmetric cell matrix applied to neighbour differences, per-cell modulus; FP32 FMA)")
        pmc = "mg_poly_step_mf_%s_pmc.md" % tag
    elif use_mg:
        # one fine-level Chebyshev smoothing step (mg_poly_step<true, __half2>): the
        # symmetric 27-point node-block stencil stored as its diagonal and 13 upper
        # blocks, each block's 9 entries as 5 FP16 pairs (280 B per node; the lower
        # blocks are the neighbours' upper ones, read once each in the ideal), the FP32
        # 3x3 block-Jacobi inverse (36 B), and per node the residual, direction and
        # correction read (3 FP32 each) and written (3 each)
        kbytes = prob.N // 3 * (14 * 5 * 4 + 9 * 4 + 6 * 12)
        kname = ("mg_poly_step<true, __half2> (fine-level Chebyshev smoothing step: half-stored symmetric "
                 "27-point node-block stencil, slot-major FP16 pairs)")
        pmc = "mg_poly_step_f16_%s_pmc.md" % tag
    else:
        # one PCG SpMV (w = A u): values (4 B/nnz), one block column per node-pair
        # block, u read and w written (4 B per DOF each)
        kbytes = 4 * nnz + 4 * prob.pattern.nblocks + 8 * N
        kname = "cg_spmv_kernel<float,3> (SELL column-form node-block SpMV w = A u fused with u.w)"
        pmc = "cg_spmv_%s_pmc.md" % tag
    achieved = kbytes / (kern_us * 1e-6) / 1e9
    if kflop is not None:
        achieved_tf = kflop / (kern_us * 1e-6) / 1e12
    rel_m = max(i["rel"] for i in infos_m)
    fcg_iters = sum(i["krylov_iters"] for i in infos_m)
    log("%s: %.4f s/step, newton %s, krylov %s, kernel %.1f us (%d samples) -> %.0f GB/s" % (
        "mg" if use_mg else "mixed", wall_m, [i["newton_iters"] for i in infos_m],
        [i["krylov_iters"] for i in infos_m], kern_us, nl.value, achieved))

    def leg(opts, name, note=None, rerun_under=1.0):
        # a leg whose run takes under rerun_under seconds is timed again (first-use
        # costs such as a graph capture would otherwise dominate it); the second run counts
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol, infos = solve_case(jfa, prob, bc, opts, nload, dmax)
        wall = time.perf_counter() - t0
        if wall < rerun_under:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sol, infos = solve_case(jfa, prob, bc, opts, nload, dmax)
            wall = time.perf_counter() - t0
        out = {"dof_per_s": N * nload / wall, "wall_s": wall,
               "newton_iters": [i["newton_iters"] for i in infos],
               "krylov_iters": [i["krylov_iters"] for i in infos],
               "fallbacks": [i.get("fallbacks", 0) for i in infos],
               "max_rel_residual": max(i["rel"] for i in infos)}
        if note:
            out["note"] = note
        log("%s: %.3f s, newton %s, krylov %s, fallbacks %s" % (
            name, wall, out["newton_iters"], out["krylov_iters"], out["fallbacks"]))
        return sol, out

    def rel_to(sol, ref):
        a, b = sol[0].reshape(-1), ref[0].reshape(-1)
        return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))

    mixed_out = None
    if use_mg and not args.no_pcg:
        _, mixed_out = leg(mixed, "mixed jacobi-pcg", "one run after the timed steps (the second when the first "
                                                      "took under 1 s)")
    learned_out, sol_l, fcg_out = None, None, None
    if not args.no_learned and use_mg:
        # the learned network as the multigrid's fine-level post-smoother (weights trained
        # on post-coarse-correction residuals of heterogeneous elasticity: 32^3 HEX8 cells,
        # or 16^3 HEX27 cells for the quadratic lattice)
        from jax_fem_amd.learned_pc import LearnedPC
        wname = "lpc_mg_smoother_hex27.npz" if prob.fe.ele_type == "HEX27" else "lpc_mg_smoother.npz"
        smoother = LearnedPC.load(os.path.join(ROOT, "jax-fem_amd", "weights", wname))
        learned = {"newton": {"precision": "mixed", "tol": 0.0, "eta_first": 1e-4,
                              "linear": {"mg": {"smoother": "learned", "model": smoother}}}}
        solve_case(jfa, prob, bc, learned, nload, dmax)        # first use: set-up and capture
        sol_l, learned_out = leg(learned, "learned smoother",
                                 "multigrid flexible CG with the learned network as the fine-level post-smoother "
                                 "(weights/%s); second run" % wname)
        learned_out["method"] = ("FP32 tangent + FP64 Newton refinement; multigrid V-cycles with Chebyshev(2) "
                                 "pre-smoothing and one learned-network step as the fine post-smoother")
    if args.learned_fcg and wl not in ("elastic-hex8-20", "neohookean-hex27-48"):   # HEX8 grids only
        from jax_fem_amd.learned_pc import LearnedPC
        model = LearnedPC.load(os.path.join(ROOT, "jax-fem_amd", "weights", "lpc_elastic.npz"))
        lfcg = {"newton": {"precision": "mixed", "tol": 0.0, "eta_first": 1e-4,
                           "linear": {"learned": {"model": model}}}}
        _, fcg_out = leg(lfcg, "learned fcg", "the learned network as the whole preconditioner of a flexible CG "
                                              "(weights/lpc_elastic.npz); single run incl. first-call setup")

    plain_out = None
    cpu = None
    rel_l2 = None
    if not args.no_plain:
        sol_p, plain_out = leg(plain, "plain")
        plain_out["method"] = ("FP64 Jacobi-BiCGSTAB 1e-10 (jax_solver)" if wl == "elastic-hex8-20"
                               else "FP64 Jacobi-PCG 1e-10")
        rel_l2 = rel_to(sol_m, sol_p)
        log("value path vs plain rel-L2 %.2e" % rel_l2)
        if learned_out is not None:
            learned_out["rel_l2_vs_plain"] = rel_to(sol_l, sol_p)
            learned_out["speedup_vs_plain"] = plain_out["wall_s"] / learned_out["wall_s"]
    ref_out = None
    if not args.no_ref and wl != "elastic-hex8-20":
        _, ref_out = leg(refdef, "reference defaults (FP64 Jacobi-BiCGSTAB)")
        ref_out["method"] = "FP64 Jacobi-BiCGSTAB, tol = atol = 1e-10, Newton rel_tol 1e-8 / tol 1e-6 (jax_solve)"
    elif wl == "elastic-hex8-20" and plain_out is not None:
        ref_out = plain_out
    if not args.no_cpu and full_size and wl in SAMPLE_CELLS and ref_out is not None:
        t0 = time.perf_counter()
        counts = {"updates": sum(k + 1 for k in ref_out["newton_iters"]), "iters": sum(ref_out["krylov_iters"])}
        cpu = cpu_baseline(wl, N, nload, counts)
        log("cpu baseline %.1f s of sampling: %s" % (time.perf_counter() - t0, cpu["sample"]))
    whole = None
    if use_mg and prob.fe.ele_type == "HEX8":
        nb = FINE_NODE_BYTES_MF if mf_on else FINE_NODE_BYTES
        wb = n0 * nb * fcg_iters
        whole = {"model": "fine-level flexible-CG traffic only (%d B per fine node per iteration%s; coarse levels, "
                          "tangent assembly and set-up not counted: a lower bound)" % (
                              nb, ", matrix-free rows" if mf_on else ""),
                 "bytes": wb, "wall_s": wall_m, "achieved": wb / wall_m / 1e9, "unit": "GB/s",
                 "frac": wb / wall_m / 1e9 / 8000.0}
    out = {
        "metric": "DOF/s + wall-to-1e-8 FP64 residual: learned-PC/FP32-refined vs plain PCG",
        "value": N * nload / wall_m,
        "unit": "DOF/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall_m * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": wl, "description": desc, "dof": N,
                   "nnz": nnz, "method": ("value: FP32 tangent, flexible CG preconditioned by geometric-multigrid "
                                          "V-cycles (Galerkin levels, Chebyshev(2) smoothing), FP64 Newton refinement "
                                          "(EW forcing); Jacobi-PCG under 'pcg_mixed', the multigrid with the learned "
                                          "network as fine-level post-smoother under 'learned_pc'") if use_mg else
                   ("value: FP32 Jacobi-PCG + FP64 Newton refinement (EW forcing); the learned-PC variant is "
                    "reported under 'learned_pc'"),
                   "wall_to_tol_s": wall_m, "max_rel_residual": rel_m,
                   "newton_iters": [i["newton_iters"] for i in infos_m],
                   "krylov_iters": [i["krylov_iters"] for i in infos_m]},
        "roofline": ({"bound": "hbm", "kernel": kname,
                      "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                      "traffic": _pmc_traffic(pmc) if full_size else None, "avg_launch_us": kern_us,
                      "launches_timed": int(nl.value), "algorithmic_bytes": kbytes} if kflop is None else
                     {"bound": "hbm", "kernel": kname,
                      "achieved": achieved, "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0,
                      "traffic": _pmc_traffic(pmc) if full_size else None, "avg_launch_us": kern_us,
                      "launches_timed": int(nl.value), "algorithmic_bytes": kbytes, "algorithmic_flop": kflop,
                      "flop_achieved_TFLOPs": achieved_tf, "flop_frac_of_fp32_peak": achieved_tf / 157.3,
                      "intensity_flop_per_byte": kflop / kbytes,
                      "note": "arithmetic intensity %.1f flop/B is below the FP32 ridge point (157.3 TFLOP/s / "
                              "8 TB/s = 19.7 flop/B), so HBM is the binding roofline; the FP32 FMA runs on the "
                              "vector ALU (no packed-FP32 instructions: at most 78.6 TFLOP/s issued)" %
                              (kflop / kbytes)}),
        "whole_solve": whole,
        "pcg_mixed": mixed_out,
        "learned_pc": learned_out,
        "learned_fcg": fcg_out,
        "plain": plain_out,
        "reference_defaults": ref_out,
        "speedup_vs_plain": (plain_out["wall_s"] / wall_m) if plain_out else None,
        "cpu_baseline": cpu,
    }
    if rel_l2 is not None:
        out["config"]["rel_l2_vs_fp64_plain"] = rel_l2
    if wl == "neohookean-hex27-48":
        out["mfma"] = hex27_jacobian_mfma(jfa, prob)
        log("hex27 jacobian: %.3f ms, %.1f TFLOP/s" % (out["mfma"]["avg_ms"], out["mfma"]["achieved"]))
        if not args.no_adjoint:
            out["adjoint"] = hex27_adjoint(jfa, prob, sol_m[0], args.fd_cells,
                                           {"mg": {"tol": 1e-12}} if use_mg else {"pcg": {"tol": 1e-12}})
            log("adjoint: %s" % json.dumps(out["adjoint"]))
    if wl == "simp-hex8-96x48x48":
        out["adjoint"] = simp_adjoint(jfa, prob, theta, args.fd_cells, main_opts,
                                      {"mg": {"tol": 1e-12}} if use_mg else None)
        log("adjoint: %s" % json.dumps(out["adjoint"]))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
