"""Geometric multigrid (csrc/mg.hip) through the C ABI against explicit sparse algebra.

* level operators vs P^T A P with P the Kronecker product of 1-D interpolation
  matrices (scipy, FP64), on lattices with odd interval counts; the coarse
  operators of an unconstrained body annihilate rigid translations;
* one V-cycle vs the same cycle written with scipy matrices;
* the preconditioned flexible CG vs the true FP64 residual and Jacobi-PCG's
  iteration count; a Newton solve with it vs the FP64 plain path;
* set-up errors (a pattern that is not the lattice stencil).

Tolerances: level 0 is an exact copy (27 slots stored) or, half-stored, within the
tangent's FP32 asymmetry (2e-6 of its largest entry); coarse levels, the Galerkin
products of the stored level 0, 2e-5
of the operator's largest entry (FP32 accumulation); the V-cycle 1e-4 of its
largest entry; solves by their FP64 residual; Newton rel-L2 1e-6."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def interp_1d(nf):
    """(nf+1, nc+1) linear interpolation from every second node (the last node kept)."""
    nc = (nf + 1) // 2 if nf >= 2 else nf
    pos = [min(2 * c, nf) for c in range(nc + 1)]
    P = np.zeros((nf + 1, nc + 1))
    for c, p in enumerate(pos):
        P[p, c] = 1.0
    for f in range(nf + 1):
        if f in pos:
            continue
        left = max(p for p in pos if p < f)
        right = min(p for p in pos if p > f)
        P[f, pos.index(left)] = (right - f) / (right - left)
        P[f, pos.index(right)] = (f - left) / (right - left)
    return P


def interp_3d(dims):
    Px, Py, Pz = (interp_1d(n - 1) for n in dims)
    P = sp.kron(sp.kron(sp.csr_matrix(Px), sp.csr_matrix(Py)), sp.csr_matrix(Pz))
    return sp.kron(P, sp.identity(3)).tocsr()


def stencil_to_csr(S, dims):
    """(27, 3, 3, n) slot-major stencil -> scipy CSR; slots leaving the lattice must hold zeros."""
    nx, ny, nz = dims
    n = nx * ny * nz
    idx = np.arange(n)
    x, y, z = idx // (ny * nz), (idx // nz) % ny, idx % nz
    rows, cols, vals = [], [], []
    for o in range(27):
        dx, dy, dz = o // 9 - 1, (o // 3) % 3 - 1, o % 3 - 1
        xx, yy, zz = x + dx, y + dy, z + dz
        ok = (xx >= 0) & (xx < nx) & (yy >= 0) & (yy < ny) & (zz >= 0) & (zz < nz)
        assert np.all(S[o][:, :, ~ok] == 0), "stencil slot %d leaves the lattice but is not zero" % o
        nb = (xx * ny + yy) * nz + zz
        for i in range(3):
            for j in range(3):
                rows.append(3 * idx[ok] + i)
                cols.append(3 * nb[ok] + j)
                vals.append(S[o, i, j, ok].astype(np.float64))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * n, 3 * n))


def box_problem(jfa, cells, material="neo", dirichlet=True, E_cell=None):
    left = lambda p: np.isclose(p[0], 0.0, atol=1e-7)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-7)
    mat = jfa.NeoHookean(E=10.0, nu=0.3) if material == "neo" else \
        jfa.LinearElastic(E=1.0, nu=0.3, param="scale_E")

    class P(jfa.Problem):
        def get_tensor_map(self):
            return mat

    bc = [[left] * 3 + [right], [0, 1, 2, 0], [lambda p: 0.0] * 3 + [lambda p: 0.05]] if dirichlet else None
    prob = P(jfa.box_mesh(*cells), vec=3, dim=3, dirichlet_bc_info=bc)
    if E_cell is not None:
        prob.internal_vars = [np.repeat(np.asarray(E_cell, dtype=np.float64)[:, None], 8, axis=1)]
    return prob


def tangent(jfa, prob, seed=0, scale=0.02):
    from importlib import import_module
    S = import_module("jax_fem_amd.solver")
    g = torch.Generator(device=DEV).manual_seed(seed)
    u = scale * torch.randn(prob.N, generator=g, device=DEV, dtype=torch.float64)
    return S.TangentMatrix(prob, prob.jacobian_bsr(u, torch.float32))


def galerkin_levels(A0, dims_list):
    As = [A0]
    for l in range(len(dims_list) - 1):
        P = interp_3d(dims_list[l])
        assert P.shape[1] == 3 * int(np.prod(dims_list[l + 1]))
        As.append((P.T @ As[-1] @ P).tocsr())
    return As


@pytest.mark.parametrize("cells,full", [((4, 3, 5), True), ((8, 6, 7), True), ((4, 3, 5), False), ((8, 6, 7), False)])
def test_galerkin_levels_vs_scipy(jfa, cells, full):
    say("mg galerkin levels", cells, "full" if full else "half")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    prob = box_problem(jfa, cells)
    A = tangent(jfa, prob)
    mg = MG.Multigrid(MG.lattice(prob), fine_full=full)
    mask = MG.dirichlet_mask(prob)
    mg.setup(A, mask)
    dims = mg.levels()
    assert dims[0] == tuple(c + 1 for c in cells) and 3 * np.prod(dims[-1]) <= 88
    m = sp.diags(mask.double().cpu().numpy())
    ref0 = (m @ A.to_scipy() @ m).toarray()
    G0 = stencil_to_csr(mg.stencil(0).cpu().numpy(), dims[0])
    if full:
        np.testing.assert_array_equal(G0.toarray(), ref0)
    else:   # upper blocks copied, lower blocks the neighbours' transposed: the tangent's FP32 asymmetry
        np.testing.assert_array_equal(np.triu(G0.toarray(), 1) != 0, np.triu(ref0, 1) != 0)
        np.testing.assert_allclose(G0.toarray(), ref0, rtol=0, atol=2e-6 * np.abs(ref0).max())
    As = galerkin_levels(G0, dims)   # the coarse levels are Galerkin products of the stored fine level
    for l in range(1, len(dims)):
        G = stencil_to_csr(mg.stencil(l).cpu().numpy(), dims[l])
        ref = As[l].toarray()
        np.testing.assert_allclose(G.toarray(), ref, rtol=0, atol=2e-5 * np.abs(ref).max(), err_msg="level %d" % l)


def test_coarse_operators_annihilate_translations(jfa):
    say("mg translations")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    prob = box_problem(jfa, (6, 5, 7), material="lin", dirichlet=False)
    A = tangent(jfa, prob, scale=0.0)
    mg = MG.Multigrid(MG.lattice(prob))
    mg.setup(A, torch.ones(prob.N, device=DEV))
    for l, d in enumerate(mg.levels()):
        G = stencil_to_csr(mg.stencil(l).cpu().numpy(), d)
        n = int(np.prod(d))
        scale = abs(G).sum(axis=1).max()
        for k in range(3):
            t = np.zeros(3 * n)
            t[k::3] = 1.0
            assert np.abs(G @ t).max() <= 1e-5 * scale, (l, k)


def block_dinv(A, unit=None):
    """Inverses of the 3x3 diagonal blocks; a block whose determinant is not positive
    (an indefinite tangent: a large random displacement) by its diagonal, as the
    set-up kernel does (det <= 1e-12 max|diag|^3 -> 1 / positive diagonal entries)."""
    n = A.shape[0] // 3
    A = A.tocsr()
    D = []
    for i in range(n):
        B = A[3 * i:3 * i + 3, 3 * i:3 * i + 3].toarray()
        if unit is not None:
            B += np.diag(1.0 - unit[3 * i:3 * i + 3])
        sc = np.abs(np.diag(B)).max()
        if np.linalg.det(B) > 1e-12 * sc ** 3:
            D.append(np.linalg.inv(B))
        else:
            D.append(np.diag([1.0 / v if v > 0 else 0.0 for v in np.diag(B)]))
    return sp.block_diag(D).tocsr()


def smoother_ref(A, D, r, e, kind, degree, param, lam):
    """`degree` steps of Chebyshev (interval [lam/param, lam]) or damped Jacobi (omega = param)
    for A e = r preconditioned by D, from e."""
    res = r - A @ e
    if kind == "jacobi":
        for _ in range(degree):
            e = e + param * (D @ res)
            res = r - A @ e
        return e
    lo, hi = lam / param, lam
    theta, delta = 0.5 * (hi + lo), 0.5 * (hi - lo)
    sigma = theta / delta
    rho = 1.0 / sigma
    d = (D @ res) / theta
    for k in range(degree):
        e = e + d
        if k == degree - 1:
            break
        res = res - A @ d
        rho_new = 1.0 / (2.0 * sigma - rho)
        d = rho_new * rho * d + 2.0 * rho_new / delta * (D @ res)
        rho = rho_new
    return e


def vcycle_ref(As, dims, mask, r, kind, degree, param, lams, A_smooth=None):
    """One V-cycle with scipy matrices; A_smooth (optional, per level, None entries
    for As[l]) are the operators the levels' smoothing and restricted residuals use
    (block inverses, bounds and the coarser levels still come from As)."""
    L = len(As)
    Dinv = [block_dinv(As[l], mask if l == 0 else None) for l in range(L - 1)]
    Ps = [interp_3d(dims[l]) for l in range(L - 1)]
    Ac = As[-1].toarray()
    Ainv = np.linalg.pinv(0.5 * (Ac + Ac.T))

    def V(l, rl):
        if l == L - 1:
            return Ainv @ rl
        A = A_smooth[l] if (A_smooth is not None and l < len(A_smooth) and A_smooth[l] is not None) else As[l]
        e = smoother_ref(A, Dinv[l], rl, np.zeros_like(rl), kind, degree, param, lams[l])
        ec = V(l + 1, Ps[l].T @ (rl - A @ e))
        e = e + (mask if l == 0 else 1.0) * (Ps[l] @ ec)
        return smoother_ref(A, Dinv[l], rl, e, kind, degree, param, lams[l])

    return V(0, r), Dinv


@pytest.mark.parametrize("kind,degree,param,full", [("chebyshev", 2, 8.0, False), ("chebyshev", 2, 8.0, True),
                                                    ("chebyshev", 3, 4.0, False), ("jacobi", 2, 0.5, True)])
def test_vcycle_vs_scipy(jfa, kind, degree, param, full):
    say("mg vcycle", kind, degree, param, "full" if full else "half")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    prob = box_problem(jfa, (6, 5, 7))
    A = tangent(jfa, prob)
    mg = MG.Multigrid(MG.lattice(prob), kind, degree, param, fine_full=full)
    mask = MG.dirichlet_mask(prob)
    mg.setup(A, mask)
    dims = mg.levels()
    m = mask.double().cpu().numpy()
    As = galerkin_levels(stencil_to_csr(mg.stencil(0).cpu().numpy(), dims[0]), dims)
    # the Gershgorin bounds of Dinv A, level by level
    Dinv = [block_dinv(As[l], m if l == 0 else None) for l in range(len(As) - 1)]
    lams = [abs(Dinv[l] @ As[l]).sum(axis=1).max() for l in range(len(As) - 1)]
    np.testing.assert_allclose(mg.bounds(), lams, rtol=1e-5)
    g = torch.Generator(device=DEV).manual_seed(11)
    r = torch.randn(prob.N, generator=g, device=DEV) * mask
    z = mg.apply(r).double().cpu().numpy()
    zr, _ = vcycle_ref(As, dims, m, r.double().cpu().numpy(), kind, degree, param, lams)
    np.testing.assert_allclose(z, zr, rtol=0, atol=1e-4 * np.abs(zr).max())


def round_f16(A):
    """The fine stencil as the FP16 smoothing passes see it: scaled by the power of two
    that puts max |A| in [2^14, 2^15), rounded to FP16 (nearest even), scaled back."""
    a = A.tocsr().copy()
    _, e = np.frexp(np.abs(a.data).max())
    sc = 2.0 ** (14 - (int(e) - 1))
    a.data = (a.data * sc).astype(np.float32).astype(np.float16).astype(np.float64) / sc
    return a


def f16_levels(dims):
    """The levels smoothed from an FP16 copy of their stencil: every half-stored one,
    i.e. every level but the dense coarsest (the grid kernels' levels and the
    one-workgroup cycle's alike)."""
    return max(len(dims) - 1, 0)


@pytest.mark.parametrize("cells", [(6, 5, 7), (8, 6, 7), (16, 12, 14)])
def test_vcycle_f16_vs_scipy(jfa, cells):
    """HF_MG_FINE_F16: the smoothing passes and restricted residuals of every level but
    the dense one use FP16-rounded stencils (packed 5 half2 words per block, lower
    blocks read transposed from the neighbour, a power-of-two scale per level), the
    block inverses, bounds and coarser levels the FP32 ones."""
    say("mg vcycle fp16 fine level", cells)
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    prob = box_problem(jfa, cells)
    A = tangent(jfa, prob, scale=0.002)   # small strains: an SPD tangent, whose V-cycle is a contraction
    mask = MG.dirichlet_mask(prob)
    mg16 = MG.Multigrid(MG.lattice(prob), fine_f16=True)
    mg16.setup(A, mask)
    dims = mg16.levels()
    m = mask.double().cpu().numpy()
    A0 = stencil_to_csr(mg16.stencil(0).cpu().numpy(), dims[0])
    As = galerkin_levels(A0, dims)
    Dinv = [block_dinv(As[l], m if l == 0 else None) for l in range(len(As) - 1)]
    lams = [abs(Dinv[l] @ As[l]).sum(axis=1).max() for l in range(len(As) - 1)]
    np.testing.assert_allclose(mg16.bounds(), lams, rtol=1e-5)
    g = torch.Generator(device=DEV).manual_seed(11)
    r = torch.randn(prob.N, generator=g, device=DEV) * mask
    rr = r.double().cpu().numpy()
    z = mg16.apply(r).double().cpu().numpy()
    lb = f16_levels(dims)
    # each smoothed level's own FP32 stencil, rounded as the set-up rounds it
    A16 = [round_f16(stencil_to_csr(mg16.stencil(l).cpu().numpy(), dims[l])) for l in range(lb)]
    say("  %d levels, %d smoothed in FP16" % (len(dims), lb))
    z16, _ = vcycle_ref(As, dims, m, rr, "chebyshev", 2, 8.0, lams, A_smooth=A16)
    z32, _ = vcycle_ref(As, dims, m, rr, "chebyshev", 2, 8.0, lams)
    e16 = np.abs(z - z16).max() / np.abs(z16).max()
    e32 = np.abs(z - z32).max() / np.abs(z32).max()
    say("  vs the FP16-stencil cycle %.2e, vs the FP32 one %.2e" % (e16, e32))
    assert e16 <= 1e-5 and e16 < 0.2 * e32


def test_mg_f16_node_kernels_newton(jfa):
    """A lattice above the lane-kernel size (74,333 nodes): the FP16 fine level in
    the thread-per-node passes.  Mixed Newton with the multigrid reaches the FP64
    residual target with and without it, to the same solution."""
    say("mg fp16 fine level, node kernels")
    lpc = __import__("jax_fem_amd.learned_pc", fromlist=["x"])
    cells = (48, 40, 36)
    E = lpc.heterogeneous_modulus(cells, 6)
    out = {}
    for f16 in (False, True):
        prob = box_problem(jfa, cells, material="lin", E_cell=E)
        u = jfa.solver(prob, {"newton": {"precision": "mixed", "tol": 0.0, "rel_tol": 1e-9,
                                         "linear": {"mg": {"fine_f16": f16, "matrix_free": False}}}})[0]
        info = prob.solve_info
        say("  fine_f16 %s: krylov %s, rel %.2e, fallbacks %d" % (f16, info["krylov_iters"], info["rel"],
                                                                   info["fallbacks"]))
        assert info["rel"] <= 1e-9 and info["fallbacks"] == 0
        out[f16] = (u, int(info["krylov_iters"]))
    rel = (torch.linalg.norm(out[True][0] - out[False][0]) / torch.linalg.norm(out[False][0])).item()
    say("  rel-L2 between the two", rel)
    assert rel <= 1e-7
    assert out[True][1] <= out[False][1] + 3


@pytest.mark.parametrize("cells", [(6, 5, 7), (44, 40, 36)])
def test_cell_operator_matches_stencil(jfa, cells):
    """A built linear-elastic law with a per-cell modulus on a box lattice: the
    stencil assembled from the cell operator (one cell matrix, 48 values by the
    cell's reflection symmetry, per-cell scales) equals the one from the element
    Jacobians, the matrix-free product equals that stencil's to FP32 rounding, and
    the V-cycle with both equals the stencil one (8-lane kernels on the small
    lattice, thread-per-node kernels on the 69,375-node one)."""
    say("mg cell operator vs stencil", cells)
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    lpc = __import__("jax_fem_amd.learned_pc", fromlist=["x"])
    prob = box_problem(jfa, cells, material="lin", E_cell=lpc.heterogeneous_modulus(cells, 3))
    mask = MG.dirichlet_mask(prob)
    mg = MG.Multigrid(MG.lattice(prob), fine_f16=False)
    u = torch.zeros(prob.N, device=DEV, dtype=torch.float64)
    g = torch.Generator(device=DEV).manual_seed(4)
    x = torch.randn(prob.N, generator=g, device=DEV)
    r = torch.randn(prob.N, generator=g, device=DEV) * mask
    # the stencil from the element Jacobians, its product and V-cycle
    L0 = MG.LatticeTangent(prob, u, mg, matrix_free=False)
    assert L0.cell_op is None
    mg.setup(L0, mask)
    y_st = L0.matvec(x)
    S0 = L0.stencil.clone()
    z_st = mg.apply(r)
    # the cell operator, the stencil assembled from it, its product and V-cycle
    L = MG.LatticeTangent(prob, u, mg)
    assert L.cell_op is not None
    K, scale = L.cell_op
    assert K.shape == (576,) and scale.shape == (prob.num_cells,)
    E = torch.as_tensor(lpc.heterogeneous_modulus(cells, 3), dtype=torch.float32, device=DEV)
    assert torch.equal(scale, E)
    es = float((L.stencil - S0).abs().max() / S0.abs().max())
    say("  stencil from the cell operator vs element Jacobians: max rel %.2e" % es)
    assert es <= 2e-6
    mg.setup(L, mask)
    y_mf = mg.cell_matvec(x)
    err = float((y_mf - y_st).abs().max() / y_st.abs().max())
    say("  product: max rel %.2e" % err)
    assert err <= 2e-6
    z_mf = mg.apply(r)
    e2 = float((z_mf - z_st).abs().max() / z_st.abs().max())
    say("  V-cycle: max rel %.2e" % e2)
    assert e2 <= 1e-5
    # not that form: a Neo-Hookean law, a distorted lattice
    neo = box_problem(jfa, (4, 3, 5))
    assert MG.LatticeTangent(neo, torch.zeros(neo.N, device=DEV, dtype=torch.float64),
                             MG.Multigrid(MG.lattice(neo))).cell_op is None
    bent = box_problem(jfa, (4, 3, 5), material="lin")
    bent.fe.points[5, 1] += 0.01
    assert bent._cell_matrix(MG.lattice(bent), bent._box_corners(MG.lattice(bent)), 0.3) is None


def test_mg_matrix_free_newton(jfa):
    """Mixed Newton with the multigrid on a 74,333-node heterogeneous lattice: the
    matrix-free fine level reaches the FP64 target with the stencil one's
    iteration count and solution."""
    say("mg matrix-free newton")
    lpc = __import__("jax_fem_amd.learned_pc", fromlist=["x"])
    cells = (48, 40, 36)
    E = lpc.heterogeneous_modulus(cells, 6)
    out = {}
    for mf in (False, True):
        prob = box_problem(jfa, cells, material="lin", E_cell=E)
        u = jfa.solver(prob, {"newton": {"precision": "mixed", "tol": 0.0, "rel_tol": 1e-9,
                                         "linear": {"mg": {"matrix_free": mf}}}})[0]
        info = prob.solve_info
        say("  matrix_free %s: krylov %s, rel %.2e, fallbacks %d" % (mf, info["krylov_iters"], info["rel"],
                                                                     info["fallbacks"]))
        assert info["rel"] <= 1e-9 and info["fallbacks"] == 0
        assert (getattr(prob._mg, "_cell_op", None) is not None) == mf
        out[mf] = (u, int(info["krylov_iters"]))
    rel = (torch.linalg.norm(out[True][0] - out[False][0]) / torch.linalg.norm(out[False][0])).item()
    say("  rel-L2 between the two", rel)
    assert rel <= 1e-6     # both within the FP64 residual target; the difference is the conditioning's
    assert out[True][1] <= out[False][1] + 2


def test_mg_fcg_converges(jfa):
    say("mg fcg solve")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    lpc = __import__("jax_fem_amd.learned_pc", fromlist=["x"])
    from importlib import import_module
    S = import_module("jax_fem_amd.solver")
    cells = (16, 12, 10)
    prob = box_problem(jfa, cells, material="lin", E_cell=lpc.heterogeneous_modulus(cells, 4))
    A = S.TangentMatrix(prob, prob.jacobian_bsr(torch.zeros(prob.N, device=DEV, dtype=torch.float64),
                                                torch.float32))
    mask = MG.dirichlet_mask(prob)
    g = torch.Generator(device=DEV).manual_seed(2)
    b = torch.randn(prob.N, generator=g, device=DEV) * mask
    mg = MG.Multigrid.for_problem(prob)
    x, info = mg.solve(A, b, rtol=1e-6, atol=0.0, maxiter=200)
    assert info.status == 0, (info.status, info.iterations)
    A64 = A.to_scipy()
    bb = b.double().cpu().numpy()
    rel = np.linalg.norm(bb - A64 @ x.double().cpu().numpy()) / np.linalg.norm(bb)
    assert rel <= 1e-5, rel
    _, jinfo = S.krylov(A, b, None, "pcg", 1e-6, 0.0, 20000, True)
    say("mg iterations", info.iterations, "jacobi-pcg", jinfo.iterations)
    assert info.iterations <= 20 and 8 * info.iterations < jinfo.iterations


@pytest.mark.parametrize("stencil", [False, True])
def test_mg_solve_graph_replays_match_eager(jfa, stencil):
    """The captured poll batch gives the eager loop's iterates bit for bit: on
    capture, on replay after unrelated eager work, with a new right-hand side,
    with another poll length, and after a set-up from another tangent (SELL
    tangent or lattice stencil as the operator)."""
    say("mg fcg graph replay vs eager", "stencil" if stencil else "sell")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    prob = box_problem(jfa, (8, 6, 7))
    mask = MG.dirichlet_mask(prob)
    mg = MG.Multigrid(MG.lattice(prob))
    g = torch.Generator(device=DEV).manual_seed(5)

    def tangent(jfa, prob, seed, scale):   # small strains: the Neo-Hookean tangent stays SPD
        gu = torch.Generator(device=DEV).manual_seed(seed)
        u = scale * torch.randn(prob.N, generator=gu, device=DEV, dtype=torch.float64)
        if stencil:
            return MG.LatticeTangent(prob, u, mg)
        S = __import__("jax_fem_amd.solver", fromlist=["x"])
        return S.TangentMatrix(prob, prob.jacobian_bsr(u, torch.float32))

    A = tangent(jfa, prob, seed=1, scale=0.002)
    mg.setup(A, mask)

    def both(A, b, poll):
        xe, ie = mg.solve(A, b, rtol=1e-7, maxiter=200, poll=poll, mask=mask, rebuild="never", eager=True)
        xg, ig = mg.solve(A, b, rtol=1e-7, maxiter=200, poll=poll, mask=mask, rebuild="never")
        assert ie.status == 0 and ig.status == 0, (ie.status, ig.status)
        assert ie.iterations == ig.iterations, (ie.iterations, ig.iterations)
        assert torch.equal(xe, xg), float((xe - xg).abs().max())
        return xg

    b1 = torch.randn(prob.N, generator=g, device=DEV) * mask
    x1 = both(A, b1, 2)
    junk = torch.randn(1 << 20, device=DEV)   # unrelated eager work between replays
    _ = (junk * 2.0).sum().item()
    torch.cuda.synchronize()
    xr, _ = mg.solve(A, b1, rtol=1e-7, maxiter=200, poll=2, mask=mask, rebuild="never")
    assert torch.equal(xr, x1)
    b2 = torch.randn(prob.N, generator=g, device=DEV) * mask
    both(A, b2, 2)
    both(A, b2, 3)
    A2 = tangent(jfa, prob, seed=2, scale=0.004)
    mg.setup(A2, mask)
    both(A2, b1, 3)


def half_stencil_to_csr(S, dims):
    """A from a half-stored stencil (14, 3, 3, n): slot k couples node n to its upper
    neighbour at offset o = 13 + k; the lower blocks are the upper ones transposed."""
    nx, ny, nz = dims
    n = nx * ny * nz
    idx = np.arange(n)
    x, y, z = idx // (ny * nz), (idx // nz) % ny, idx % nz
    rows, cols, vals = [], [], []
    for k in range(14):
        o = 13 + k
        dx, dy, dz = o // 9 - 1, (o // 3) % 3 - 1, o % 3 - 1
        xx, yy, zz = x + dx, y + dy, z + dz
        ok = (xx >= 0) & (xx < nx) & (yy >= 0) & (yy < ny) & (zz >= 0) & (zz < nz)
        assert np.all(S[k][:, :, ~ok] == 0), "slot %d leaves the lattice but is not zero" % o
        nb = (xx * ny + yy) * nz + zz
        for i in range(3):
            for j in range(3):
                v = S[k, i, j, ok].astype(np.float64)
                rows.append(3 * idx[ok] + i); cols.append(3 * nb[ok] + j); vals.append(v)
                if k:
                    rows.append(3 * nb[ok] + j); cols.append(3 * idx[ok] + i); vals.append(v)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * n, 3 * n))


def test_lattice_stencil_assembly_and_operator(jfa):
    """The stencil assembled from the element Jacobians holds the SELL tangent's
    diagonal and upper blocks bit for bit (same cell-order sums, Dirichlet rows
    kept), and its operator is the row-eliminated tangent."""
    say("mg lattice stencil assembly")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    S_ = __import__("jax_fem_amd.solver", fromlist=["x"])
    cells = (6, 5, 7)
    prob = box_problem(jfa, cells)
    mask = MG.dirichlet_mask(prob)
    mg = MG.Multigrid.for_problem(prob)
    g = torch.Generator(device=DEV).manual_seed(3)
    u = 0.002 * torch.randn(prob.N, generator=g, device=DEV, dtype=torch.float64)
    L = MG.LatticeTangent(prob, u, mg)
    assert prob._box_corners(mg.grid) is not None   # a box_mesh: the coordinate-driven kernel ran
    # the kernel that reads the pattern's contribution lists gives the same stencil bit for bit
    _lib = __import__("jax_fem_amd._lib", fromlist=["x"])
    S_gen = torch.empty_like(L.stencil)
    _lib.call("hf_mg_assemble_hex8_f32", mg.handle, prob.pattern.handle, _lib.ptr(prob._dev()[1]), prob.num_cells,
              None, _lib.ptr(prob._bufs["elem_jac_upper32"]), _lib.ptr(S_gen), _lib.stream())
    assert torch.equal(S_gen, L.stencil)
    dims = tuple(c + 1 for c in cells)
    n = int(np.prod(dims))
    S = L.stencil.reshape(14, 3, 3, n).cpu().numpy()
    raw = S_.TangentMatrix(prob, prob.jacobian_bsr(u, torch.float32, bc=False)).to_scipy().tocsr()
    H = half_stencil_to_csr(S, dims)
    upper = lambda M: sp.triu(sp.csr_matrix(M), k=0).toarray()
    # block-upper part: node-pair blocks (n, m) with m >= n in lattice order
    nodes = np.arange(3 * n) // 3
    keep = nodes[:, None] <= nodes[None, :]
    np.testing.assert_array_equal(np.where(keep, H.toarray(), 0), np.where(keep, raw.toarray(), 0))
    mg.setup(L, mask)
    A = S_.TangentMatrix(prob, prob.jacobian_bsr(u, torch.float32))
    x = torch.randn(prob.N, generator=g, device=DEV)
    x[mask == 0] = torch.randn(int((mask == 0).sum()), generator=g, device=DEV)
    y_st, y_sell = L.matvec(x), A.matvec(x)
    err = float((y_st - y_sell).abs().max() / y_sell.abs().max())
    say("stencil operator vs row-eliminated SELL: max rel", err)
    assert err <= 1e-5
    # the hierarchy built from the stencil equals the one built from the SELL tangent
    mg2 = MG.Multigrid(MG.lattice(prob))
    mg2.setup(A, mask)
    for l in range(1, len(mg.levels())):
        a, b = mg.stencil(l), mg2.stencil(l)
        np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=0,
                                   atol=1e-5 * float(b.abs().max()), err_msg="level %d" % l)


@pytest.mark.parametrize("stencil", [True, False])
def test_newton_with_mg_matches_fp64(jfa, stencil):
    say("mg newton", "lattice stencil" if stencil else "SELL tangent")
    cells = (10, 8, 8)
    prob = box_problem(jfa, cells)
    u64 = jfa.solver(prob, {"newton": {"tol": 0.0, "rel_tol": 1e-10,
                                       "linear": {"pcg": {"tol": 1e-12, "atol": 1e-14}}}})[0]
    prob2 = box_problem(jfa, cells)
    u = jfa.solver(prob2, {"newton": {"precision": "mixed", "tol": 0.0, "rel_tol": 1e-10,
                                      "linear": {"mg": {"stencil": stencil}}}})[0]
    assert prob2.solve_info["rel"] <= 1e-10
    assert prob2.solve_info["fallbacks"] == 0
    rel = (torch.linalg.norm(u - u64) / torch.linalg.norm(u64)).item()
    say("newton mg rel-L2", rel, "krylov", prob2.solve_info["krylov_iters"])
    assert rel <= 1e-6



def test_linear_law_keeps_one_tangent_per_solve(jfa):
    """A linear law's tangent does not depend on u: the mixed-precision Newton
    refinement assembles it once per solve and still reaches the FP64 solution."""
    say("linear law tangent reuse")
    cells = (10, 8, 8)
    E = np.exp(np.random.default_rng(4).normal(0.0, 1.0, int(np.prod(cells))))
    prob = box_problem(jfa, cells, material="lin", E_cell=E)
    u64 = jfa.solver(prob, {"newton": {"tol": 0.0, "rel_tol": 1e-10,
                                       "linear": {"pcg": {"tol": 1e-13, "atol": 0.0}}}})[0]
    prob2 = box_problem(jfa, cells, material="lin", E_cell=E)
    v0 = getattr(prob2, "_stencil_version", 0)
    u = jfa.solver(prob2, {"newton": {"precision": "mixed", "tol": 0.0, "rel_tol": 1e-10, "linear": {"mg": {}}}})[0]
    assert prob2.solve_info["newton_iters"] >= 2      # FP32 solves refined in FP64
    assert prob2._stencil_version - v0 == 1
    assert prob2.solve_info["rel"] <= 1e-10
    rel = (torch.linalg.norm(u - u64) / torch.linalg.norm(u64)).item()
    assert rel <= 1e-8, rel

def test_mg_adjoint_matches_fp64_pcg(jfa):
    say("mg adjoint")
    from importlib import import_module
    S = import_module("jax_fem_amd.solver")
    lpc = import_module("jax_fem_amd.learned_pc")
    cells = (9, 7, 8)
    prob = box_problem(jfa, cells, material="lin", E_cell=lpc.heterogeneous_modulus(cells, 5))
    u = jfa.solver(prob, {"newton": {"tol": 0.0, "rel_tol": 1e-10, "linear": {"pcg": {"tol": 1e-12}}}})[0].reshape(-1)
    g = torch.Generator(device=DEV).manual_seed(3)
    v = torch.randn(prob.N, generator=g, device=DEV, dtype=torch.float64)
    lam_ref, _ = S.adjoint_solve(prob, u, v, {"pcg": {"tol": 1e-13, "atol": 0.0}})
    lam, _ = S.adjoint_solve(prob, u, v, {"mg": {"tol": 1e-12}})
    info = prob.adjoint_info
    rel = (torch.linalg.norm(lam - lam_ref) / torch.linalg.norm(lam_ref)).item()
    say("mg adjoint rel", rel, info)
    assert info["rel"] <= 1e-12 and info["refinements"] <= 6
    assert rel <= 1e-9


def _hex27_problem(jfa, cells, material):
    left = lambda p: np.isclose(p[0], 0.0, atol=1e-7)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-7)

    class P(jfa.Problem):
        def get_tensor_map(self):
            return material

    bc = [[left] * 3 + [right], [0, 1, 2, 0], [lambda p: 0.0] * 3 + [lambda p: 0.05]]
    return P(jfa.box_mesh_hex27(*cells), vec=3, dim=3, ele_type="HEX27", quadrature_order=4, dirichlet_bc_info=bc)


def test_hex27_level1_is_the_hex8_operator(jfa):
    """P^T A P of the Q2 stiffness with P the interpolation from every second lattice
    node (the cell vertices) is the Q1 (HEX8) stiffness of the same cells on a box:
    Q1 functions are Q2 functions, and both Gauss rules integrate their form exactly."""
    say("mg hex27 level 1 = hex8 operator")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    cells = (4, 3, 5)
    mat = jfa.LinearElastic(E=10.0, nu=0.3)
    p27 = _hex27_problem(jfa, cells, mat)
    A27 = tangent(jfa, p27, scale=0.0)
    mg = MG.Multigrid(MG.lattice(p27), fine_sell=True)
    mg.setup(A27, MG.dirichlet_mask(p27))
    dims = mg.levels()
    assert dims[0] == tuple(2 * c + 1 for c in cells) and dims[1] == tuple(c + 1 for c in cells)
    G1 = stencil_to_csr(mg.stencil(1).cpu().numpy(), dims[1]).toarray()
    p8 = box_problem(jfa, cells, material="lin")
    p8.material = mat

    class P8(jfa.Problem):
        def get_tensor_map(self):
            return mat

    left = lambda p: np.isclose(p[0], 0.0, atol=1e-7)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-7)
    p8 = P8(jfa.box_mesh(*cells), vec=3, dim=3,
            dirichlet_bc_info=[[left] * 3 + [right], [0, 1, 2, 0], [lambda p: 0.0] * 3 + [lambda p: 0.05]])
    A8 = tangent(jfa, p8, scale=0.0).to_scipy().toarray()
    free = MG.dirichlet_mask(p8).cpu().numpy() > 0
    ref = A8[np.ix_(free, free)]
    np.testing.assert_allclose(G1[np.ix_(free, free)], ref, rtol=0, atol=2e-5 * np.abs(ref).max())
    # and level 1 is the Galerkin product of the masked SELL tangent (scipy)
    m = sp.diags(MG.dirichlet_mask(p27).double().cpu().numpy())
    As = galerkin_levels((m @ A27.to_scipy() @ m).tocsr(), dims)
    for l in range(1, len(dims)):
        G = stencil_to_csr(mg.stencil(l).cpu().numpy(), dims[l]).toarray()
        r = As[l].toarray()
        np.testing.assert_allclose(G, r, rtol=0, atol=2e-5 * np.abs(r).max(), err_msg="level %d" % l)


@pytest.mark.parametrize("f16", [False, True])
def test_hex27_vcycle_and_newton(jfa, f16):
    """HEX27 (the fine level is the SELL tangent): the V-cycle against scipy, with the
    fine smoothing products in FP32 or from the FP16 copy of the SELL values (the
    reference then smooths with M round16(A) M, the row-eliminated tangent rounded
    with the same power-of-two scale); Newton through it against the FP64 path."""
    say("mg hex27 vcycle + newton", "fp16 values" if f16 else "fp32 values")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    cells = (3, 2, 3)
    p27 = _hex27_problem(jfa, cells, jfa.NeoHookean(E=10.0, nu=0.3))
    A = tangent(jfa, p27, scale=0.01)
    mg = MG.Multigrid(MG.lattice(p27), fine_sell=True, fine_f16=f16)
    mask = MG.dirichlet_mask(p27)
    mg.setup(A, mask)
    dims = mg.levels()
    m = mask.double().cpu().numpy()
    As = galerkin_levels((sp.diags(m) @ A.to_scipy() @ sp.diags(m)).tocsr(), dims)
    Dinv = [block_dinv(As[l], m if l == 0 else None) for l in range(len(As) - 1)]
    lams = [abs(Dinv[l] @ As[l]).sum(axis=1).max() for l in range(len(As) - 1)]
    np.testing.assert_allclose(mg.bounds(), lams, rtol=1e-5)
    g = torch.Generator(device=DEV).manual_seed(4)
    r = torch.randn(p27.N, generator=g, device=DEV) * mask
    z = mg.apply(r).double().cpu().numpy()
    A16 = [(sp.diags(m) @ round_f16(A.to_scipy()) @ sp.diags(m)).tocsr()] if f16 else None
    zr, _ = vcycle_ref(As, dims, m, r.double().cpu().numpy(), "chebyshev", 2, 8.0, lams, A_smooth=A16)
    err = np.abs(z - zr).max() / np.abs(zr).max()
    say("  V-cycle vs scipy: max rel %.2e" % err)
    assert err <= (1e-5 if f16 else 1e-4)
    # Newton through the multigrid vs the FP64 path
    u64 = jfa.solver(p27, {"newton": {"tol": 0.0, "rel_tol": 1e-10, "linear": {"pcg": {"tol": 1e-12, "atol": 1e-14}}}})[0]
    p2 = _hex27_problem(jfa, cells, jfa.NeoHookean(E=10.0, nu=0.3))
    u = jfa.solver(p2, {"newton": {"precision": "mixed", "tol": 0.0, "rel_tol": 1e-10, "linear": {"mg": {}}}})[0]
    assert p2.solve_info["rel"] <= 1e-10 and p2.solve_info["fallbacks"] == 0
    rel = (torch.linalg.norm(u - u64) / torch.linalg.norm(u64)).item()
    say("hex27 newton mg rel-L2", rel, "krylov", p2.solve_info["krylov_iters"])
    assert rel <= 1e-6


def test_setup_rejects_non_lattice_pattern(jfa):
    say("mg non-lattice")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    cells = (3, 3, 3)
    mesh = jfa.box_mesh(*cells)
    perm = np.random.default_rng(0).permutation(len(mesh.points))
    inv = np.argsort(perm)
    mesh2 = jfa.Mesh(mesh.points[perm], inv[mesh.cells], "HEX8")   # same body, nodes relabelled

    class P(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=1.0, nu=0.3)

    prob = P(mesh2, vec=3, dim=3)
    A = tangent(jfa, prob, scale=0.0)
    mg = MG.Multigrid((4, 4, 4))
    with pytest.raises(jfa.HipFemError, match="lattice"):
        mg.setup(A, torch.ones(prob.N, device=DEV))
    with pytest.raises(jfa.HipFemError, match="2 nodes"):
        MG.Multigrid((1, 4, 4))


@pytest.mark.parametrize("cells", [(2, 2, 2), (3, 3, 3), (5, 4, 6)])
def test_mg_solve_edge_cases(jfa, cells):
    """The solve loop at its edges: a one-level hierarchy (the dense FP64 solve
    alone, 2x2x2 cells), a fine level over the dense one only (3x3x3), a zero
    right-hand side (done at the start, x untouched), maxiter = 0 (status 1), and
    poll batches longer than the solve needs: every launch after convergence
    exits at once, so x and the iteration count equal those of poll = 1 bit for bit,
    replayed or eager."""
    say("mg edge cases", cells)
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    prob = box_problem(jfa, cells)
    A = tangent(jfa, prob, scale=0.002)
    mask = MG.dirichlet_mask(prob)
    mg = MG.Multigrid(MG.lattice(prob))
    nlev = len(mg.levels())
    assert nlev == {(2, 2, 2): 1, (3, 3, 3): 2}.get(cells, nlev) and nlev >= 1
    mg.setup(A, mask)
    g = torch.Generator(device=DEV).manual_seed(5)
    b = torch.randn(prob.N, generator=g, device=DEV) * mask
    A64 = A.to_scipy()
    bb = b.double().cpu().numpy()
    x1, i1 = mg.solve(A, b, rtol=1e-6, atol=0.0, maxiter=100, rebuild="never")
    assert i1.status == 0, (i1.status, i1.iterations)
    rel = np.linalg.norm(bb - A64 @ x1.double().cpu().numpy()) / np.linalg.norm(bb)
    assert rel <= 1e-5, rel
    if nlev == 1:
        assert i1.iterations <= 2, i1.iterations   # the preconditioner is the exact inverse
    for poll, eager in ((3, False), (3, True), (7, False)):
        xp, ip = mg.solve(A, b, rtol=1e-6, atol=0.0, maxiter=100, rebuild="never", poll=poll, eager=eager)
        assert (ip.status, ip.iterations) == (i1.status, i1.iterations), (poll, eager, ip.iterations)
        assert torch.equal(xp, x1), (poll, eager)
    z = torch.zeros_like(b)
    xz, iz = mg.solve(A, z, x0=torch.zeros_like(b), rtol=1e-6, atol=0.0, maxiter=100, rebuild="never")
    assert iz.status == 0 and iz.iterations == 0 and bool((xz == 0).all())
    xm, im = mg.solve(A, b, rtol=1e-6, atol=0.0, maxiter=0, rebuild="never")
    assert im.status == 1 and im.iterations == 0 and bool((xm == 0).all())


@pytest.mark.parametrize("load", [False, True])
def test_box_residual_gather_bit_exact(jfa, load):
    """On a box_mesh the residual is gathered by lattice coordinates
    (hf_gather_residual_box_f64); it equals the pattern-list gather bit for bit,
    with and without an external load vector."""
    say("box residual gather", "load" if load else "no load")
    _lib = __import__("jax_fem_amd._lib", fromlist=["x"])
    mat = jfa.NeoHookean(E=10.0, nu=0.3)

    class P(jfa.Problem):
        def get_tensor_map(self):
            return mat

        def get_surface_maps(self):
            return [lambda u, x: np.array([0.0, 0.3, -1.0])] if load else []

    right = lambda p: np.isclose(p[0], 1.0, atol=1e-7)
    prob = P(jfa.box_mesh(5, 4, 6), vec=3, dim=3, location_fns=[right] if load else None)
    assert prob._gather_box() is not None
    g = torch.Generator(device=DEV).manual_seed(9)
    u = 0.01 * torch.randn(prob.N, generator=g, device=DEV, dtype=torch.float64)
    R_box = prob.residual_vec(u)
    fext = prob._external()
    assert (fext is not None) == load
    R_gen = torch.empty_like(R_box)
    _lib.call("hf_gather_residual_f64", prob.pattern.handle, _lib.ptr(prob._bufs["elem_res"]), _lib.ptr(fext),
              _lib.ptr(R_gen), _lib.stream())
    assert torch.equal(R_box, R_gen)


@pytest.mark.parametrize("kind", ["stencil", "hex27"])
def test_split_direction_scalars_match_fused(jfa, kind):
    """The one-workgroup direction-scalar kernel (the form operators of more than
    2048 workgroups take: c2, c4) against the fused form on a small lattice: same
    status and iteration count, iterates equal to FP32 rounding, FP64 residual at
    the target; a breakdown (status 2: the hierarchy of A preconditioning -A, so
    p.q < 0) and a non-finite product (status 3: one value set to 3e38) are reported
    by both forms."""
    say("mg split vs fused direction scalars", kind)
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    S = __import__("jax_fem_amd.solver", fromlist=["x"])
    if kind == "hex27":
        prob = _hex27_problem(jfa, (3, 2, 3), jfa.NeoHookean(E=10.0, nu=0.3))
        mg = MG.Multigrid(MG.lattice(prob), fine_sell=True)
        A = tangent(jfa, prob, scale=0.005)
    else:
        prob = box_problem(jfa, (9, 6, 7))
        mg = MG.Multigrid(MG.lattice(prob))
        g0 = torch.Generator(device=DEV).manual_seed(3)
        u = 0.002 * torch.randn(prob.N, generator=g0, device=DEV, dtype=torch.float64)
        A = MG.LatticeTangent(prob, u, mg)
    mask = MG.dirichlet_mask(prob)
    mg.setup(A, mask)
    g = torch.Generator(device=DEV).manual_seed(6)
    b = torch.randn(prob.N, generator=g, device=DEV) * mask
    out = {}
    for split in (False, True):
        for eager in (False, True):
            x, info = mg.solve(A, b, rtol=1e-6, atol=0.0, maxiter=200, rebuild="never", eager=eager,
                               split_scalars=split)
            out[(split, eager)] = (x, info.status, info.iterations)
    for key, (x, st, it) in out.items():
        assert st == 0, (key, st)
    xf, _, itf = out[(False, False)]
    xs, _, its = out[(True, False)]
    assert abs(itf - its) <= 1, (itf, its)
    assert torch.equal(out[(True, True)][0], xs)          # replayed = eager, split form too
    rel = (torch.linalg.norm(xs - xf) / torch.linalg.norm(xf)).item()
    assert rel <= 1e-5, rel
    A64 = (A.to_sell() if kind == "stencil" else A).to_scipy()
    bb = b.double().cpu().numpy()
    res = np.linalg.norm(bb - A64 @ xs.double().cpu().numpy()) / np.linalg.norm(bb)
    assert res <= 1e-5, res
    # breakdown and non-finite values, both forms (SELL operator)
    As = A.to_sell() if kind == "stencil" else A
    An = S.TangentMatrix(prob, -As.vals)
    big = As.vals.clone()
    big[torch.argmax(big.abs())] = 3e38
    Ab = S.TangentMatrix(prob, big)
    for split in (False, True):
        # a failed solve marks the hierarchy stale: set it up from A again before each
        mg.setup(As, mask)
        _, i2 = mg.solve(An, b, rtol=1e-6, atol=0.0, maxiter=50, rebuild="never", split_scalars=split)
        assert i2.status == 2, (split, i2.status, i2.iterations)
        mg.setup(As, mask)
        # right-hand side scaled so that A z overflows FP32 at that entry
        _, i3 = mg.solve(Ab, 1e6 * b, rtol=1e-6, atol=0.0, maxiter=50, rebuild="never", split_scalars=split)
        assert i3.status == 3, (split, i3.status, i3.iterations)
