"""HEX27 (3x3x3 Gauss, quadrature_order 4) through the C ABI vs the FP64 oracle:
cell residual, FP64 element Jacobian (vector ALU), FP32 element Jacobian (MFMA),
parameter VJP, a Neo-Hookean Newton solve, and the adjoint w.r.t. per-cell mu
against the oracle's implicit_vjp and central differences.

Tolerances as test_gpu_parity: FP64 kernels 1e-11 relative, FP32 2e-5 relative,
solutions vs the oracle 1e-8, gradients 1e-6 vs the oracle and 1e-5 vs FD."""
import sys

import numpy as np
import pytest
import torch

from oracle import fe as ofe
from oracle import kernels as ok
from oracle import solver as osv

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def iscl(a, b):
    return np.abs(a - b) < 1e-5


def _hex27_box(n=(2, 2, 2), seed=0, jitter=0.03):
    """HEX27 box with interior nodes jittered (boundary planes stay planar)."""
    pts, cells = ofe.box_mesh_hex27(*n)
    rng = np.random.default_rng(seed)
    interior = np.all((pts > 1e-9) & (pts < 1 - 1e-9), axis=1)
    pts = pts.copy()
    pts[interior] += jitter * rng.uniform(-1, 1, (interior.sum(), 3)) / max(n)
    return pts, cells


def _mats(jfa, kind, param):
    kw = dict(E=10.0, nu=0.3, param=param)
    if param == "mu":
        kw["kappa"] = 8.0
    if kind == "linear_elastic":
        return jfa.LinearElastic(**kw), ok.Material(kind, E=10.0, nu=0.3, param=param)
    return jfa.NeoHookean(**kw), ok.Material(kind, E=10.0, nu=0.3, param=param, kappa=8.0)


def _problem(jfa, pts, cells, mat, params=None, dirichlet=None):
    class P(jfa.Problem):
        def get_tensor_map(self):
            return mat

        def set_params(self, p):
            self.internal_vars = [p[:, None].expand(-1, 27) if torch.is_tensor(p) else p]

    prob = P(jfa.Mesh(pts, cells, "HEX27"), vec=3, dim=3, ele_type="HEX27", quadrature_order=4,
             dirichlet_bc_info=dirichlet)
    if params is not None:
        prob.internal_vars = [params]
    return prob


@pytest.mark.parametrize("kind,param", [("linear_elastic", None), ("neo_hookean", None), ("neo_hookean", "mu"),
                                        ("linear_elastic", "simp")])
def test_hex27_cell_kernels_vs_oracle(jfa, kind, param):
    say("hex27 cell kernels", kind, param)
    pts, cells = _hex27_box()
    mat, omat = _mats(jfa, kind, param)
    rng = np.random.default_rng(1)
    prm = rng.uniform(0.3, 1.0, (len(cells), 27)) if param else None
    prob = _problem(jfa, pts, cells, mat, params=prm)
    oprob = osv.OracleProblem(pts, cells, 3, omat, ele_type="HEX27", quadrature_order=4, params=prm)
    assert prob.fe.num_quads == 27 and oprob.JxW.shape == (len(cells), 27)
    u = 0.03 * rng.standard_normal(oprob.N)
    R = prob.residual_vec(torch.tensor(u, device=DEV)).cpu().numpy()
    Ro = oprob.compute_residual(u)
    np.testing.assert_allclose(R, Ro, rtol=1e-11, atol=1e-11 * np.abs(Ro).max())
    Ko = oprob.element_jacobians(u)
    K64 = prob.element_jacobians(torch.tensor(u, device=DEV)).cpu().numpy()
    np.testing.assert_allclose(K64, Ko, rtol=1e-11, atol=1e-11 * np.abs(Ko).max())
    K32 = prob.element_jacobians(torch.tensor(u, device=DEV), torch.float32).cpu().numpy()
    np.testing.assert_allclose(K32, Ko, rtol=2e-5, atol=2e-5 * np.abs(Ko).max())
    if param:
        lam = rng.standard_normal(oprob.N)
        from importlib import import_module
        g = import_module("jax_fem_amd.solver").param_vjp(prob, torch.tensor(u, device=DEV),
                                                          torch.tensor(lam, device=DEV)).cpu().numpy()
        go = -ok.cell_param_vjp(omat, oprob.cell_u(u), oprob.grads, oprob.JxW, prm, lam.reshape(-1, 3)[cells])
        np.testing.assert_allclose(g, go, rtol=1e-10, atol=1e-12 * np.abs(go).max())


def test_hex27_default_quadrature_runs_pointwise(jfa):
    """HEX27 with the reference's default quadrature order 10 (6^3 points), for which
    no built kernel exists: the law runs on the pointwise path; residual and element
    Jacobian against the oracle at the same rule."""
    say("hex27 default quadrature, pointwise path")
    pts, cells = ofe.box_mesh_hex27(2, 1, 1)
    mat, omat = _mats(jfa, "neo_hookean", None)

    class P(jfa.Problem):
        def get_tensor_map(self):
            return mat

    prob = P(jfa.Mesh(pts, cells, "HEX27"), vec=3, dim=3, ele_type="HEX27")   # default order 10: 6^3 points
    assert prob.fe.num_quads == 216 and prob._plan()["law"] == "generic"
    op = osv.OracleProblem(pts, cells, 3, omat, ele_type="HEX27", quadrature_order=10)
    u = 0.03 * np.random.default_rng(4).standard_normal(op.N)
    R = prob.residual_vec(torch.tensor(u, device=DEV)).cpu().numpy()
    Ro = op.compute_residual(u)
    np.testing.assert_allclose(R, Ro, rtol=1e-11, atol=1e-11 * np.abs(Ro).max())
    K = prob.element_jacobians(torch.tensor(u, device=DEV)).cpu().numpy()
    Ko = op.element_jacobians(u)
    np.testing.assert_allclose(K, Ko, rtol=1e-11, atol=1e-11 * np.abs(Ko).max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hex27_newton_vs_oracle(jfa, precision):
    say("hex27 newton", precision)
    pts, cells = _hex27_box((3, 2, 2), 4)
    mat, omat = _mats(jfa, "neo_hookean", None)
    left = lambda p: iscl(p[0], 0.0)
    right = lambda p: iscl(p[0], 1.0)
    dirichlet = [[left, left, left, right], [0, 1, 2, 0], [lambda p: 0.0] * 3 + [lambda p: 0.05]]
    prob = _problem(jfa, pts, cells, mat, dirichlet=dirichlet)
    oprob = osv.OracleProblem(pts, cells, 3, omat, ele_type="HEX27", quadrature_order=4,
                              dirichlet_bc_info=dirichlet)
    uo, _ = osv.newton_solve(oprob, rel_tol=1e-12, tol=0.0)
    sol = jfa.solver(prob, {"newton": {"precision": precision, "tol": 0.0, "rel_tol": 1e-10,
                                       "linear": {"pcg": {"tol": 1e-12, "atol": 1e-14}}}})
    u = sol[0].reshape(-1).cpu().numpy()
    assert prob.solve_info["rel"] <= 1e-10
    np.testing.assert_allclose(u, uo, rtol=0, atol=1e-8 * np.abs(uo).max())


def test_hex27_adjoint_mu_vs_oracle_and_fd(jfa):
    say("hex27 adjoint mu")
    pts, cells = _hex27_box((2, 2, 2), 7)
    rng = np.random.default_rng(9)
    E, nu = 10.0, 0.3
    mu0 = E / (2 * (1 + nu))
    mu = mu0 * (1.0 + 0.1 * rng.standard_normal(len(cells)))
    mat = jfa.NeoHookean(E=E, nu=nu, param="mu", kappa=E / (3 * (1 - 2 * nu)))
    left = lambda p: iscl(p[0], 0.0)
    right = lambda p: iscl(p[0], 1.0)
    dirichlet = [[left, left, left, right], [0, 1, 2, 0], [lambda p: 0.0] * 3 + [lambda p: 0.05]]
    prob = _problem(jfa, pts, cells, mat, dirichlet=dirichlet)
    th = torch.tensor(mu, device=DEV, requires_grad=True)
    opts = {"newton": {"tol": 0.0, "rel_tol": 1e-12, "linear": {"pcg": {"tol": 1e-13, "atol": 1e-16}}}}
    fwd = jfa.ad_wrapper(prob, opts, {"bicgstab": {"tol": 1e-13, "atol": 1e-16}})
    u = fwd(th)[0]
    obj = (u ** 2).sum()          # not self-adjoint: lambda_I is far from R_I(u*) = 0
    obj.backward()
    g = th.grad.cpu().numpy()
    # oracle
    om = ok.Material("neo_hookean", E=E, nu=nu, param="mu", kappa=E / (3 * (1 - 2 * nu)))
    op = osv.OracleProblem(pts, cells, 3, om, ele_type="HEX27", quadrature_order=4, dirichlet_bc_info=dirichlet,
                           params=np.repeat(mu[:, None], 27, axis=1))
    uo, _ = osv.newton_solve(op, rel_tol=1e-12, tol=0.0)
    np.testing.assert_allclose(u.detach().cpu().numpy().ravel(), uo, rtol=0, atol=1e-8 * np.abs(uo).max())
    gq, _, _ = osv.implicit_vjp(op, uo, 2.0 * uo)
    go = gq.sum(axis=1)
    np.testing.assert_allclose(g, go, rtol=1e-6, atol=1e-9 * np.abs(go).max())
    for c in rng.choice(len(cells), 3, replace=False):
        vals = []
        for s in (1, -1):
            t2 = th.detach().clone()
            t2[c] += s * 1e-4 * mu0
            prob.set_params(t2)
            uu = jfa.solver(prob, opts)[0]
            vals.append(float((uu ** 2).sum()))
        fd = (vals[0] - vals[1]) / (2e-4 * mu0)
        assert abs(fd - g[c]) <= 1e-5 * abs(g[c]) + 1e-12, (c, fd, g[c])


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_hex27_row_assembly_bit_exact(jfa, dtype):
    """hf_assemble on a HEX27 vec-3 pattern (each block the segmented sum of its
    contribution list) against the same sums formed in NumPy in the same order: every
    block is the sum of its (cell, a, b) contributions in ascending cell order, from
    zero, in the element matrices' own precision - bit for bit."""
    say("hex27 row assembly", dtype)
    pts, cells = _hex27_box((3, 2, 2), seed=5)
    mat, _ = _mats(jfa, "neo_hookean", None)
    prob = _problem(jfa, pts, cells, mat)
    S = __import__("jax_fem_amd.solver", fromlist=["x"])
    rng = np.random.default_rng(2)
    u = torch.as_tensor(0.02 * rng.standard_normal(prob.N), device=DEV)
    _, ej = prob._element_terms(u, True, dtype)
    E = ej.reshape(len(cells), 27, 27, 3, 3).cpu().numpy()
    vals = prob.jacobian_bsr(u, dtype, bc=False)
    A = S.TangentMatrix(prob, vals).to_scipy().tocsr()
    npdt = np.float32 if dtype == torch.float32 else np.float64
    # block sums in ascending cell order (np.add.at adds in index order, unbuffered)
    nn = prob.fe.num_total_nodes
    keys = (cells[:, :, None].astype(np.int64) * nn + cells[:, None, :]).reshape(-1)
    uniq, inv = np.unique(keys, return_inverse=True)
    blocks = np.zeros((len(uniq), 3, 3), dtype=npdt)
    np.add.at(blocks, inv, E.reshape(-1, 3, 3).astype(npdt))
    rows, cols = uniq // nn, uniq % nn
    for i in range(3):
        for k in range(3):
            got = np.asarray(A[3 * rows + i, 3 * cols + k]).reshape(-1).astype(npdt)
            assert np.array_equal(got, blocks[:, i, k]), (i, k, np.abs(got - blocks[:, i, k]).max())
    say("  %d blocks equal bit for bit" % len(uniq))
