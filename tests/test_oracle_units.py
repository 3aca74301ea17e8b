"""Oracle unit checks fixed by mathematics (closed forms, invariants, FD)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import basis, fe, kernels
from oracle import solver as osv


@pytest.mark.parametrize("ele", ["HEX8", "HEX27"])
def test_basis_invariants(ele):
    order = 2 if ele == "HEX8" else 4
    v, g, w = basis.get_shape_vals_and_grads(ele, order)
    assert abs(w.sum() - 1) < 1e-14
    np.testing.assert_allclose(v.sum(1), 1, atol=1e-13)
    np.testing.assert_allclose(g.sum(1), 0, atol=1e-12)
    X, _ = basis.hex_quadrature(order)
    nc = basis.node_coords(ele)
    # reproduction of polynomial fields of the element's degree
    f = lambda x: 1 + 2 * x[:, 0] - x[:, 1] + 0.5 * x[:, 2] + (x[:, 0] * x[:, 1] * x[:, 2] if ele == "HEX8" else x[:, 0] ** 2 * x[:, 1] ** 2 * x[:, 2] ** 2)
    np.testing.assert_allclose(v @ f(nc), f(X), atol=1e-12)
    # Kronecker property at nodes
    vn, _ = basis.tabulate(ele, nc)
    np.testing.assert_allclose(vn, np.eye(len(nc)), atol=1e-12)


def test_quad4_basis_invariants_and_edges():
    for order in (0, 2, 3, 5):
        v, g, w = basis.get_shape_vals_and_grads("QUAD4", order)
        assert abs(w.sum() - 1) < 1e-14 and v.shape[1] == 4 and g.shape[2] == 2
        np.testing.assert_allclose(v.sum(1), 1, atol=1e-14)
        np.testing.assert_allclose(g.sum(1), 0, atol=1e-13)
        X, _ = basis.quad_quadrature(order)
        nc = basis.node_coords("QUAD4")
        f = lambda x: 1 + 2 * x[:, 0] - x[:, 1] + 0.5 * x[:, 0] * x[:, 1]       # bilinear: reproduced
        np.testing.assert_allclose(v @ f(nc), f(X), atol=1e-13)
    # counter-clockwise corners (re_order [0,1,3,2] of basix's lexicographic vertices)
    np.testing.assert_array_equal(basis.node_coords("QUAD4"), [[0, 0], [1, 0], [1, 1], [0, 1]])
    vn, _ = basis.tabulate("QUAD4", basis.node_coords("QUAD4"))
    np.testing.assert_allclose(vn, np.eye(4), atol=1e-14)
    pts, cells = fe.rectangle_mesh(4, 3, 2.0, 1.5)
    assert np.allclose(pts[cells[5]], [[0.5, 1.0], [1.0, 1.0], [1.0, 1.5], [0.5, 1.5]])
    v, g, w = basis.get_shape_vals_and_grads("QUAD4")
    _, JxW = fe.shape_grads(pts, cells, g, w)
    assert abs(JxW.sum() - 3.0) < 1e-13 and np.all(JxW > 0)
    fsv, fsg, fw, fn, finds = basis.get_face_shape_vals_and_grads("QUAD4")
    for loc, length, nface in ((lambda p: abs(p[1] - 1.5) < 1e-9, 2.0, 4), (lambda p: abs(p[0]) < 1e-9, 1.5, 3)):
        bi = fe.boundary_inds(pts, cells, [loc], finds)[0]
        assert len(bi) == nface
        _, nanson = fe.face_shape_grads(pts, cells, bi, fsg, fn, fw)
        assert abs(nanson.sum() - length) < 1e-13
        xs = fe.physical_surface_quad_points(pts, cells, bi, fsv)
        assert all(loc(x) for x in xs.reshape(-1, 2))


def test_source_field_notebook_gradient():
    """docs/source/learn/source_field_identification/example.ipynb: Poisson on a
    64x64 QUAD4 square, mass map m = theta, u = 0 at x = 0 and x = 1; the
    gradient of J = sum (u_pred - u_true)^2 JxW at theta_ini = 1e5 is
    0.006816018916280742 at [10, 3] (finite differences: 0.0068160189166665075)."""
    pts, cells = fe.rectangle_mesh(64, 64)
    left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-5)
    info = [[left, right], [0, 0], [lambda p: 0.0, lambda p: 0.0]]
    mass = lambda x, th: th[..., None]

    def make(theta):
        return osv.OracleProblem(pts, cells, 1, kernels.Material("poisson", k=1.0), ele_type="QUAD4",
                                 dirichlet_bc_info=info, mass_map=mass, mass_params=theta)
    xq = make(np.zeros((len(cells), 4))).quad_points
    th_true = -1e5 * np.exp(-((xq[..., 0] - 0.5) ** 2 + (xq[..., 1] - 0.5) ** 2) / 0.02)
    u_true, _ = osv.newton_solve(make(th_true))
    pp = make(1e5 * np.ones_like(th_true))
    u_pred, _ = osv.newton_solve(pp)
    uq = lambda u: np.einsum("cn,qn->cq", u[cells], pp.shape_vals)
    d = uq(u_pred) - uq(u_true)
    v = np.zeros(pp.N)
    np.add.at(v, cells.reshape(-1), np.einsum("cq,qn,cq->cn", 2 * d, pp.shape_vals, pp.JxW).reshape(-1))
    g, _, _ = osv.implicit_vjp(pp, u_pred, v, wrt="mass")
    assert g.shape == (64 * 64, 4)
    assert abs(g[10, 3] - 0.006816018916280742) < 1e-9 * 0.0068


def test_gauss_exactness_and_hex27_order():
    x, w = basis.gauss_legendre_01(3)
    for k in range(6):
        assert abs(np.sum(w * x ** k) - 1.0 / (k + 1)) < 1e-14
    # HEX27 mesh order = VTK triquadratic hexahedron (basis.py:60-61 with basix numbering)
    nc = basis.node_coords("HEX27")
    assert np.allclose(nc[8], [0.5, 0, 0]) and np.allclose(nc[11], [0, 0.5, 0])
    assert np.allclose(nc[16], [0, 0, 0.5]) and np.allclose(nc[20], [0, 0.5, 0.5])
    assert np.allclose(nc[24], [0.5, 0.5, 0]) and np.allclose(nc[26], [0.5, 0.5, 0.5])


def test_face_areas_and_volume():
    pts, cells = fe.box_mesh(3, 2, 4, 1.5, 1.0, 2.0)
    v, g, w = basis.get_shape_vals_and_grads("HEX8")
    _, JxW = fe.shape_grads(pts, cells, g, w)
    assert abs(JxW.sum() - 3.0) < 1e-12
    fsv, fsg, fw, fn, finds = basis.get_face_shape_vals_and_grads("HEX8")
    bi = fe.boundary_inds(pts, cells, [lambda p: abs(p[0] - 1.5) < 1e-9], finds)[0]
    assert len(bi) == 8
    _, nanson = fe.face_shape_grads(pts, cells, bi, fsg, fn, fw)
    assert abs(nanson.sum() - 2.0) < 1e-12


@pytest.mark.parametrize("kind", ["linear_elastic", "neo_hookean"])
def test_tangent_is_fd_of_stress(kind):
    mat = kernels.Material(kind, E=10.0, nu=0.3)
    rng = np.random.default_rng(1)
    ug = 0.1 * rng.standard_normal((4, 3, 3))
    m1, m2, _, _ = mat.moduli(None)
    A = mat.tangent(ug, m1, m2)
    h = 1e-6
    for k in range(3):
        for l in range(3):
            d = np.zeros((3, 3)); d[k, l] = h
            fd = (mat.stress(ug + d, m1, m2) - mat.stress(ug - d, m1, m2)) / (2 * h)
            np.testing.assert_allclose(A[..., k, l], fd, rtol=1e-6, atol=1e-7)


def test_stress_is_fd_of_energy():
    mat = kernels.Material("neo_hookean", E=10.0, nu=0.3)
    rng = np.random.default_rng(2)
    ug = 0.1 * rng.standard_normal((3, 3))
    m1, m2, _, _ = mat.moduli(None)
    P = mat.stress(ug, m1, m2)
    h = 1e-6
    for k in range(3):
        for l in range(3):
            d = np.zeros((3, 3)); d[k, l] = h
            fd = (mat.energy(ug + d, m1, m2) - mat.energy(ug - d, m1, m2)) / (2 * h)
            assert abs(fd - P[k, l]) < 1e-6


def _small_problem(kind="neo_hookean", ele="HEX8"):
    if ele == "HEX8":
        pts, cells = fe.box_mesh(2, 2, 2)
        order = None
    else:
        pts, cells = fe.box_mesh_hex27(1, 1, 2)
        order = 4
    P = osv.OracleProblem(pts, cells, 3, kernels.Material(kind, E=10.0, nu=0.3), ele_type=ele,
                          quadrature_order=order)
    rng = np.random.default_rng(3)
    u = 0.05 * rng.standard_normal(P.N)
    return P, u


@pytest.mark.parametrize("ele", ["HEX8", "HEX27"])
def test_element_jacobian_fd_symmetry_rigid(ele):
    P, u = _small_problem("neo_hookean", ele)
    cu = P.cell_u(u)
    K = kernels.cell_jacobian(P.mat, cu, P.grads, P.JxW)
    nd = K.shape[1]
    h = 1e-6
    for c in range(nd):
        du = np.zeros_like(cu).reshape(len(cu), -1)
        du[:, c] = h
        rp = kernels.cell_residual(P.mat, cu + du.reshape(cu.shape), P.grads, P.JxW)
        rm = kernels.cell_residual(P.mat, cu - du.reshape(cu.shape), P.grads, P.JxW)
        np.testing.assert_allclose(K[:, :, c], (rp - rm) / (2 * h), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(K, np.swapaxes(K, 1, 2), atol=1e-10)
    # at u = 0 the rigid translations are in the null space of K (linear elastic)
    Kl = kernels.cell_jacobian(kernels.Material("linear_elastic", E=10.0, nu=0.3), 0 * cu, P.grads, P.JxW)
    t = np.tile([1.0, 0.0, 0.0], nd // 3)
    assert np.abs(Kl @ t).max() < 1e-10


def test_residual_is_fd_of_energy():
    P, u = _small_problem()
    R = P.compute_residual(u)
    h = 1e-6
    for i in [0, 5, 17, 40]:
        e = np.zeros_like(u); e[i] = h
        fd = (kernels.energy_total(P.mat, P.cell_u(u + e), P.grads, P.JxW)
              - kernels.energy_total(P.mat, P.cell_u(u - e), P.grads, P.JxW)) / (2 * h)
        assert abs(fd - R[i]) < 1e-6


def test_textbook_cube_stiffness_trace():
    # unit-cube hex8, linear elastic, 2x2x2 Gauss: K is exact for the trilinear
    # field; check the energy of a uniform strain equals volume * sigma:eps / 2.
    pts, cells = fe.box_mesh(1, 1, 1)
    P = osv.OracleProblem(pts, cells, 3, kernels.Material("linear_elastic", E=7.0, nu=0.25))
    K = kernels.cell_jacobian(P.mat, np.zeros((1, 8, 3)), P.grads, P.JxW)[0]
    eps = np.array([[1e-3, 2e-4, 0], [2e-4, -5e-4, 1e-4], [0, 1e-4, 3e-4]])
    u = (pts @ eps.T)[cells[0]].reshape(-1)
    lam = 7.0 * 0.25 / (1.25 * 0.5); mu = 7.0 / 2.5
    sig = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    assert abs(0.5 * u @ K @ u - 0.5 * np.sum(sig * eps)) < 1e-14


def test_pattern_and_row_elimination():
    P, u = _small_problem()
    P.update_bcs([[lambda p: abs(p[0]) < 1e-9] * 2, [0, 2], [lambda p: 0.0, lambda p: 0.1]])
    A = P.assemble(u)
    I, J = P.coo_pattern()
    ref = sp.coo_matrix((np.ones(len(I)), (I, J)), shape=(P.N, P.N)).tocsr()
    ref.sum_duplicates(); ref.sort_indices()
    assert np.array_equal(A.indptr, ref.indptr) and np.array_equal(A.indices, ref.indices)
    # structured closed form nnz = 9 (3nx+1)(3ny+1)(3nz+1)
    assert A.nnz == 9 * 7 ** 3
    for rows in P.bc_rows():
        for r in rows:
            row = A.getrow(r)
            assert row.nnz == A.indptr[r + 1] - A.indptr[r]           # pattern kept
            assert row[0, r] == 1.0 and abs(row).sum() == 1.0


def test_bicgstab_matches_direct_solve():
    P, u = _small_problem("linear_elastic")
    P.update_bcs([[lambda p: abs(p[0]) < 1e-9] * 3, [0, 1, 2], [lambda p: 0.0] * 3])
    A = P.assemble(u)
    b = np.random.default_rng(4).standard_normal(P.N)
    b[np.concatenate(P.bc_rows())] = 0
    x, k, info = osv.bicgstab(A, b, np.zeros_like(b), lambda v: v / A.diagonal())
    import scipy.sparse.linalg as spla
    xd = spla.spsolve(A.tocsc(), b)
    assert info == 0 and np.abs(x - xd).max() < 1e-8 * np.abs(xd).max()


def test_cnn_oracle_tap_convention():
    # torch conv3d (cross-correlation, padding 1) == the defining sum over taps
    import torch
    from oracle import cnn as ocnn
    rng = np.random.default_rng(0)
    grid = (4, 3, 5)
    nn = int(np.prod(grid))
    x = rng.standard_normal((nn, 4))
    W = rng.standard_normal((27, 4, 6))
    b = rng.standard_normal(6)
    y = ocnn.conv(torch.as_tensor(x)[None], ocnn.to_torch_weight(W), torch.as_tensor(b), grid)[0].numpy()
    for n in (0, 7, 31, nn - 1):
        np.testing.assert_allclose(y[n], ocnn.conv_direct_node(x, W, b, grid, n), rtol=1e-12, atol=1e-12)


def test_line_search_rule_and_root():
    """The oracle's line search (solver.py:424-464) on a tip-loaded Neo-Hookean
    cantilever: the step length is the first of 1, 1/2, 1/4, 1/8 after which the
    next halving raises the residual norm (checked by evaluating all four norms),
    and the damped Newton reaches the same root as the undamped one."""
    left = lambda p: np.isclose(p[0], 0.0, atol=1e-5)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-5)
    pts, cells = fe.box_mesh(3, 3, 3)
    op = osv.OracleProblem(pts, cells, 3, kernels.Material("neo_hookean", E=10.0, nu=0.3),
                           dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.0] * 3], location_fns=[right],
                           surface_maps=[lambda x: np.broadcast_to(np.array([0.0, 0.0, 1.0]), x.shape)])
    u = np.zeros(op.N)
    R = op.apply_bc_vec(op.compute_residual(u), u)
    A = op.assemble(u)
    inc, _ = osv.linear_solve(A, -R, op.assign_bc(np.zeros(op.N)) - op.copy_bc(u), "spsolve")
    alpha, unew = osv.line_search(op, u, inc)
    with np.errstate(invalid="ignore"):
        norms = [np.linalg.norm(op.apply_bc_vec(op.compute_residual(u + a * inc), u + a * inc))
                 for a in (1.0, 0.5, 0.25, 0.125)]
    expect = 0.125
    for k in range(3):
        if norms[k + 1] > norms[k]:
            expect = 0.5 ** k
            break
    assert alpha == expect and alpha < 1.0
    np.testing.assert_array_equal(unew, u + alpha * inc)
    with np.errstate(invalid="ignore"):
        u1, i1 = osv.newton_solve(op, method="spsolve", line_search_flag=True)
        u0, _ = osv.newton_solve(op, method="spsolve")
    assert min(i1["alphas"]) < 1.0
    assert np.linalg.norm(u1 - u0) / np.linalg.norm(u0) < 1e-8
