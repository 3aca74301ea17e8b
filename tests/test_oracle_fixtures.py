"""Oracle vs the reference's own gold data (FEniCSx fixtures, notebook gradients).

Fixtures: tests/benchmarks/*/fenicsx (converted by tests/golden/make_fixtures.py).
Tolerances are the reference tests' own: assert_array_almost_equal decimal=5
(|d| < 1.5e-5), cylinder decimal=3, area decimal=4, traction decimal=5.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import basis, fe, kernels
from oracle import solver as osv


def iscl(a, b):
    return abs(a - b) < 1e-5


def fixture(name):
    return np.load(os.path.join(GOLDEN, "fenicsx_%s.npz" % name), allow_pickle=False)


def test_linear_poisson():
    # tests/benchmarks/linear_poisson/test_linear_poisson.py:15-65
    d = fixture("linear_poisson")
    P = osv.OracleProblem(d["points"], d["cells"], 1, kernels.Material("poisson"),
                          dirichlet_bc_info=[[lambda p: iscl(p[0], 0), lambda p: iscl(p[0], 1)],
                                             [0, 0], [lambda p: 0.0, lambda p: 1.0]])
    u, _ = osv.newton_solve(P)
    np.testing.assert_array_almost_equal(d["sol"].ravel(), u, decimal=5)
    np.testing.assert_allclose(u, d["points"][:, 0], atol=1e-8)      # closed form u = x


def test_linear_elasticity_cube():
    # tests/benchmarks/linear_elasticity_cube/test_linear_elasticity_cube.py:15-83
    d = fixture("linear_elasticity_cube")
    left = lambda p: iscl(p[0], 0.0)
    right = lambda p: iscl(p[0], 1.0)
    P = osv.OracleProblem(d["points"], d["cells"], 3, kernels.Material("linear_elastic", E=70e3, nu=0.3),
                          dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 1.0] * 3],
                          location_fns=[right],
                          surface_maps=[lambda x: -np.array([10.0, 0, 0])],
                          mass_map=lambda x: -np.array([0.0, 10.0, 10.0]))
    u, _ = osv.newton_solve(P)
    np.testing.assert_array_almost_equal(d["sol"].ravel(), u, decimal=5)


def _cyl_elastic():
    d = fixture("linear_elasticity_cylinder")
    top = lambda p: iscl(p[2], 10.0)
    bot = lambda p: iscl(p[2], 0.0)
    z = lambda x: 0.0 * x[..., 0]
    P = osv.OracleProblem(
        d["points"], d["cells"], 3, kernels.Material("linear_elastic", E=70e3, nu=0.3),
        dirichlet_bc_info=[[bot] * 3, [0, 1, 2], [lambda p: 0.0] * 3], location_fns=[top],
        surface_maps=[lambda x: -np.stack([1e3 * x[..., 0] ** 2 + 1e3 * x[..., 1] ** 2, z(x), z(x)], -1)],
        mass_map=lambda x: -np.stack([1e3 * x[..., 0], 2e3 * x[..., 1], 3e3 * x[..., 2]], -1))
    return d, P


def test_linear_elasticity_cylinder():
    # tests/benchmarks/linear_elasticity_cylinder/test_linear_elasticity_cylinder.py:15-132
    d, P = _cyl_elastic()
    u, _ = osv.newton_solve(P)
    np.testing.assert_array_almost_equal(d["sol"].ravel(), u, decimal=3)
    area = P.face_data[0][2].sum()
    np.testing.assert_almost_equal(float(d["surface_area"]), area, decimal=4)
    assert abs(area - 76.536686473018) < 1e-10


def test_hyperelasticity_cylinder_and_traction():
    # tests/benchmarks/hyperelasticity/test_hyper_elasticity.py:15-127
    d = fixture("hyperelasticity")
    top = lambda p: iscl(p[2], 10.0)
    bot = lambda p: iscl(p[2], 0.0)
    mat = kernels.Material("neo_hookean", E=1e3, nu=0.3)
    P = osv.OracleProblem(d["points"], d["cells"], 3, mat,
                          dirichlet_bc_info=[[bot, bot, bot, top, top, top], [0, 1, 2, 0, 1, 2],
                                             [lambda p: 0.0] * 5 + [lambda p: 1.0]],
                          location_fns=[top], surface_maps=[lambda x: 0.0 * x])
    u, info = osv.newton_solve(P)
    assert info["newton_iters"] >= 3
    np.testing.assert_array_almost_equal(d["sol"].ravel(), u, decimal=5)
    # compute_traction (test_hyper_elasticity.py:48-70): int P . e_z dA over the top face
    fsv, fsg, fw, fn, finds = basis.get_face_shape_vals_and_grads("HEX8")
    bi = P.boundary_inds_list[0]
    g, nanson = fe.face_shape_grads(P.points, P.cells, bi, fsg, fn, fw)
    ug = np.einsum("sni,sqnj->sqij", u.reshape(-1, 3)[P.cells[bi[:, 0]]], g)
    m1, m2, _, _ = mat.moduli(None)
    tz = np.einsum("sqij,j,sq->i", mat.stress(ug, m1, m2), np.array([0, 0, 1.0]), nanson)[2]
    np.testing.assert_almost_equal(float(d["traction"]), tz, decimal=5)


def _gradients_problem(scale, rho):
    """docs/source/learn/compute_gradients/example.ipynb cells 11-21 (5^3 box)."""
    pts, cells = fe.box_mesh(5, 5, 5)
    bottom = lambda p: iscl(p[2], 0.0)
    topz = lambda p: iscl(p[2], 1.0)
    return osv.OracleProblem(
        pts, cells, 3, kernels.Material("neo_hookean", E=1e6, nu=0.3, param="scale_E"),
        dirichlet_bc_info=[[bottom] * 3, [0, 1, 2], [lambda p: 0.0, lambda p: 0.0, lambda p: scale]],
        location_fns=[topz], surface_maps=[lambda x: np.broadcast_to(np.array([0, 0, 1e3]), x.shape)],
        params=rho)


def test_notebook_adjoint_golden_values():
    # example.ipynb md cell 31: dscale_d = 431.59223609853564, drho[0,0] = 0.002266954599447443
    rho = 0.5 * np.ones((125, 8))
    P = _gradients_problem(1.0, rho)
    u, _ = osv.newton_solve(P)
    rows_z = P.bc_rows()[2]
    gp, lam, dg = osv.implicit_vjp(P, u, 2 * u, dirichlet_param_grads=[(rows_z, np.ones(len(rows_z)))])
    assert abs(dg[0] / 431.59223609853564 - 1) < 1e-9
    # drho[0,0] depends on gmsh's cell 0 / basix's quad 0: the 4 symmetric bottom corners
    # (box_mesh cells 0, 20, 100, 120 at their corner-nearest quad points) all carry it.
    for c, q in ((0, 0), (20, 2), (100, 4), (120, 6)):
        assert abs(gp[c, q] / 0.002266954599447443 - 1) < 1e-9


def test_notebook_adjoint_vs_central_fd():
    rho = 0.5 * np.ones((125, 8))
    P = _gradients_problem(1.0, rho)
    u, _ = osv.newton_solve(P, rel_tol=1e-12)
    gp, _, _ = osv.implicit_vjp(P, u, 2 * u)
    rng = np.random.default_rng(0)
    for c in rng.choice(125, 3, replace=False):
        h = 1e-4
        vals = []
        for s in (+1, -1):
            r2 = rho.copy()
            r2[c, 0] += s * h
            P2 = _gradients_problem(1.0, r2)
            u2, _ = osv.newton_solve(P2, u0=u, rel_tol=1e-13, tol=1e-13)
            vals.append(np.sum(u2 ** 2))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - gp[c, 0]) <= 1e-5 * abs(gp[c, 0]) + 1e-12
