"""CPU checks of the drop-in boundary: C-ABI exports, host set-up vs the oracle."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from oracle import basis as obasis
from oracle import fe as ofe
from oracle import kernels as ok
from oracle import solver as osv


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "hipfem.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(jfa):
    lib = jfa._lib.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    # the binding covers every declared entry point
    assert set(jfa._lib.declared_symbols()) == set(names)


def test_tables_match_oracle(jfa):
    from importlib import import_module
    fe = import_module("jax_fem_amd.fe")
    for ele, order in (("HEX8", 2), ("HEX27", 4)):
        v, g, w = obasis.get_shape_vals_and_grads(ele, order)
        F = fe.FiniteElement(jfa.box_mesh(1, 1, 1) if ele == "HEX8" else jfa.box_mesh_hex27(1, 1, 1), 3, 3, ele,
                             quadrature_order=order)
        np.testing.assert_allclose(F.shape_vals, v, atol=1e-14)
        np.testing.assert_allclose(F.shape_grads_ref, g, atol=1e-13)
        np.testing.assert_allclose(F.quad_weights, w, atol=1e-15)
        fv, fg, fw, fn, fi = obasis.get_face_shape_vals_and_grads(ele, order)
        np.testing.assert_allclose(F.face_shape_vals, fv, atol=1e-14)
        np.testing.assert_allclose(F.face_quad_weights, fw, atol=1e-15)
        np.testing.assert_array_equal(F.face_inds, fi)
        np.testing.assert_array_equal(F.face_normals, fn)


def _iscl(a, b):
    return np.abs(a - b) < 1e-5


def test_bc_and_faces_match_oracle(jfa):
    d = np.load(os.path.join(GOLDEN, "fenicsx_linear_elasticity_cylinder.npz"))
    mesh = jfa.Mesh(d["points"], d["cells"])
    top = lambda p: np.isclose(p[2], 10.0, atol=1e-5)
    bot = lambda p: np.isclose(p[2], 0.0, atol=1e-5)
    info = [[bot, bot, top], [0, 2, 1], [lambda p: 0.0, lambda p: 0.5, lambda p: p[0] ** 2]]
    from importlib import import_module
    F = import_module("jax_fem_amd.fe").FiniteElement(mesh, 3, 3, "HEX8", dirichlet_bc_info=info)
    on, ov, ova = ofe.dirichlet_bcs(d["points"], info)
    for a, b in zip(F.node_inds_list, on):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(F.vals_list, ova):
        np.testing.assert_allclose(a, b, atol=0)
    bi = F.get_boundary_conditions_inds([top])[0]
    obi = ofe.boundary_inds(d["points"], d["cells"], [top], obasis.get_face_shape_vals_and_grads("HEX8")[4])[0]
    np.testing.assert_array_equal(bi, obi)
    _, nanson = F.get_face_shape_grads(bi)
    assert abs(nanson.sum() - 76.536686473018) < 1e-10


def test_problem_rejects_plain_python_map(jfa):
    class P(jfa.Problem):
        def get_tensor_map(self):
            return lambda g: g
    with pytest.raises(NotImplementedError):
        P(jfa.box_mesh(1, 1, 1), vec=1, dim=3)


def test_coo_pattern_matches_oracle(jfa):
    class P(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=1.0, nu=0.3)

        def get_surface_maps(self):
            return [lambda u, x: np.array([0.0, 0.0, 1.0])]
    m = jfa.box_mesh(2, 1, 2)
    prob = P(m, vec=3, dim=3, location_fns=[lambda p: np.isclose(p[0], 2.0 / 2.0, atol=1e-5)])
    op = osv.OracleProblem(m.points, m.cells, 3, ok.Material("linear_elastic"),
                           location_fns=[lambda p: np.isclose(p[0], 1.0, atol=1e-5)],
                           surface_maps=[lambda x: np.array([0.0, 0.0, 1.0])])
    I, J = op.coo_pattern()
    np.testing.assert_array_equal(prob.I, I)
    np.testing.assert_array_equal(prob.J, J)


def test_solver_options_resolution(jfa):
    from importlib import import_module
    s = import_module("jax_fem_amd.solver")
    assert s._resolve_solver_options({}) == ("newton", {})
    m, cfg = s._resolve_solver_options({"petsc_solver": {}, "tol": 1e-5})
    assert m == "newton" and cfg == {"tol": 1e-5, "linear": {"petsc_solver": {}}}
    with pytest.raises(ValueError):
        s._resolve_solver_options({"newton": {}, "arc_length": {}})
    with pytest.raises(ValueError):
        s._resolve_solver_options({"newton": 3})


def test_multigrid_lattice_detection(jfa):
    from importlib import import_module
    mgm = import_module("jax_fem_amd.multigrid")

    class P(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=1.0, nu=0.3)

    assert mgm.lattice(P(jfa.box_mesh(3, 2, 4), vec=3, dim=3)) == (4, 3, 5)
    assert mgm.lattice(P(jfa.box_mesh_hex27(2, 1, 2), vec=3, dim=3, ele_type="HEX27",
                         quadrature_order=4)) == (5, 3, 5)
    m = jfa.box_mesh(2, 2, 2)
    with pytest.raises(NotImplementedError):   # the same cells without the lattice metadata
        mgm.lattice(P(jfa.Mesh(m.points, m.cells, "HEX8"), vec=3, dim=3))
    class Q(jfa.Problem):
        def get_tensor_map(self):
            return jfa.Poisson()

    with pytest.raises(NotImplementedError):   # vec = 1
        mgm.lattice(Q(jfa.box_mesh(2, 2, 2), vec=1, dim=3))


def test_box_corner_table(jfa):
    """The corner table that lets the stencil assembly take a node's cells from the
    coordinates: corner (dx, dy, dz) of every cell sits at local index
    table[4dx + 2dy + dz]; meshes whose cells are not in lattice order get none."""
    class P(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=1.0, nu=0.3)

    nxc, nyc, nzc = 3, 2, 4
    prob = P(jfa.box_mesh(nxc, nyc, nzc), vec=3, dim=3)
    grid = (nxc + 1, nyc + 1, nzc + 1)
    table = prob._box_corners(grid)
    assert table is not None and sorted(table) == list(range(8))
    cells = np.asarray(prob.fe.cells)
    ny, nz = grid[1], grid[2]
    for c in (0, 5, cells.shape[0] - 1):   # brute force: every corner's node id from the lattice
        i, j, k = c // (nyc * nzc), (c // nzc) % nyc, c % nzc
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    assert cells[c, table[4 * dx + 2 * dy + dz]] == ((i + dx) * ny + (j + dy)) * nz + (k + dz)
    prob.fe.cells = cells[::-1].copy()      # the same cells, not in lattice order
    prob._corner_key = None
    assert prob._box_corners(grid) is None
    prob.fe.cells = cells[:, [1, 0, 2, 3, 4, 5, 6, 7]].copy()   # one consistent corner order: still a table
    prob._corner_key = None
    t2 = prob._box_corners(grid)
    assert t2 is not None and list(t2) == [1 if t == 0 else 0 if t == 1 else t for t in table]
