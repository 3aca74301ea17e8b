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


def test_quad4_tables_and_rectangle_match_oracle(jfa):
    from importlib import import_module
    fe = import_module("jax_fem_amd.fe")
    m = jfa.rectangle_mesh(5, 3, 2.0, 1.5)
    pts, cells = ofe.rectangle_mesh(5, 3, 2.0, 1.5)
    np.testing.assert_array_equal(m.points, pts)
    np.testing.assert_array_equal(m.cells_dict[jfa.get_meshio_cell_type("QUAD4")], cells)
    top = lambda p: np.isclose(p[1], 1.5, atol=1e-5)
    for order in (None, 1, 3):
        F = fe.FiniteElement(jfa.Mesh(m.points, m.cells), 1, 2, "QUAD4", quadrature_order=order)
        v, g, w = obasis.get_shape_vals_and_grads("QUAD4", order)
        np.testing.assert_allclose(F.shape_vals, v, atol=1e-15)
        np.testing.assert_allclose(F.shape_grads_ref, g, atol=1e-14)
        np.testing.assert_allclose(F.quad_weights, w, atol=1e-15)
        fv, fg, fw, fn, fi = obasis.get_face_shape_vals_and_grads("QUAD4", order)
        np.testing.assert_allclose(F.face_shape_vals, fv, atol=1e-15)
        np.testing.assert_allclose(F.face_quad_weights, fw, atol=1e-15)
        np.testing.assert_array_equal(F.face_inds, fi)
        np.testing.assert_array_equal(F.face_normals, fn)
        sg, jxw = ofe.shape_grads(pts, cells, g, w)
        np.testing.assert_allclose(F.shape_grads, sg, atol=1e-13)
        np.testing.assert_allclose(F.JxW, jxw, atol=1e-15)
        bi = F.get_boundary_conditions_inds([top])[0]
        np.testing.assert_array_equal(bi, ofe.boundary_inds(pts, cells, [top], fi)[0])
        assert abs(F.get_face_shape_grads(bi)[1].sum() - 2.0) < 1e-13
    with pytest.raises(ValueError):
        fe.FiniteElement(m, 1, 3, "QUAD4")


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


def test_plain_maps_probed_and_planned(jfa):
    """A plain-function tensor map (the reference's idiom: jax.grad of an energy,
    test_hyper_elasticity.py:15-34) is recognised as the built law it equals, with
    moduli fitted to 1e-12; a map that is no built law is planned for the pointwise
    (device) path; mass and surface maps that depend on u become element terms,
    u-independent ones a load vector."""
    import jax_fem_amd.numpy as np_
    from jax_fem_amd import autodiff as ad

    def psi(F):
        E, nu = 1e3, 0.3
        mu, kappa = E / (2. * (1. + nu)), E / (3. * (1. - 2. * nu))
        J = np_.linalg.det(F)
        return (mu / 2.) * (J ** (-2. / 3.) * np_.trace(F.T @ F) - 3.) + (kappa / 2.) * (J - 1.) ** 2.
    P_fn = ad.grad(psi)

    class Hyper(jfa.Problem):
        def get_tensor_map(self):
            return lambda u_grad: P_fn(u_grad + np_.eye(3))

    p = Hyper(jfa.box_mesh(2, 2, 2), vec=3, dim=3)
    plan = p._plan()
    assert plan["law"] == "fast" and plan["pure_fast"]
    assert type(p.material).__name__ == "NeoHookean"
    assert abs(p.material.E / 1e3 - 1) < 1e-12 and abs(p.material.nu / 0.3 - 1) < 1e-12

    class Cubic(jfa.Problem):
        def get_tensor_map(self):
            return lambda u_grad: u_grad + u_grad ** 3

        def get_mass_map(self):
            return lambda u, x: 3.0 * u

        def get_surface_maps(self):
            return [lambda u, x: np.array([0.0, 0.0, 1.0])]

    q = Cubic(jfa.box_mesh(2, 2, 2), vec=3, dim=3, location_fns=[lambda x: np.isclose(x[0], 1.0)])
    plan = q._plan()
    assert plan["law"] == "generic" and plan["mass"] == "cell" and plan["surfaces"] == ["load"]
    assert not plan["pure_fast"]


def test_complex_step_and_torch_transforms_agree(jfa):
    """grad / jacfwd of a NumPy-written map by the complex step equal torch.func's
    on the same code (jax_fem_amd.numpy dispatches on the argument type), and the
    Neo-Hookean first Piola stress of the built law."""
    import torch
    import jax_fem_amd.numpy as np_
    from jax_fem_amd import autodiff as ad

    def psi(F):
        J = np_.linalg.det(F)
        return 3.0 * (J ** (-2. / 3.) * np_.sum(F * F) - 3.) + 20.0 * (J - 1.) ** 2 + np_.sqrt(np_.trace(F))
    g = np.random.default_rng(4).standard_normal((3, 3)) * 0.1 + np.eye(3)
    pn = ad.grad(psi)(g)
    pt = ad.grad(psi)(torch.tensor(g)).numpy()
    np.testing.assert_allclose(pn, pt, rtol=1e-12, atol=1e-12)
    f = lambda gg: ad.grad(psi)(gg)
    Cn = ad.jacfwd(f)(g)
    Ct = ad.jacfwd(f)(torch.tensor(g)).numpy()
    np.testing.assert_allclose(Cn, Ct, rtol=1e-10, atol=1e-10)
    U = torch.tensor(np.random.default_rng(5).standard_normal((4, 3, 3)) * 0.05)
    nh = jfa.NeoHookean(E=10.0, nu=0.3)

    def energy(F):
        mu, kappa = nh.moduli()
        J = np_.linalg.det(F)
        return mu / 2 * (J ** (-2. / 3.) * np_.trace(F.T @ F) - 3.) + kappa / 2 * (J - 1.) ** 2
    P = ad.vmap(lambda ug: ad.grad(energy)(ug + np_.eye(3)))(U)
    np.testing.assert_allclose(P.numpy(), nh(U.numpy()), rtol=1e-12, atol=1e-12)


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
    # re-wrapped the reference's way: the lattice is detected from the coordinates
    m2 = jfa.Mesh(m.points, m.cells_dict[jfa.get_meshio_cell_type("HEX8")])
    assert mgm.lattice(P(m2, vec=3, dim=3)) == (3, 3, 3)
    perm = np.random.default_rng(0).permutation(len(m.points))
    with pytest.raises(NotImplementedError):   # the same body, nodes relabelled: no lattice
        mgm.lattice(P(jfa.Mesh(m.points[perm], np.argsort(perm)[m.cells], "HEX8"), vec=3, dim=3))
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


def test_vtu_round_trip_and_binary_reader(jfa, tmp_path):
    """save_sol writes the mesh and 'sol' (float32, as the reference); read_vtu gives
    the points, cells_dict and point data back; a base64 zlib-compressed array laid
    out as the VTK XML format specifies (UInt64 header: block count, block size,
    last block size, compressed sizes) decodes to the same numbers."""
    import base64
    import struct
    import zlib
    from importlib import import_module
    fe_mod = import_module("jax_fem_amd.fe")
    m = jfa.box_mesh(3, 2, 2, 1.0, 0.5, 0.25)
    fe = fe_mod.FiniteElement(m, 3, 3, "HEX8")
    sol = np.random.default_rng(0).standard_normal((len(m.points), 3))
    path = str(tmp_path / "out" / "u.vtu")
    jfa.save_sol(fe, sol, path, cell_infos=[("rho", np.arange(fe.num_cells, dtype=float))],
                 point_infos=[("T", sol[:, 0])])
    r = jfa.read_vtu(path)
    np.testing.assert_array_equal(r.points, m.points)
    np.testing.assert_array_equal(r.cells_dict["hexahedron"], m.cells)
    np.testing.assert_array_equal(r.point_data["sol"], sol.astype(np.float32))
    np.testing.assert_array_equal(r.cell_data["rho"], np.arange(fe.num_cells, dtype=np.float32))
    mesh = jfa.Mesh(r.points, r.cells_dict[jfa.get_meshio_cell_type("HEX8")])
    assert mesh.node_grid == (4, 3, 3)
    # a compressed binary file
    vals = np.linspace(0.0, 1.0, 8 * 3)
    raw = vals.astype("<f8").tobytes()
    comp = zlib.compress(raw)
    blob = base64.b64encode(struct.pack("<QQQQ", 1, len(raw), len(raw), len(comp)) + comp).decode()
    conn = base64.b64encode(struct.pack("<QQQQ", 1, 64, 64, len(zlib.compress(np.arange(8, dtype="<i8").tobytes())))
                            + zlib.compress(np.arange(8, dtype="<i8").tobytes())).decode()
    off = np.array([8], dtype="<i8").tobytes()
    offb = base64.b64encode(struct.pack("<QQQQ", 1, 8, 8, len(zlib.compress(off))) + zlib.compress(off)).decode()
    typ = np.array([12], dtype="<u1").tobytes()
    typb = base64.b64encode(struct.pack("<QQQQ", 1, 1, 1, len(zlib.compress(typ))) + zlib.compress(typ)).decode()
    xml = ('<?xml version="1.0"?><VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
           'header_type="UInt64" compressor="vtkZLibDataCompressor"><UnstructuredGrid>'
           '<Piece NumberOfPoints="8" NumberOfCells="1"><Points><DataArray type="Float64" NumberOfComponents="3" '
           'format="binary">%s</DataArray></Points><Cells><DataArray type="Int64" Name="connectivity" '
           'format="binary">%s</DataArray><DataArray type="Int64" Name="offsets" format="binary">%s</DataArray>'
           '<DataArray type="UInt8" Name="types" format="binary">%s</DataArray></Cells></Piece></UnstructuredGrid>'
           '</VTKFile>' % (blob, conn, offb, typb))
    p2 = tmp_path / "b.vtu"
    p2.write_text(xml)
    r2 = jfa.read_vtu(str(p2))
    np.testing.assert_array_equal(r2.points.reshape(-1), vals)
    np.testing.assert_array_equal(r2.cells_dict["hexahedron"], np.arange(8)[None])
