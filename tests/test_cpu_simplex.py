"""Simplex elements (TET4, TET10, TRI3, TRI6) on the host, against what the
mathematics fixes: Kronecker property at the nodes (VTK / meshio node order),
reproduction of polynomials of the element's degree, gradients by central
differences, exactness of the volume rules on monomials (int x^a y^b z^c over
the unit simplex = a! b! c! / (a+b+c+dim)!), the facet rules through the
divergence theorem (sum_f int n dA = 0, sum_f int x.n dA = dim |T|), and the
structured meshes' volumes, face areas, orientation and node use."""
from math import factorial
from importlib import import_module

import numpy as np
import pytest


@pytest.mark.parametrize("ele", ["TET4", "TET10", "TRI3", "TRI6"])
def test_simplex_tables(jfa, ele):
    fe = import_module("jax_fem_amd.fe")
    el = fe.ELEMENTS[ele]
    d = el.nodes.shape[1]
    vol = 0.5 if d == 2 else 1.0 / 6.0
    vn, _ = fe.tabulate(ele, el.nodes)
    np.testing.assert_allclose(vn, np.eye(len(el.nodes)), atol=1e-14)
    for order in range(7):
        P, W = fe.volume_rule(order, d, "simplex")
        assert abs(W.sum() - vol) < 1e-14
        for a in range(order + 1):
            for b in range(order + 1 - a):
                for c in (range(order + 1 - a - b) if d == 3 else [0]):
                    f = P[:, 0] ** a * P[:, 1] ** b * (P[:, 2] ** c if d == 3 else 1.0)
                    exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + d)
                    assert abs(f @ W - exact) < 1e-14, (order, a, b, c)
    X = np.random.default_rng(0).uniform(0, 0.3, (20, d))
    vx, gx = fe.tabulate(ele, X)
    np.testing.assert_allclose(vx.sum(1), 1, atol=1e-14)
    np.testing.assert_allclose(gx.sum(1), 0, atol=1e-13)
    if el.degree == 1:
        f = lambda x: 1 + 2 * x[:, 0] - x[:, 1] + (0.5 * x[:, 2] if d == 3 else 0)
    else:
        f = lambda x: 1 + x[:, 0] ** 2 - 3 * x[:, 0] * x[:, 1] + x[:, -1] ** 2
    np.testing.assert_allclose(vx @ f(el.nodes), f(X), atol=1e-13)
    h = 1e-6
    for e in range(d):
        Xp, Xm = X.copy(), X.copy()
        Xp[:, e] += h
        Xm[:, e] -= h
        fd = (fe.tabulate(ele, Xp)[0] - fe.tabulate(ele, Xm)[0]) / (2 * h)
        np.testing.assert_allclose(fd, gx[:, :, e], atol=1e-8)
    fp, fw = fe.face_rule(el.order, d, "simplex")
    normals = fe.FACETS[("simplex", d)][1]
    tot = sum(fw[k].sum() * normals[k] for k in range(len(fw)))
    xn = sum((fp[k] @ normals[k]) @ fw[k] for k in range(len(fw)))
    np.testing.assert_allclose(tot, 0, atol=1e-14)
    assert abs(xn - d * vol) < 1e-14


@pytest.mark.parametrize("ele,dims,vol,xface", [("TET4", (3, 2, 2, 1.5, 1.0, 1.0), 1.5, (8, 1.0)),
                                                ("TET10", (2, 2, 3, 1.0, 1.0, 1.5), 1.5, (12, 1.5)),
                                                ("TRI3", (3, 4, 2.0, 1.0), 2.0, (4, 1.0)),
                                                ("TRI6", (2, 3, 1.0, 1.0), 1.0, (3, 1.0))])
def test_simplex_meshes(jfa, ele, dims, vol, xface):
    fe = import_module("jax_fem_amd.fe")
    m = jfa.box_mesh_tet(*dims, ele) if ele.startswith("TET") else jfa.rectangle_mesh_tri(*dims, ele)
    assert list(m.cells_dict) == [jfa.get_meshio_cell_type(ele)]
    F = fe.FiniteElement(jfa.Mesh(m.points, m.cells_dict[jfa.get_meshio_cell_type(ele)]), 1, m.points.shape[1],
                         ele)
    assert np.all(F.JxW > 0)
    assert abs(F.JxW.sum() - vol) < 1e-12
    bi = F.get_boundary_conditions_inds([lambda p: np.isclose(p[0], 0.0)])[0]
    _, nanson = F.get_face_shape_grads(bi)
    assert len(bi) == xface[0] and abs(nanson.sum() - xface[1]) < 1e-12
    assert len(np.unique(m.cells)) == len(m.points)
    # conforming: every interior facet (by its vertex set) is shared by exactly two cells
    faces = np.sort(m.cells[:, F.face_inds].reshape(-1, F.face_inds.shape[1]), axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert counts.max() == 2
    if ele == "TET4":
        assert np.array_equal(jfa.box_mesh_gmsh(*dims, None, ele).cells, m.cells)


def test_quadrature_rule_other_than_default_is_refused(jfa):
    fe = import_module("jax_fem_amd.fe")
    with pytest.raises(NotImplementedError):
        fe.FiniteElement(jfa.box_mesh(1, 1, 1), 3, 3, "HEX8", quadrature_rule="GLL")
