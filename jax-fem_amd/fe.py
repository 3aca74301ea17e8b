"""Set-up side of a finite-element variable (host, once per mesh / load step).

Mirrors ``jax_fem.fe.FiniteElement`` (fe.py:20-322) for the Lagrange elements
HEX8, HEX27, QUAD4 (tensor products on [0,1]^dim) and TET4, TET10, TRI3, TRI6
(barycentric, on the unit simplex): reference tables (basis.py:141-250),
Dirichlet index/value extraction (fe.py:220-269), boundary-face selection
(fe.py:271-322), and the face / volume load vectors of u-independent surface
and mass maps (problem.py:216-259).  None of this is on the solve path: the
per-iteration work (cell kernels, assembly, Krylov) runs in libhipfem.

Quadrature as basix's default rule for each cell (basis.py:154-160): Gauss
(Gauss-Jacobi = Gauss-Legendre) with (order+2)//2 points per direction on
intervals, quadrilaterals and hexahedra, first coordinate slowest; on simplices
the symmetric rules basix takes from Xiao-Gimbutas for degrees <= 2 (1 point,
3 / 4 points at the closed-form positions, equal weights) and, for higher
orders, collapsed Gauss-Jacobi (Stroud) rules of the same polynomial exactness
(basix uses Xiao-Gimbutas there too: different points, the same integrals of
polynomials up to the order).
"""
from collections import namedtuple

import numpy as np

from .mesh import HEX8_OFFSETS, HEX27_OFFSETS, QUAD4_OFFSETS

Element = namedtuple("Element", "shape nodes degree order")

_TET_VERTS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
_TRI_VERTS = np.array([[0, 0], [1, 0], [0, 1]], dtype=np.float64)
# mid-edge nodes in VTK / meshio order (the reference's re_order maps them onto
# basix's edge order, basis.py:80, :104)
_TET10_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
_TRI6_EDGES = [(0, 1), (1, 2), (0, 2)]


def _with_edges(verts, edges):
    return np.vstack([verts] + [0.5 * (verts[a] + verts[b])[None] for a, b in edges])


# Nodes are in the mesh's (VTK / meshio) order; the reference maps that order
# onto basix's with re_order (basis.py:53, :60-61, :74-110), which is the same as
# tabulating in mesh order directly.
ELEMENTS = {
    # shape, reference coordinates of the nodes, degree, default quadrature order
    # (basis.py:56, :64, :77, :83, :89, :101, :107)
    "HEX8": Element("cube", HEX8_OFFSETS / 1.0, 1, 2),
    "HEX27": Element("cube", HEX27_OFFSETS / 2.0, 2, 10),
    "QUAD4": Element("cube", QUAD4_OFFSETS / 1.0, 1, 2),
    "TET4": Element("simplex", _TET_VERTS, 1, 0),
    "TET10": Element("simplex", _with_edges(_TET_VERTS, _TET10_EDGES), 2, 2),
    "TRI3": Element("simplex", _TRI_VERTS, 1, 0),
    "TRI6": Element("simplex", _with_edges(_TRI_VERTS, _TRI6_EDGES), 2, 2),
}
# basix facets as mesh-order vertex lists (in basix's vertex order) and their
# outward normals.  Hexahedron: z=0, y=0, x=0, x=1, y=1, z=1; quadrilateral:
# y=0, x=0, x=1, y=1; simplices: facet i opposite vertex i.
_S3 = 1.0 / np.sqrt(3.0)
_S2 = 1.0 / np.sqrt(2.0)
FACETS = {
    ("cube", 3): (np.array([[0, 1, 3, 2], [0, 1, 4, 5], [0, 3, 4, 7], [1, 2, 5, 6], [3, 2, 7, 6], [4, 5, 7, 6]]),
                  np.array([[0, 0, -1], [0, -1, 0], [-1, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)),
    ("cube", 2): (np.array([[0, 1], [0, 3], [1, 2], [3, 2]]),
                  np.array([[0, -1], [-1, 0], [1, 0], [0, 1]], dtype=np.float64)),
    ("simplex", 3): (np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]]),
                     np.array([[_S3, _S3, _S3], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], dtype=np.float64)),
    ("simplex", 2): (np.array([[1, 2], [0, 2], [0, 1]]),
                     np.array([[_S2, _S2], [-1, 0], [0, -1]], dtype=np.float64)),
}
FACES, FACE_NORMALS = FACETS[("cube", 3)]


def element_dim(ele_type):
    return ELEMENTS[ele_type].nodes.shape[1]


def gauss_points(order):
    """Gauss-Legendre on [0,1] with (order+2)//2 points, ascending; weights sum to 1."""
    m = (order + 2) // 2
    t, w = np.polynomial.legendre.leggauss(m)
    return (t + 1.0) / 2.0, w / 2.0


def _gauss_jacobi01(n, alpha):
    """n-point Gauss-Jacobi rule on [0,1] for the weight (1-t)^alpha."""
    from scipy.special import roots_jacobi
    x, w = roots_jacobi(n, alpha, 0.0)
    return (x + 1.0) / 2.0, w / 2.0 ** (alpha + 1)


def simplex_rule(order, dim):
    """Quadrature on the unit triangle / tetrahedron (points (n, dim), weights summing
    to the volume 1/2 or 1/6)."""
    if order <= 1:
        return np.full((1, dim), 1.0 / (dim + 1)), np.array([1.0 / (2 if dim == 2 else 6)])
    if order == 2 and dim == 2:
        a, b = 1.0 / 6.0, 2.0 / 3.0
        return np.array([[a, a], [a, b], [b, a]]), np.full(3, 1.0 / 6.0)
    if order == 2 and dim == 3:
        a, b = (5.0 - np.sqrt(5.0)) / 20.0, (5.0 + 3.0 * np.sqrt(5.0)) / 20.0
        return np.array([[b, a, a], [a, a, a], [a, a, b], [a, b, a]]), np.full(4, 1.0 / 24.0)
    n = (order + 2) // 2
    if dim == 2:
        x, wx = _gauss_jacobi01(n, 1.0)
        y, wy = _gauss_jacobi01(n, 0.0)
        X, Y = np.meshgrid(x, y, indexing="ij")
        P = np.stack([X, Y * (1 - X)], -1).reshape(-1, 2)
        return P, np.outer(wx, wy).reshape(-1)
    x, wx = _gauss_jacobi01(n, 2.0)
    y, wy = _gauss_jacobi01(n, 1.0)
    z, wz = _gauss_jacobi01(n, 0.0)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    P = np.stack([X, Y * (1 - X), Z * (1 - X) * (1 - Y)], -1).reshape(-1, 3)
    return P, np.einsum("i,j,k->ijk", wx, wy, wz).reshape(-1)


def lagrange(deg, t):
    """1-D Lagrange values and derivatives through equispaced nodes on [0,1]: (len(t), deg+1)."""
    t = np.asarray(t, dtype=np.float64)[:, None]
    if deg == 1:
        return np.hstack([1 - t, t]), np.hstack([-np.ones_like(t), np.ones_like(t)])
    v = np.hstack([2 * (t - 0.5) * (t - 1), -4 * t * (t - 1), 2 * t * (t - 0.5)])
    d = np.hstack([4 * t - 3, 4 - 8 * t, 4 * t - 1])
    return v, d


def _tabulate_simplex(el, X):
    d = el.nodes.shape[1]
    lam = np.hstack([1.0 - X.sum(axis=1, keepdims=True), X])                 # (np, d+1)
    dlam = np.vstack([-np.ones((1, d)), np.eye(d)])                          # (d+1, d)
    nv = d + 1
    if el.degree == 1:
        return lam, np.broadcast_to(dlam[None], (len(X), nv, d)).copy()
    edges = _TET10_EDGES if d == 3 else _TRI6_EDGES
    vals = [lam[:, i] * (2 * lam[:, i] - 1) for i in range(nv)]
    grads = [(4 * lam[:, i] - 1)[:, None] * dlam[i][None] for i in range(nv)]
    for a, b in edges:
        vals.append(4 * lam[:, a] * lam[:, b])
        grads.append(4 * (lam[:, b][:, None] * dlam[a][None] + lam[:, a][:, None] * dlam[b][None]))
    return np.stack(vals, axis=1), np.stack(grads, axis=1)


def tabulate(ele_type, X):
    """Shape values (np, nn) and reference gradients (np, nn, dim) at points X (np, dim):
    products of 1-D Lagrange factors (cube cells) or barycentric Lagrange
    polynomials (simplices)."""
    el = ELEMENTS[ele_type]
    X = np.asarray(X, dtype=np.float64)
    if el.shape == "simplex":
        return _tabulate_simplex(el, X)
    nodes, deg = el.nodes, el.degree
    d = nodes.shape[1]
    ijk = np.rint(nodes * deg).astype(int)
    V, D = zip(*[lagrange(deg, X[:, c]) for c in range(d)])
    v = [V[c][:, ijk[:, c]] for c in range(d)]
    g = [D[c][:, ijk[:, c]] for c in range(d)]
    vals = np.prod(np.stack(v), axis=0)
    grads = np.stack([np.prod(np.stack([g[c] if c == e else v[c] for c in range(d)]), axis=0)
                      for e in range(d)], axis=-1)
    return vals, grads


def volume_rule(order, dim=3, shape="cube"):
    """The cell's quadrature rule: tensor Gauss on [0,1]^dim, first coordinate slowest
    (basix's order), or the simplex rule."""
    if shape == "simplex":
        return simplex_rule(order, dim)
    t, w = gauss_points(order)
    P = np.stack(np.meshgrid(*([t] * dim), indexing="ij"), -1).reshape(-1, dim)
    W = np.prod(np.stack(np.meshgrid(*([w] * dim), indexing="ij"), -1).reshape(-1, dim), axis=-1)
    return P, W


def face_rule(order, dim=3, shape="cube"):
    """Facet quadrature points (nfacet, nfq, dim) and weights (nfacet, nfq): the
    facet cell's rule (interval, quadrilateral or triangle) mapped affinely from the
    facet's first vertex along its edges to the second (and third) vertex, the
    weights scaled by the map's length / area factor (basis.py:205-231)."""
    if dim == 2:
        t, w = gauss_points(order)
        s, ws = t[:, None], w
    elif shape == "cube":
        t, w = gauss_points(order)
        s = np.stack(np.meshgrid(t, t, indexing="ij"), -1).reshape(-1, 2)
        ws = np.outer(w, w).reshape(-1)
    else:
        s, ws = simplex_rule(order, 2)
    if shape == "cube":
        corners = (HEX8_OFFSETS if dim == 3 else QUAD4_OFFSETS).astype(np.float64)
    else:
        corners = _TET_VERTS if dim == 3 else _TRI_VERTS
    facets = FACETS[(shape, dim)][0]
    pts, wts = [], []
    for f in facets:
        v0 = corners[f[0]]
        E = np.stack([corners[f[a + 1]] - v0 for a in range(dim - 1)], axis=1)   # (dim, dim-1)
        pts.append(v0[None] + s @ E.T)
        scale = np.linalg.norm(E[:, 0]) if dim == 2 else np.linalg.norm(np.cross(E[:, 0], E[:, 1]))
        wts.append(ws * scale)
    return np.stack(pts), np.stack(wts)


def _eval_points(fn, pts, extra=()):
    """Evaluate a per-point function over many points.

    Tries one call on the transposed batch (x[0] is then the vector of first
    coordinates, which is what element-wise per-point code computes), verifies
    it on a sample of single-point calls, and otherwise loops point by point."""
    n = len(pts)
    if n == 0:
        return np.zeros((0,))
    single = lambda i: np.asarray(fn(pts[i], *[e[i] for e in extra]), dtype=np.float64)
    try:
        out = np.asarray(fn(pts.T, *[np.asarray(e).T for e in extra]))
        if out.dtype == object:
            raise ValueError
        out = np.asarray(out, dtype=np.float64)
        ref0 = single(0)
        if out.shape == ref0.shape:           # constant map
            out = np.broadcast_to(out, (n,) + ref0.shape)
        else:
            out = np.moveaxis(out.reshape(ref0.shape + (n,)), -1, 0) if out.ndim else np.full(n, float(out))
        for i in sorted({(n - 1) * k // 6 for k in range(7)}):
            if not np.array_equal(out[i], single(i)):   # exact, NaN unequal
                raise ValueError
        return np.ascontiguousarray(out)
    except Exception:
        return np.stack([single(i) for i in range(n)])


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _tensor_device(obj):
    if _is_torch(obj) and hasattr(obj, "device"):
        return obj.device
    if isinstance(obj, (list, tuple)):
        for o in obj:
            d = _tensor_device(o)
            if d is not None:
                return d
    return None


class HostField(np.ndarray):
    """A NumPy array of set-up data (JxW, physical shape gradients, quadrature
    points) that meets torch tensors on their device: in an operation with a
    tensor it takes part as a tensor on that tensor's device, so reference-style
    post-processing such as ``np.sum((u_q - v_q)**2 * fe.JxW[:, :, None])`` runs
    on the GPU and stays differentiable.  On its own it is an ordinary array."""

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        import torch
        kwargs = kwargs or {}
        dev = _tensor_device(list(args) + list(kwargs.values())) or torch.device("cpu")

        def conv(a):
            if isinstance(a, HostField):
                return torch.as_tensor(np.asarray(a), device=dev)
            if isinstance(a, (list, tuple)):
                return type(a)(conv(b) for b in a)
            return a
        return func(*conv(args), **{k: conv(v) for k, v in kwargs.items()})


def _field(a):
    return np.asarray(a).view(HostField)


def _location_flags(fn, points):
    nargs = fn.__code__.co_argcount
    if nargs == 1:
        f = lambda x: fn(x)
        return _eval_points(f, points).astype(bool).reshape(-1)
    if nargs == 2:
        idx = np.arange(len(points))
        return _eval_points(fn, points, (idx,)).astype(bool).reshape(-1)
    raise ValueError("Wrong number of arguments for location_fn: must be 1 or 2, get %d" % nargs)


class FiniteElement:
    """One FE variable: mesh, tables, Dirichlet data, boundary faces (fe.py:20-110)."""

    def __init__(self, mesh, vec, dim, ele_type="HEX8", quadrature_rule=None, quadrature_order=None,
                 dirichlet_bc_info=None):
        if ele_type not in ELEMENTS:
            raise NotImplementedError("element %s (supported: %s)" % (ele_type, ", ".join(ELEMENTS)))
        if element_dim(ele_type) != dim:
            raise ValueError("element %s is %d-dimensional, dim=%d" % (ele_type, element_dim(ele_type), dim))
        if quadrature_rule not in (None, "default"):
            raise NotImplementedError("quadrature_rule %r: only basix's default rule is built" % (quadrature_rule,))
        self.mesh, self.vec, self.dim, self.ele_type = mesh, vec, dim, ele_type
        self.points = np.asarray(mesh.points, dtype=np.float64)
        self.cells = np.asarray(mesh.cells, dtype=np.int64)
        self.num_cells = len(self.cells)
        self.num_total_nodes = len(self.points)
        self.num_total_dofs = self.num_total_nodes * vec
        order = ELEMENTS[ele_type].order if quadrature_order is None else quadrature_order
        self.quadrature_order = order
        shape = ELEMENTS[ele_type].shape
        qp, qw = volume_rule(order, dim, shape)
        sv, sg = tabulate(ele_type, qp)
        fp, fw = face_rule(order, dim, shape)
        nf, nfq, _ = fp.shape
        fv, fg = tabulate(ele_type, fp.reshape(-1, dim))
        self.quad_weights, self.shape_vals, self.shape_grads_ref = _field(qw), _field(sv), _field(sg)
        self.face_quad_weights = _field(fw)
        self.face_shape_vals = _field(fv.reshape(nf, nfq, -1))
        self.face_shape_grads_ref = _field(fg.reshape(nf, nfq, -1, dim))
        self.face_inds, normals = FACETS[(shape, dim)]
        self.face_normals = _field(normals)
        self.num_quads, self.num_nodes = self.shape_vals.shape
        self.num_faces, self.num_face_quads = nf, nfq
        self.dirichlet_bc_info = dirichlet_bc_info
        self._loc_cache = {}
        self.node_inds_list, self.vec_inds_list, self.vals_list = self.Dirichlet_boundary_conditions(
            dirichlet_bc_info)

    def _loc_nodes(self, loc):
        """Nodes selected by a location function.  A location function is a pure
        function of the point (and index), and the mesh of an element is fixed, so
        the selection is evaluated once per function object; the cache holds the
        function itself, so its id cannot be reused by another one."""
        return self._loc_entry(loc)[1]

    def _loc_entry(self, loc):
        """(loc, nodes, their coordinates), cached per location function."""
        hit = self._loc_cache.get(id(loc))
        if hit is not None and hit[0] is loc:
            return hit
        nodes = np.flatnonzero(_location_flags(loc, self.points))
        hit = (loc, nodes, np.ascontiguousarray(self.points[nodes]))
        self._loc_cache[id(loc)] = hit
        return hit

    # -- Dirichlet data (fe.py:220-269) --------------------------------------
    def Dirichlet_boundary_conditions(self, dirichlet_bc_info):
        node_inds_list, vec_inds_list, vals_list = [], [], []
        if dirichlet_bc_info is None:
            return node_inds_list, vec_inds_list, vals_list
        location_fns, vecs, value_fns = dirichlet_bc_info
        assert len(location_fns) == len(value_fns) and len(value_fns) == len(vecs)
        for loc, comp, val in zip(location_fns, vecs, value_fns):
            _, nodes, pts = self._loc_entry(loc)
            probe = val(pts[0]) if len(nodes) else 0.0
            if _is_torch(probe):
                # differentiable Dirichlet values (ad_wrapper through set_params)
                import torch
                values = torch.stack([torch.as_tensor(val(self.points[i]), dtype=torch.float64).reshape(())
                                      for i in nodes.tolist()]) if len(nodes) else torch.zeros(0)
            else:
                values = _eval_points(val, pts).reshape(-1) if len(nodes) else np.zeros(0)
                values = np.asarray(values, dtype=np.float64)
            node_inds_list.append(nodes)
            vec_inds_list.append(np.full(len(nodes), comp, dtype=np.int32))
            vals_list.append(values)
        return node_inds_list, vec_inds_list, vals_list

    def update_Dirichlet_boundary_conditions(self, dirichlet_bc_info):
        self.dirichlet_bc_info = dirichlet_bc_info
        self.node_inds_list, self.vec_inds_list, self.vals_list = self.Dirichlet_boundary_conditions(
            dirichlet_bc_info)

    # -- boundary faces (fe.py:271-322) --------------------------------------
    def get_boundary_conditions_inds(self, location_fns):
        out = []
        for loc in (location_fns or []):
            flags = _location_flags(loc, self.points)
            on = np.all(flags[self.cells[:, self.face_inds]], axis=-1)       # (ncell, nfacet)
            out.append(np.argwhere(on))
        return out

    # -- geometry (fe.py:112-218), host FP64 ----------------------------------
    def _jac(self, coords, grads_ref):
        Jm = np.einsum("...nd,...ne->...de", coords, grads_ref)
        return Jm, np.linalg.det(Jm), np.linalg.inv(Jm)

    def get_shape_grads(self, points=None):
        pts = self.points if points is None else points
        X = pts[self.cells][:, None]                                         # (c,1,n,dim)
        _, det, inv = self._jac(X, self.shape_grads_ref[None])
        grads = np.einsum("qnd,cqde->cqne", self.shape_grads_ref, inv)
        return grads, det * self.quad_weights[None]

    def _geometry(self):
        if getattr(self, "_geom", None) is None:
            self._geom = self.get_shape_grads()
        return self._geom

    @property
    def shape_grads(self):
        """Physical shape gradients (num_cells, num_quads, num_nodes, dim) (problem.py:137)."""
        return _field(self._geometry()[0])

    @property
    def JxW(self):
        """det(J) * w per (cell, quadrature point) (num_cells, num_quads) (problem.py:137)."""
        return _field(self._geometry()[1])

    @property
    def v_grads_JxW(self):
        """(num_cells, num_quads, num_nodes, 1, dim) (problem.py:139)."""
        g, jxw = self._geometry()
        return _field(g[:, :, :, None, :] * jxw[:, :, None, None, None])

    def _cell_values(self, sol):
        """sol (num_total_nodes, vec) gathered per cell; torch stays torch (on its device)."""
        if _is_torch(sol):
            import torch
            return sol.reshape(self.num_total_nodes, self.vec)[torch.as_tensor(self.cells, device=sol.device)]
        return np.asarray(sol).reshape(self.num_total_nodes, self.vec)[self.cells]

    def sol_to_grad(self, sol):
        """u_grads (num_cells, num_quads, vec, dim) from a nodal solution (fe.py:366-382)."""
        cs = self._cell_values(sol)
        if _is_torch(cs):
            import torch
            g = torch.as_tensor(self.shape_grads, dtype=cs.dtype, device=cs.device)
            return torch.einsum("cni,cqnd->cqid", cs, g)
        return np.einsum("cni,cqnd->cqid", cs, self.shape_grads)

    def convert_from_dof_to_quad(self, sol):
        """u at the quadrature points (num_cells, num_quads, vec) (fe.py:324-341)."""
        cs = self._cell_values(sol)
        if _is_torch(cs):
            import torch
            v = torch.as_tensor(self.shape_vals, dtype=cs.dtype, device=cs.device)
            return torch.einsum("cni,qn->cqi", cs, v)
        return np.einsum("cni,qn->cqi", cs, self.shape_vals)

    def convert_from_dof_to_face_quad(self, sol, boundary_inds):
        """u at the face quadrature points (num_selected_faces, num_face_quads, vec) (fe.py:343-364)."""
        cs = self._cell_values(sol)[boundary_inds[:, 0]]
        vals = self.face_shape_vals[boundary_inds[:, 1]]
        if _is_torch(cs):
            import torch
            return torch.einsum("sni,sqn->sqi", cs, torch.as_tensor(vals, dtype=cs.dtype, device=cs.device))
        return np.einsum("sni,sqn->sqi", cs, vals)

    def get_physical_quad_points(self):
        return _field(np.einsum("qn,cnd->cqd", self.shape_vals, self.points[self.cells]))

    def get_face_shape_grads(self, boundary_inds):
        X = self.points[self.cells[boundary_inds[:, 0]]][:, None]
        g = self.face_shape_grads_ref[boundary_inds[:, 1]]
        _, det, inv = self._jac(X, g)
        grads = np.einsum("sqnd,sqde->sqne", g, inv)
        nJ = np.einsum("sd,sqde->sqe", self.face_normals[boundary_inds[:, 1]], inv)
        nanson = np.linalg.norm(nJ, axis=-1) * det * np.asarray(self.face_quad_weights)[boundary_inds[:, 1]]
        return _field(grads), _field(nanson)

    def get_physical_surface_quad_points(self, boundary_inds):
        return _field(np.einsum("sqn,snd->sqd", self.face_shape_vals[boundary_inds[:, 1]],
                                self.points[self.cells[boundary_inds[:, 0]]]))
