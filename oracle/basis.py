"""Reference-cell basis and quadrature for HEX8 / HEX27 / QUAD4 (oracle).

Restates ``jax_fem/basis.py``:
  * ``get_elements`` (basis.py:19-114): HEX8 re_order [0,1,3,2,4,5,7,6] (:53),
    default quadrature order 2 (:56); HEX27 re_order (:60-61), default order 10 (:64);
    QUAD4 re_order [0,1,3,2] (:87), default order 2 (:89), interval facets.
  * ``get_shape_vals_and_grads`` (basis.py:141-175): Lagrange basis of degree 1/2
    on the [0,1]^3 reference cell at Gauss-Jacobi (= Gauss-Legendre for the hex)
    points, columns permuted into mesh (VTK) node order.
  * ``get_face_shape_vals_and_grads`` (basis.py:178-250): face quadrature mapped
    by the facet vertices, facet weights scaled by the facet Jacobian, outward
    normals, face vertex lists in mesh numbering.

basix (0.10.0, absent here) provides the tabulation in the reference.  The
restatement works from the element definition instead: the basis function of a
node is the tensor product of 1-D Lagrange polynomials through the node's
reference coordinates, and the node coordinates in mesh order follow from
basix's hexahedron numbering (vertices lexicographic with x fastest, then the 12
edges, 6 facets and the interior) or quadrilateral numbering (vertices
lexicographic with x fastest) composed with ``re_order``.

Pinned by: partition of unity, sum of gradients = 0, exact reproduction of
linear (HEX8, QUAD4) and quadratic (HEX27) fields, weights summing to 1,
exactness of the Gauss rule on monomials, and end-to-end by the FEniCSx fixtures
and the source-field notebook's gradient (QUAD4).
"""
import numpy as np

# basix hexahedron vertex coordinates (lexicographic, x fastest).
_BASIX_VERTS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                         [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]], dtype=np.float64)
# basix hexahedron edges and facets (vertex lists).
_BASIX_EDGES = [[0, 1], [0, 2], [0, 4], [1, 3], [1, 5], [2, 3],
                [2, 6], [3, 7], [4, 5], [4, 6], [5, 7], [6, 7]]
_BASIX_FACETS = [[0, 1, 2, 3], [0, 1, 4, 5], [0, 2, 4, 6],
                 [1, 3, 5, 7], [2, 3, 6, 7], [4, 5, 6, 7]]
# basix outward facet normals, in facet order.
_BASIX_FACET_NORMALS = np.array([[0, 0, -1], [0, -1, 0], [-1, 0, 0],
                                 [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)

# basix quadrilateral vertices (x fastest), facets (intervals) and outward normals.
_BASIX_QUAD_VERTS = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], dtype=np.float64)
_BASIX_QUAD_FACETS = [[0, 1], [0, 2], [1, 3], [2, 3]]
_BASIX_QUAD_FACET_NORMALS = np.array([[0, -1], [-1, 0], [1, 0], [0, 1]], dtype=np.float64)

RE_ORDER = {
    "HEX8": [0, 1, 3, 2, 4, 5, 7, 6],                                    # basis.py:53
    "HEX27": [0, 1, 3, 2, 4, 5, 7, 6, 8, 11, 13, 9, 16, 18, 19,
              17, 10, 12, 15, 14, 22, 23, 21, 24, 20, 25, 26],           # basis.py:60-61
    "QUAD4": [0, 1, 3, 2],                                               # basis.py:87
}
DEFAULT_ORDER = {"HEX8": 2, "HEX27": 10, "QUAD4": 2}                     # basis.py:56,64,89
DEGREE = {"HEX8": 1, "HEX27": 2, "QUAD4": 1}
DIM = {"HEX8": 3, "HEX27": 3, "QUAD4": 2}


def basix_node_coords(degree):
    """Reference coordinates of basix's P1/P2 hexahedron DOFs in basix order."""
    pts = [v for v in _BASIX_VERTS]
    if degree == 2:
        pts += [0.5 * (_BASIX_VERTS[a] + _BASIX_VERTS[b]) for a, b in _BASIX_EDGES]
        pts += [_BASIX_VERTS[f].mean(axis=0) for f in _BASIX_FACETS]
        pts += [np.array([0.5, 0.5, 0.5])]
    return np.array(pts)


def node_coords(ele_type):
    """Reference coordinates of the element nodes in mesh order (basis.py:171)."""
    if DIM[ele_type] == 2:
        return _BASIX_QUAD_VERTS[RE_ORDER[ele_type]]
    return basix_node_coords(DEGREE[ele_type])[RE_ORDER[ele_type]]


def gauss_legendre_01(m):
    """m-point Gauss-Legendre rule on [0,1], points ascending, weights sum to 1."""
    x, w = np.polynomial.legendre.leggauss(m)
    return 0.5 * (x + 1.0), 0.5 * w


def points_per_direction(order):
    """Gauss-Jacobi points per direction for a given polynomial order (order+2)//2."""
    return (order + 2) // 2


def hex_quadrature(order):
    """Tensor Gauss rule on [0,1]^3; point index = (i*m + j)*m + k with x from i."""
    x1, w1 = gauss_legendre_01(points_per_direction(order))
    X = np.stack(np.meshgrid(x1, x1, x1, indexing="ij"), axis=-1).reshape(-1, 3)
    W = (w1[:, None, None] * w1[None, :, None] * w1[None, None, :]).reshape(-1)
    return X, W


def quad_quadrature(order):
    x1, w1 = gauss_legendre_01(points_per_direction(order))
    X = np.stack(np.meshgrid(x1, x1, indexing="ij"), axis=-1).reshape(-1, 2)
    W = (w1[:, None] * w1[None, :]).reshape(-1)
    return X, W


def _lagrange_1d(nodes, x):
    """Values and derivatives of the 1-D Lagrange polynomials through `nodes` at x."""
    n = len(nodes)
    x = np.asarray(x, dtype=np.float64)
    vals = np.ones((len(x), n))
    ders = np.zeros((len(x), n))
    for a in range(n):
        for b in range(n):
            if b == a:
                continue
            vals[:, a] *= (x - nodes[b]) / (nodes[a] - nodes[b])
        for c in range(n):
            if c == a:
                continue
            term = np.full(len(x), 1.0 / (nodes[a] - nodes[c]))
            for b in range(n):
                if b == a or b == c:
                    continue
                term *= (x - nodes[b]) / (nodes[a] - nodes[b])
            ders[:, a] += term
    return vals, ders


def tabulate(ele_type, X):
    """Shape values (np, nn) and reference gradients (np, nn, dim) at points X."""
    deg = DEGREE[ele_type]
    dim = DIM[ele_type]
    nodes_1d = np.linspace(0.0, 1.0, deg + 1)
    nc = node_coords(ele_type)
    idx = np.rint(nc * deg).astype(int)           # 1-D node index per direction
    v = []
    d = []
    for c in range(dim):
        vc, dc = _lagrange_1d(nodes_1d, X[:, c])
        v.append(vc[:, idx[:, c]])
        d.append(dc[:, idx[:, c]])
    vals = np.ones_like(v[0])
    for c in range(dim):
        vals = vals * v[c]
    grads = []
    for e in range(dim):
        g = np.ones_like(v[0])
        for c in range(dim):
            g = g * (d[c] if c == e else v[c])
        grads.append(g)
    return vals, np.stack(grads, axis=-1)


def get_shape_vals_and_grads(ele_type, quadrature_order=None):
    """Restates basis.py:141-175 -> (shape_vals (nq,nn), shape_grads_ref (nq,nn,dim), weights (nq))."""
    order = DEFAULT_ORDER[ele_type] if quadrature_order is None else quadrature_order
    X, W = hex_quadrature(order) if DIM[ele_type] == 3 else quad_quadrature(order)
    vals, grads = tabulate(ele_type, X)
    return vals, grads, W


def get_face_shape_vals_and_grads(ele_type, quadrature_order=None):
    """Restates basis.py:178-250.

    Returns face_shape_vals (6,nfq,nn), face_shape_grads_ref (6,nfq,nn,3),
    face_weights (6,nfq), face_normals (6,3), face_inds (6,4) in mesh numbering.
    """
    order = DEFAULT_ORDER[ele_type] if quadrature_order is None else quadrature_order
    if DIM[ele_type] == 2:
        return _quad_facets(ele_type, order)
    P2, W2 = quad_quadrature(order)
    # P1 map of the reference quadrilateral (basix vertex order (0,0),(1,0),(0,1),(1,1)).
    s, t = P2[:, 0], P2[:, 1]
    lin = np.stack([(1 - s) * (1 - t), s * (1 - t), (1 - s) * t, s * t], axis=-1)
    re = np.array(RE_ORDER["HEX8"])
    face_pts, face_w, face_inds = [], [], []
    for f, facet in enumerate(_BASIX_FACETS):
        V = _BASIX_VERTS[facet]
        face_pts.append(lin @ V)
        e0 = V[1] - V[0]
        e1 = V[2] - V[0]
        face_w.append(W2 * np.linalg.norm(np.cross(e0, e1)))
        # basix vertex number -> mesh vertex number (reorder_inds, basis.py:117-124)
        face_inds.append([int(np.argwhere(re == v)[0, 0]) for v in facet])
    face_pts = np.stack(face_pts)
    nf, nfq, _ = face_pts.shape
    vals, grads = tabulate(ele_type, face_pts.reshape(-1, 3))
    nn = vals.shape[1]
    return (vals.reshape(nf, nfq, nn), grads.reshape(nf, nfq, nn, 3),
            np.stack(face_w), _BASIX_FACET_NORMALS.copy(), np.array(face_inds))


def _quad_facets(ele_type, order):
    """Interval facets of the quadrilateral: Gauss points on [0,1] mapped from the
    facet's first basix vertex to its second, weights scaled by the edge length."""
    x1, w1 = gauss_legendre_01(points_per_direction(order))
    re = np.array(RE_ORDER[ele_type])
    face_pts, face_w, face_inds = [], [], []
    for facet in _BASIX_QUAD_FACETS:
        V = _BASIX_QUAD_VERTS[facet]
        face_pts.append(V[0][None] + x1[:, None] * (V[1] - V[0])[None])
        face_w.append(w1 * np.linalg.norm(V[1] - V[0]))
        face_inds.append([int(np.argwhere(re == v)[0, 0]) for v in facet])
    face_pts = np.stack(face_pts)
    nf, nfq, _ = face_pts.shape
    vals, grads = tabulate(ele_type, face_pts.reshape(-1, 2))
    nn = vals.shape[1]
    return (vals.reshape(nf, nfq, nn), grads.reshape(nf, nfq, nn, 2),
            np.stack(face_w), _BASIX_QUAD_FACET_NORMALS.copy(), np.array(face_inds))
