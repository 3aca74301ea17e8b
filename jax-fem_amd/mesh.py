"""Mesh container and structured box generators.

``Mesh`` mirrors jax_fem/generate_mesh.py:13-27 (points (nnode, dim), cells
(ncell, nn)).  ``rectangle_mesh`` reproduces generate_mesh.py:120-148 (node
index i*(Ny+1)+j, QUAD4 corners counter-clockwise).  ``box_mesh`` reproduces generate_mesh.py:151-189 (node index
(i*(Ny+1)+j)*(Nz+1)+k, HEX8 corners in VTK order) and returns a ``Mesh``
directly (the reference wraps it in a meshio object).  ``box_mesh_tet`` /
``rectangle_mesh_tri`` / ``box_mesh_gmsh`` are structured simplex and gmsh-free
counterparts of the reference's gmsh boxes; ``box_mesh_hex27`` is
the structured HEX27 counterpart (the reference only has a gmsh path,
generate_mesh.py:192-262): nodes on the doubled lattice, cell nodes in the
VTK triquadratic order that the reference's HEX27 re_order maps to basix
(basis.py:60-61).
"""
import numpy as np

# VTK triquadratic-hexahedron node offsets on the doubled lattice (units of h/2):
# 8 corners, 12 edge mid-points, 6 face centres (x-, x+, y-, y+, z-, z+), centre.
HEX27_OFFSETS = np.array([
    [0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2],
    [1, 0, 0], [2, 1, 0], [1, 2, 0], [0, 1, 0], [1, 0, 2], [2, 1, 2], [1, 2, 2], [0, 1, 2],
    [0, 0, 1], [2, 0, 1], [2, 2, 1], [0, 2, 1],
    [0, 1, 1], [2, 1, 1], [1, 0, 1], [1, 2, 1], [1, 1, 0], [1, 1, 2], [1, 1, 1]])
HEX8_OFFSETS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                         [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
# VTK quadrilateral: counter-clockwise corners
QUAD4_OFFSETS = np.array([[0, 0], [1, 0], [1, 1], [0, 1]])


_CELL_TYPES = {"HEX8": "hexahedron", "HEX27": "hexahedron27", "HEX20": "hexahedron20", "TET4": "tetra",
               "TET10": "tetra10", "QUAD4": "quad", "QUAD8": "quad8", "TRI3": "triangle", "TRI6": "triangle6"}
_NN_TO_TYPE = {8: "HEX8", 27: "HEX27", 20: "HEX20", 10: "TET10", 3: "TRI3"}


class Mesh:
    """points (nnode, dim) float64, cells (ncell, nn) int, optional ele_type.

    Also the reading side of a meshio mesh as the reference's generators return
    it (generate_mesh.py:120-189): ``cells_dict`` maps the meshio cell-type name to
    the connectivity, so ``Mesh(m.points, m.cells_dict[get_meshio_cell_type(e)])``
    works on the meshes made here.  ``node_grid`` is the structured node lattice
    (nx, ny, nz) when the points are one, in box_mesh order
    n = (i*ny + j)*nz + k (set by the box generators, otherwise detected from the
    coordinates), else None."""

    def __init__(self, points, cells, ele_type=None):
        self.points = np.asarray(points, dtype=np.float64)
        self.cells = np.asarray(cells)
        self.ele_type = ele_type
        self.point_data = {}

    @property
    def cells_dict(self):
        cd = self.__dict__.get("_cells_dict")
        if cd is not None:
            return cd
        nn, d = self.cells.shape[1], self.points.shape[1]
        e = self.ele_type or {4: "TET4" if d == 3 else "QUAD4", 6: "TRI6"}.get(nn) or _NN_TO_TYPE[nn]
        return {_CELL_TYPES[e]: self.cells}

    @property
    def node_grid(self):
        if "_node_grid" not in self.__dict__:
            self._node_grid = _detect_lattice(self.points)
        return self._node_grid

    @node_grid.setter
    def node_grid(self, g):
        self._node_grid = g

    def count_selected_faces(self, location_fn):
        from .fe import FiniteElement
        fe = FiniteElement(self, vec=1, dim=self.points.shape[1], ele_type=self.ele_type or "HEX8")
        return len(fe.get_boundary_conditions_inds([location_fn])[0])


def _detect_lattice(points):
    """(nx, ny, nz) when the points are the tensor lattice of their sorted unique
    coordinates in box_mesh order (x slowest, z fastest), else None."""
    if points.ndim != 2 or points.shape[1] != 3 or len(points) < 8:
        return None
    ax = [np.unique(points[:, d]) for d in range(3)]
    dims = tuple(len(a) for a in ax)
    if int(np.prod(dims)) != len(points) or min(dims) < 2:
        return None
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    ref = np.stack((X, Y, Z), axis=-1).reshape(-1, 3)
    scale = max(1.0, float(np.abs(points).max()))
    if not np.allclose(points, ref, rtol=0.0, atol=1e-12 * scale):
        return None
    return dims


def _lattice(n, L):
    return np.linspace(0.0, L, n + 1)


def rectangle_mesh(Nx, Ny, domain_x=1.0, domain_y=1.0):
    """QUAD4 rectangle with Nx x Ny cells (generate_mesh.py:120-148); like the
    reference's meshio object it has ``points`` and ``cells_dict['quad']``."""
    X, Y = np.meshgrid(_lattice(Nx, domain_x), _lattice(Ny, domain_y), indexing="ij")
    points = np.stack((X, Y), axis=-1).reshape(-1, 2)
    idx = np.arange(len(points)).reshape(Nx + 1, Ny + 1)
    i, j = [a.reshape(-1) for a in np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")]
    o = QUAD4_OFFSETS
    cells = idx[i[:, None] + o[None, :, 0], j[:, None] + o[None, :, 1]]
    mesh = Mesh(points, cells.astype(np.int64), "QUAD4")
    mesh.cell_grid = (Nx, Ny)
    return mesh


def box_mesh(Nx, Ny, Nz, domain_x=1.0, domain_y=1.0, domain_z=1.0):
    """HEX8 box with Nx x Ny x Nz cells (generate_mesh.py:151-189)."""
    X, Y, Z = np.meshgrid(_lattice(Nx, domain_x), _lattice(Ny, domain_y), _lattice(Nz, domain_z),
                          indexing="ij")
    points = np.stack((X, Y, Z), axis=-1).reshape(-1, 3)
    idx = np.arange(len(points)).reshape(Nx + 1, Ny + 1, Nz + 1)
    i, j, k = [a.reshape(-1) for a in np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")]
    o = HEX8_OFFSETS
    cells = idx[i[:, None] + o[None, :, 0], j[:, None] + o[None, :, 1], k[:, None] + o[None, :, 2]]
    mesh = Mesh(points, cells.astype(np.int64), "HEX8")
    mesh.node_grid = (Nx + 1, Ny + 1, Nz + 1)   # structured node lattice, node n = (i*(Ny+1)+j)*(Nz+1)+k
    mesh.cell_grid = (Nx, Ny, Nz)
    return mesh


def box_mesh_hex27(Nx, Ny, Nz, domain_x=1.0, domain_y=1.0, domain_z=1.0):
    """HEX27 box with Nx x Ny x Nz cells on the (2N+1)^3 node lattice."""
    X, Y, Z = np.meshgrid(_lattice(2 * Nx, domain_x), _lattice(2 * Ny, domain_y), _lattice(2 * Nz, domain_z),
                          indexing="ij")
    points = np.stack((X, Y, Z), axis=-1).reshape(-1, 3)
    idx = np.arange(len(points)).reshape(2 * Nx + 1, 2 * Ny + 1, 2 * Nz + 1)
    i, j, k = [a.reshape(-1) for a in np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")]
    o = HEX27_OFFSETS
    cells = idx[2 * i[:, None] + o[None, :, 0], 2 * j[:, None] + o[None, :, 1], 2 * k[:, None] + o[None, :, 2]]
    mesh = Mesh(points, cells.astype(np.int64), "HEX27")
    mesh.node_grid = (2 * Nx + 1, 2 * Ny + 1, 2 * Nz + 1)   # node n = (i*(2Ny+1)+j)*(2Nz+1)+k
    mesh.cell_grid = (Nx, Ny, Nz)
    return mesh


def _freudenthal():
    """The 6 tetrahedra of a unit cube along its (0,0,0)-(1,1,1) diagonal, one per
    axis order, each positively oriented, as lists of 4 corner offsets."""
    import itertools
    tets = []
    for perm in itertools.permutations(range(3)):
        p = [np.zeros(3, dtype=int)]
        for ax in perm:
            q = p[-1].copy()
            q[ax] = 1
            p.append(q)
        if np.linalg.det(np.stack([p[1] - p[0], p[2] - p[0], p[3] - p[0]])) < 0:
            p[1], p[2] = p[2], p[1]
        tets.append(p)
    return tets


def box_mesh_tet(Nx, Ny, Nz, domain_x=1.0, domain_y=1.0, domain_z=1.0, ele_type="TET4"):
    """Structured TET4 / TET10 box: every cell of the Nx x Ny x Nz lattice split into 6
    tetrahedra along its main diagonal (a conforming Freudenthal triangulation),
    vertices and, for TET10, mid-edge nodes (VTK order) on the doubled lattice.
    The reference builds these boxes with gmsh (generate_mesh.py:192-262); this is a
    structured equivalent (same domain and element type, different triangulation)."""
    if ele_type not in ("TET4", "TET10"):
        raise ValueError(ele_type)
    f = 1 if ele_type == "TET4" else 2
    X, Y, Z = np.meshgrid(_lattice(f * Nx, domain_x), _lattice(f * Ny, domain_y), _lattice(f * Nz, domain_z),
                          indexing="ij")
    points = np.stack((X, Y, Z), axis=-1).reshape(-1, 3)
    idx = np.arange(len(points)).reshape(f * Nx + 1, f * Ny + 1, f * Nz + 1)
    i, j, k = [a.reshape(-1) for a in np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")]
    cells = []
    for p in _freudenthal():
        verts = [f * np.stack([i + v[0], j + v[1], k + v[2]], -1) for v in p]     # doubled-lattice coords
        nodes = list(verts)
        if f == 2:
            for a, b in ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)):
                nodes.append((verts[a] + verts[b]) // 2)
        cells.append(np.stack([idx[n[:, 0], n[:, 1], n[:, 2]] for n in nodes], axis=1))
    cells = np.stack(cells, axis=1).reshape(-1, 4 if f == 1 else 10)
    return Mesh(points, cells.astype(np.int64), ele_type)


def rectangle_mesh_tri(Nx, Ny, domain_x=1.0, domain_y=1.0, ele_type="TRI3"):
    """Structured TRI3 / TRI6 rectangle: every cell split along its (0,0)-(1,1)
    diagonal into two counter-clockwise triangles; TRI6 mid-edge nodes (VTK order)
    on the doubled lattice."""
    if ele_type not in ("TRI3", "TRI6"):
        raise ValueError(ele_type)
    f = 1 if ele_type == "TRI3" else 2
    X, Y = np.meshgrid(_lattice(f * Nx, domain_x), _lattice(f * Ny, domain_y), indexing="ij")
    points = np.stack((X, Y), axis=-1).reshape(-1, 2)
    idx = np.arange(len(points)).reshape(f * Nx + 1, f * Ny + 1)
    i, j = [a.reshape(-1) for a in np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")]
    cells = []
    for tri in (((0, 0), (1, 0), (1, 1)), ((0, 0), (1, 1), (0, 1))):
        verts = [f * np.stack([i + v[0], j + v[1]], -1) for v in tri]
        nodes = list(verts)
        if f == 2:
            for a, b in ((0, 1), (1, 2), (0, 2)):
                nodes.append((verts[a] + verts[b]) // 2)
        cells.append(np.stack([idx[n[:, 0], n[:, 1]] for n in nodes], axis=1))
    cells = np.stack(cells, axis=1).reshape(-1, 3 if f == 1 else 6)
    return Mesh(points, cells.astype(np.int64), ele_type)


def box_mesh_gmsh(Nx, Ny, Nz, domain_x, domain_y, domain_z, data_dir=None, ele_type="HEX8"):
    """The reference's box_mesh_gmsh (generate_mesh.py:192-262) without gmsh: the
    structured box of the element type (HEX8 box_mesh, HEX27 box_mesh_hex27,
    TET4/TET10 box_mesh_tet), returned as a Mesh with meshio's cells_dict; nothing
    is written to data_dir."""
    if ele_type == "HEX8":
        return box_mesh(Nx, Ny, Nz, domain_x, domain_y, domain_z)
    if ele_type == "HEX27":
        return box_mesh_hex27(Nx, Ny, Nz, domain_x, domain_y, domain_z)
    if ele_type in ("TET4", "TET10"):
        return box_mesh_tet(Nx, Ny, Nz, domain_x, domain_y, domain_z, ele_type)
    raise NotImplementedError("box_mesh_gmsh: element %s" % ele_type)


def get_meshio_cell_type(ele_type):
    """meshio's name of the cell block of an element type (generate_mesh.py:89-117)."""
    if ele_type not in _CELL_TYPES:
        raise NotImplementedError(ele_type)
    return _CELL_TYPES[ele_type]
