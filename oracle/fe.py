"""Meshes, geometry, Dirichlet data and boundary-face selection (oracle).

Restates:
  * ``box_mesh`` (generate_mesh.py:151-189): structured HEX8 grid, node index
    (i*(Ny+1) + j)*(Nz+1) + k, cell corners in VTK order; ``rectangle_mesh``
    (generate_mesh.py:120-148), its QUAD4 counterpart.
  * a structured HEX27 grid in the same style (no reference counterpart: the
    reference builds HEX27 boxes with gmsh, generate_mesh.py:192-262); nodes on
    the doubled lattice, cell nodes in the VTK/mesh order of ``basis.node_coords``.
  * ``FiniteElement.get_shape_grads`` (fe.py:112-141): J = sum_a x_a (x) dN_a/dxi,
    dN/dx = dN/dxi . J^-1, JxW = det(J) * w (no abs).
  * ``get_face_shape_grads`` (fe.py:143-183): Nanson scale ||n . J^-1|| det(J) w_face.
  * ``get_physical_quad_points`` / ``get_physical_surface_quad_points`` (fe.py:185-218).
  * ``Dirichlet_boundary_conditions`` (fe.py:220-258): node_inds = argwhere(loc(x)),
    vec_inds = const, vals = value_fn(x).
  * ``get_boundary_conditions_inds`` (fe.py:271-322): faces whose every vertex
    satisfies the location function, as argwhere over (cell, local face).

Pinned by: the cylinder top-face area 76.536686473018 (fixture), unit-cube face
areas, volume = sum JxW, and end-to-end by the FEniCSx fixtures.
"""
import numpy as np

from . import basis


def box_mesh(Nx, Ny, Nz, Lx=1.0, Ly=1.0, Lz=1.0):
    """HEX8 box; Nx,Ny,Nz are cell counts (generate_mesh.py:151-189)."""
    x = np.linspace(0, Lx, Nx + 1)
    y = np.linspace(0, Ly, Ny + 1)
    z = np.linspace(0, Lz, Nz + 1)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    points = np.stack((X, Y, Z), axis=3).reshape(-1, 3)
    ids = np.arange(len(points)).reshape(Nx + 1, Ny + 1, Nz + 1)
    c = [ids[:-1, :-1, :-1], ids[1:, :-1, :-1], ids[1:, 1:, :-1], ids[:-1, 1:, :-1],
         ids[:-1, :-1, 1:], ids[1:, :-1, 1:], ids[1:, 1:, 1:], ids[:-1, 1:, 1:]]
    cells = np.stack(c, axis=3).reshape(-1, 8)
    return points, cells


def rectangle_mesh(Nx, Ny, Lx=1.0, Ly=1.0):
    """QUAD4 rectangle (generate_mesh.py:120-148): node index i*(Ny+1) + j,
    corners (i,j), (i+1,j), (i+1,j+1), (i,j+1)."""
    x = np.linspace(0, Lx, Nx + 1)
    y = np.linspace(0, Ly, Ny + 1)
    X, Y = np.meshgrid(x, y, indexing="ij")
    points = np.stack((X, Y), axis=2).reshape(-1, 2)
    ids = np.arange(len(points)).reshape(Nx + 1, Ny + 1)
    c = [ids[:-1, :-1], ids[1:, :-1], ids[1:, 1:], ids[:-1, 1:]]
    return points, np.stack(c, axis=2).reshape(-1, 4)


def box_mesh_hex27(Nx, Ny, Nz, Lx=1.0, Ly=1.0, Lz=1.0):
    """HEX27 box on the (2Nx+1)(2Ny+1)(2Nz+1) lattice, cells in mesh node order."""
    x = np.linspace(0, Lx, 2 * Nx + 1)
    y = np.linspace(0, Ly, 2 * Ny + 1)
    z = np.linspace(0, Lz, 2 * Nz + 1)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    points = np.stack((X, Y, Z), axis=3).reshape(-1, 3)
    ids = np.arange(len(points)).reshape(2 * Nx + 1, 2 * Ny + 1, 2 * Nz + 1)
    off = np.rint(basis.node_coords("HEX27") * 2).astype(int)      # (27,3) in {0,1,2}
    ci, cj, ck = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
    ci, cj, ck = ci.reshape(-1), cj.reshape(-1), ck.reshape(-1)
    cells = ids[2 * ci[:, None] + off[None, :, 0],
                2 * cj[:, None] + off[None, :, 1],
                2 * ck[:, None] + off[None, :, 2]]
    return points, cells


def shape_grads(points, cells, shape_grads_ref, weights):
    """fe.py:112-141 -> (ncell,nq,nn,3) physical gradients, (ncell,nq) JxW."""
    X = points[cells]                                             # (c,n,d)
    Jm = np.einsum("cnd,qne->cqde", X, shape_grads_ref)           # dx_d/dxi_e
    det = np.linalg.det(Jm)
    inv = np.linalg.inv(Jm)                                       # dxi_d/dx_e
    grads = np.einsum("qnd,cqde->cqne", shape_grads_ref, inv)
    return grads, det * weights[None, :]


def face_shape_grads(points, cells, boundary_inds, face_shape_grads_ref, face_normals, face_weights):
    """fe.py:143-183 -> face gradients (nsel,nfq,nn,3), nanson_scale (nsel,nfq)."""
    X = points[cells][boundary_inds[:, 0]]
    g = face_shape_grads_ref[boundary_inds[:, 1]]
    Jm = np.einsum("snd,sqne->sqde", X, g)
    det = np.linalg.det(Jm)
    inv = np.linalg.inv(Jm)
    grads = np.einsum("sqnd,sqde->sqne", g, inv)
    n = face_normals[boundary_inds[:, 1]]
    nJ = np.einsum("sd,sqde->sqe", n, inv)
    nanson = np.linalg.norm(nJ, axis=-1) * det * face_weights[boundary_inds[:, 1]]
    return grads, nanson


def physical_quad_points(points, cells, shape_vals):
    return np.einsum("qn,cnd->cqd", shape_vals, points[cells])


def physical_surface_quad_points(points, cells, boundary_inds, face_shape_vals):
    return np.einsum("sqn,snd->sqd", face_shape_vals[boundary_inds[:, 1]],
                     points[cells][boundary_inds[:, 0]])


def _call_loc(fn, points):
    """Vectorised location function; accepts f(x) or f(x, ind) (fe.py:244-250)."""
    n = fn.__code__.co_argcount
    idx = np.arange(len(points))
    if n == 1:
        return np.array([bool(fn(p)) for p in points])
    if n == 2:
        return np.array([bool(fn(p, i)) for p, i in zip(points, idx)])
    raise ValueError("Wrong number of arguments for location_fn: must be 1 or 2, get %d" % n)


def dirichlet_bcs(points, dirichlet_bc_info):
    """fe.py:220-258 -> node_inds_list, vec_inds_list, vals_list."""
    node_inds_list, vec_inds_list, vals_list = [], [], []
    if dirichlet_bc_info is None:
        return node_inds_list, vec_inds_list, vals_list
    location_fns, vecs, value_fns = dirichlet_bc_info
    assert len(location_fns) == len(value_fns) and len(value_fns) == len(vecs)
    for i in range(len(location_fns)):
        node_inds = np.argwhere(_call_loc(location_fns[i], points)).reshape(-1)
        vec_inds = np.full(len(node_inds), vecs[i], dtype=np.int32)
        vals = np.array([float(value_fns[i](p)) for p in points[node_inds]], dtype=np.float64)
        node_inds_list.append(node_inds)
        vec_inds_list.append(vec_inds)
        vals_list.append(vals)
    return node_inds_list, vec_inds_list, vals_list


def boundary_inds(points, cells, location_fns, face_inds):
    """fe.py:271-322 -> list of (nsel,2) [cell, local face] arrays in argwhere order."""
    out = []
    if location_fns is None:
        return out
    for fn in location_fns:
        flags_node = _call_loc(fn, points)                         # per node
        face_nodes = cells[:, face_inds]                           # (c,6,4)
        flags = np.all(flags_node[face_nodes], axis=-1)
        out.append(np.argwhere(flags))
    return out
