"""``Problem``: the reference's weak-form class, driving libhipfem on one GPU.

Mirrors jax_fem/problem.py:15-611 for one variable (vec = 1 or 3, HEX8 or
HEX27): same constructor, same user hooks (``get_tensor_map``,
``get_mass_map``, ``get_surface_maps``, ``custom_init``, ``set_params``) and the
engine methods ``compute_residual`` (:462) and ``newton_update`` (:477).

Differences by design:
  * ``get_tensor_map()`` returns a ``jax_fem_amd.materials`` law (callable like
    the reference's map) so that the cell kernels know the law and its tangent;
    a plain Python function raises ``NotImplementedError``.
  * Mass and surface maps must not depend on u (they contribute a load vector
    and no Jacobian); they are evaluated once per set-up or parameter change
    (problem.py:216-259 evaluates them inside the element loop).
  * ``newton_update`` assembles the tangent straight into a device BSR matrix
    (3x3 node-pair blocks) instead of host COO values ``V``; ``problem.V``,
    ``problem.I``, ``problem.J`` are still available on request (host copies).
State on the device (torch tensors): points (nnode,3) f64, cells (ncell,nn)
i32, per-quadrature parameters (ncell,nq) f64, Dirichlet DOFs/values, the load
vector, the sparsity pattern (built once by a device radix sort) and the
element/BSR buffers reused across Newton iterations.
"""
import ctypes
from dataclasses import dataclass
from typing import Any

import numpy as np
import torch

from . import _lib
from .fe import FiniteElement, _eval_points
from .materials import Material
from .mesh import Mesh


def _fingerprint(obj):
    """Content key of a nest of lists/tuples of arrays: torch tensors by (storage
    pointer, version counter, shape) -- an in-place edit bumps the version --,
    NumPy arrays and scalars by a digest of their bytes."""
    import hashlib
    if obj is None:
        return None
    if isinstance(obj, (list, tuple)):
        return tuple(_fingerprint(o) for o in obj)
    if torch.is_tensor(obj):
        return ("t", obj.data_ptr(), obj._version, tuple(obj.shape), str(obj.dtype), str(obj.device))
    a = np.ascontiguousarray(np.asarray(obj))
    if a.dtype == object:
        return ("o", id(obj))
    return ("a", a.shape, a.dtype.str, hashlib.blake2b(a.view(np.uint8).reshape(-1), digest_size=16).digest())


@dataclass
class Problem:
    mesh: Mesh
    vec: int
    dim: int
    ele_type: str = "HEX8"
    quadrature_rule: Any = None
    quadrature_order: Any = None
    dirichlet_bc_info: Any = None
    location_fns: Any = None
    additional_info: Any = ()
    device: Any = None

    def __post_init__(self):
        multi = isinstance(self.mesh, list) and len(self.mesh) > 1
        if isinstance(self.mesh, list) and not multi:
            self.mesh, self.vec, self.ele_type = self.mesh[0], self.vec[0], self.ele_type[0]
            qo = self.quadrature_order
            self.quadrature_order = qo[0] if isinstance(qo, list) else qo
            dbc = self.dirichlet_bc_info
            self.dirichlet_bc_info = dbc[0] if isinstance(dbc, list) and dbc and isinstance(dbc[0], list) \
                and len(dbc) == 1 else dbc
        if self.device is None:
            self.device = torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() \
                else torch.device("cpu")
        if multi:
            # several variables (problem.py:55-128): one FiniteElement each on the same cells
            nv = len(self.mesh)
            per = lambda a: a[:] if isinstance(a, list) else [a] * nv
            self.vec, self.ele_type = per(self.vec), per(self.ele_type)
            qos, dbcs = per(self.quadrature_order), per(self.dirichlet_bc_info)
            self.fes = [FiniteElement(self.mesh[i], self.vec[i], self.dim, self.ele_type[i], self.quadrature_rule,
                                      qos[i], dbcs[i]) for i in range(nv)]
            if len({fe.num_cells for fe in self.fes}) != 1 or len({fe.num_quads for fe in self.fes}) != 1:
                raise ValueError("the variables of a problem must share the cells and the quadrature points")
        else:
            self.fes = [FiniteElement(self.mesh, self.vec, self.dim, self.ele_type, self.quadrature_rule,
                                      self.quadrature_order, self.dirichlet_bc_info)]
        self.fe = self.fes[0]
        self.num_vars = len(self.fes)
        self.cells_list = [fe.cells for fe in self.fes]
        self.num_cells = self.fe.num_cells
        self.offset = [0]
        for fe in self.fes[:-1]:
            self.offset.append(self.offset[-1] + fe.num_total_dofs)
        self.num_total_dofs_all_vars = sum(fe.num_total_dofs for fe in self.fes)
        self.num_nodes_cumsum = np.cumsum([0] + [fe.num_nodes for fe in self.fes])
        # each cell's DOFs, variable by variable, node-major (the reference's ravel_pytree order)
        self._dof_cells = np.hstack([(o + fe.vec * fe.cells[:, :, None] + np.arange(fe.vec)[None, None, :])
                                     .reshape(fe.num_cells, -1) for o, fe in zip(self.offset, self.fes)])
        self.boundary_inds_list = self.fe.get_boundary_conditions_inds(self.location_fns)
        self.internal_vars = ()
        self.internal_vars_surfaces = [() for _ in self.boundary_inds_list]
        self._pattern = None
        self._bufs = {}
        self._ext_key = None
        self._bc_key = None
        self._bc_struct = None     # node arrays + components the BC DOF set was built from
        self.custom_init(*self.additional_info)
        self._check_hooks()

    # -------------------------------------------------- layout of the unknowns
    @property
    def _multi(self):
        return self.num_vars > 1

    @property
    def _enn(self):
        """Entities per cell in the device pattern: nodes, or DOFs for several variables."""
        return self._dof_cells.shape[1] if self._multi else self.fe.num_nodes

    @property
    def _evec(self):
        """Components per pattern entity: vec, or 1 (DOF-level pattern) for several variables."""
        return 1 if self._multi else self.vec

    def unflatten_fn_dof(self, cell_flat):
        """A cell's flat DOF vector -> [(num_nodes_i, vec_i), ...] (problem.py:117-118);
        torch or NumPy, also under vmap."""
        out, k = [], 0
        for fe in self.fes:
            n = fe.num_nodes * fe.vec
            out.append(cell_flat[..., k:k + n].reshape(tuple(cell_flat.shape[:-1]) + (fe.num_nodes, fe.vec)))
            k += n
        return out

    # ------------------------------------------------------------------ hooks
    def custom_init(self, *args):
        pass

    def set_params(self, params):
        raise NotImplementedError("Child class must implement this function!")

    def _check_hooks(self):
        """Collect the user hooks (problem.py:277-346).  A tensor map that is a
        jax_fem_amd.materials law runs on the built cell kernels; a plain function is
        probed at its first evaluation (a built law it equals runs on those kernels)
        and otherwise evaluated pointwise on the device (generic.PointwiseTerms)."""
        nsurf = len(self.boundary_inds_list)
        if hasattr(self, "get_surface_maps"):
            assert nsurf == len(self.get_surface_maps())
        elif hasattr(self, "get_universal_kernels_surface"):
            assert nsurf == len(self.get_universal_kernels_surface())
        else:
            assert nsurf == 0, "Missing definitions for surface integral"
        self.material = None
        self._tensor_fn = None
        if hasattr(self, "get_tensor_map"):
            m = self.get_tensor_map()
            if isinstance(m, Material):
                if m.vec != self.fe.vec:
                    raise ValueError("material expects vec=%d, problem has vec=%d" % (m.vec, self.fe.vec))
                self.material = m
            elif callable(m):
                self._tensor_fn = m
            else:
                raise TypeError("get_tensor_map() must return a callable")
        elif not (hasattr(self, "get_universal_kernel") or hasattr(self, "get_mass_map")):
            raise NotImplementedError("the problem defines no cell term (tensor map, mass map or universal kernel)")
        self._plan_cache = None

    # ------------------------------------------------------------ the plan
    probe_laws = True   # recognise a plain-function tensor map as a built law when it equals one

    def cell_internal_vars(self):
        """internal_vars as device FP64 tensors in the shapes given, each with leading
        dimension num_cells: what a universal kernel receives per cell (problem.py:
        335-357 vmap it over the cells unchanged: per-cell node values (ncell, nn, vec),
        per-cell scalars (ncell,), per-point fields (ncell, nq, ...)).  Held for a whole solve."""
        held = getattr(self, "_held_raw", None)
        if held is not None:
            return held
        out = []
        for a in (self.internal_vars or ()):
            t = torch.as_tensor(a.detach() if torch.is_tensor(a) else np.asarray(a), dtype=torch.float64,
                                device=self.device)
            if t.dim() == 0 or t.shape[0] != self.num_cells:
                raise ValueError("internal_vars entries must have leading dimension num_cells")
            out.append(t.contiguous())
        return out

    def volume_internal_vars(self):
        """internal_vars as device FP64 tensors of shape (ncell, nq, ...) (a per-cell
        array (ncell,) is repeated over the points): what tensor and mass maps receive
        per quadrature point.  Held for a whole solve."""
        held = getattr(self, "_held_ivs", None)
        if held is not None:
            return held
        out = []
        for t in self.cell_internal_vars():
            if t.dim() == 1:
                t = t[:, None].expand(self.num_cells, self.fe.num_quads)
            if tuple(t.shape[:2]) != (self.num_cells, self.fe.num_quads):
                raise ValueError("internal_vars entries of tensor and mass maps must have leading shape "
                                 "(num_cells, num_quads)")
            out.append(t.contiguous())
        return out

    def _plan(self):
        """How each term of the weak form is evaluated, decided once per internal
        variable layout: (law, mass, surfaces, universal).  law: 'fast' (built cell
        kernels) or 'generic'; mass: None, 'load' (u-independent: a load vector) or
        'cell'; surfaces: per surface 'load' or 'face' (u-dependent or universal)."""
        ivs = self.internal_vars or ()
        layout = tuple(tuple(np.shape(a)) for a in ivs)
        if self._plan_cache is not None and self._plan_cache[0] == layout:
            return self._plan_cache[1]
        from .materials import identify
        law = "generic"
        fn = self._tensor_fn
        if self.material is not None:
            law = "fast" if self._fast_supported(self.material) else "generic"
            fn = self.material            # a built law runs pointwise where its kernels do not apply
        elif self._tensor_fn is not None and self.probe_laws and not self._multi:
            scalar = len(layout) == 1 and len(layout[0]) == 2
            if len(layout) == 0 or scalar:
                m = identify(self._tensor_fn, self.vec, self.dim, scalar_param=scalar)
                if m is not None and self._fast_supported(m):
                    self.material = m
                    law = "fast"
                    import logging
                    logging.getLogger("jax_fem_amd").info(
                        "tensor map recognised as %s(E=%.6g, nu=%.6g%s): built cell kernels" % (
                            type(m).__name__, getattr(m, "E", 0.0), getattr(m, "nu", 0.0),
                            ", param=%s" % m.param if m.param else ""))
        if law == "fast" and (len(layout) > 1 or (len(layout) == 1 and len(layout[0]) not in (1, 2))):
            raise ValueError("a built law takes one scalar parameter per quadrature point")
        mass = None
        if hasattr(self, "get_mass_map"):
            mass = "cell" if (len(layout) > 0 or self._depends_on_u(self.get_mass_map(), "mass")) else "load"
        surfaces = []
        smaps = self.get_surface_maps() if hasattr(self, "get_surface_maps") else [None] * len(self.boundary_inds_list)
        for k, smap in enumerate(smaps):
            if hasattr(self, "get_universal_kernels_surface"):
                surfaces.append("face")
            elif smap is None:
                surfaces.append(None)
            else:
                surfaces.append("face" if self._depends_on_u(smap, "surface", k) else "load")
        plan = {"law": law, "fn": fn if law == "generic" else None, "mass": mass, "surfaces": surfaces,
                "universal": hasattr(self, "get_universal_kernel")}
        plan["pure_fast"] = law == "fast" and mass != "cell" and "face" not in surfaces and not plan["universal"]
        self._plan_cache = (layout, plan)
        self._terms = None
        return plan

    def _fast_supported(self, m):
        if self._multi or self.dim != 3 or self.fe.ele_type not in ("HEX8", "HEX27") or m.vec != self.vec:
            return False
        q = (int(self.fe.quadrature_order) + 2) // 2
        return (self.fe.ele_type == "HEX8" and q == 2) or (self.fe.ele_type == "HEX27" and q == 3)

    def _depends_on_u(self, fn, kind, k=None):
        """Probe a mass or surface map at a few points with random u against u = 0."""
        fe = self.fe
        rng = np.random.default_rng(1)
        if kind == "mass":
            x = fe.get_physical_quad_points().reshape(-1, self.dim)
            extra = []
        else:
            bi = self.boundary_inds_list[k]
            if len(bi) == 0:
                return False
            x = fe.get_physical_surface_quad_points(bi).reshape(-1, self.dim)
            extra = [np.asarray(a.detach().cpu() if torch.is_tensor(a) else a).reshape(len(x), -1)
                     for a in (self.internal_vars_surfaces[k] or ())]
        idx = sorted({(len(x) - 1) * j // 3 for j in range(4)})
        for i in idx:
            e = [a[i].reshape(np.shape(a[i])) for a in extra]
            try:
                y0 = np.asarray(fn(np.zeros(fe.vec), x[i], *e), dtype=np.float64)
                y1 = np.asarray(fn(rng.standard_normal(fe.vec), x[i], *e), dtype=np.float64)
            except Exception:
                return True     # not evaluable on NumPy arrays: take the general (device) path
            if not np.array_equal(y0, y1):
                return True
        return False

    def _pointwise(self):
        """The device evaluator of the problem's pointwise maps (built on first use)."""
        plan = self._plan()
        if getattr(self, "_terms", None) is None:
            from .generic import PointwiseTerms
            t = PointwiseTerms(self)
            t.tensor_fn = plan["fn"]
            t.mass_fn = self.get_mass_map() if plan["mass"] == "cell" else None
            t.universal = self.get_universal_kernel() if plan["universal"] else None
            fe = self.fe
            smaps = self.get_surface_maps() if hasattr(self, "get_surface_maps") else None
            usurf = self.get_universal_kernels_surface() if hasattr(self, "get_universal_kernels_surface") else None
            for k, (bi, how) in enumerate(zip(self.boundary_inds_list, plan["surfaces"])):
                if how != "face" or len(bi) == 0:
                    t.faces.append(None)
                    continue
                grads, nanson = fe.get_face_shape_grads(bi)
                dev = self.device
                T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
                isv = []
                for a in (self.internal_vars_surfaces[k] or ()):
                    isv.append(torch.as_tensor(a.detach() if torch.is_tensor(a) else np.asarray(a),
                                               dtype=torch.float64, device=dev))
                fd = {"cells": torch.as_tensor(bi[:, 0], dtype=torch.long, device=dev),
                      "cell_nodes": torch.as_tensor(fe.cells[bi[:, 0]], dtype=torch.long, device=dev),
                      "vals": T(fe.face_shape_vals[bi[:, 1]]), "grads": T(grads), "nanson": T(nanson),
                      "x": T(fe.get_physical_surface_quad_points(bi)), "isv": isv,
                      "nsel": len(bi), "nfq": fe.num_face_quads,
                      "map": smaps[k] if smaps is not None else None,
                      "universal": usurf[k] if usurf is not None else None}
                if self._multi:
                    # universal surface kernels see every variable (problem.py:153-182)
                    gn = [f.get_face_shape_grads(bi) for f in self.fes]
                    fd["sol_idx"] = torch.as_tensor(self._dof_cells[bi[:, 0]], dtype=torch.long, device=dev)
                    fd["vals_all"] = T(np.concatenate([np.asarray(f.face_shape_vals)[bi[:, 1]] for f in self.fes],
                                                      axis=2))
                    fd["grads_all"] = T(np.concatenate([np.asarray(g) for g, _ in gn], axis=2))
                    fd["nanson_all"] = T(np.stack([np.asarray(n) for _, n in gn], axis=1))
                t.faces.append(fd)
            self._terms = t
        return self._terms

    # ------------------------------------------- geometric quantities (problem.py:130-182)
    def _geometric(self):
        g = getattr(self, "_geom_q", None)
        if g is None:
            from .fe import _field
            fe = self.fe
            g = {"physical_quad_points": fe.get_physical_quad_points(),
                 "JxW": _field(np.asarray(fe.JxW)[:, None, :]), "shape_grads": fe.shape_grads,
                 "v_grads_JxW": fe.v_grads_JxW, "selected_face_shape_grads": [], "nanson_scale": [],
                 "selected_face_shape_vals": [], "physical_surface_quad_points": []}
            for bi in self.boundary_inds_list:
                grads, nanson = fe.get_face_shape_grads(bi)
                g["selected_face_shape_grads"].append(_field(grads))
                g["nanson_scale"].append(_field(np.asarray(nanson)[:, None, :]))
                g["selected_face_shape_vals"].append(_field(np.asarray(fe.face_shape_vals)[bi[:, 1]]))
                g["physical_surface_quad_points"].append(fe.get_physical_surface_quad_points(bi))
            self._geom_q = g
        return g

    @property
    def physical_quad_points(self):
        """(num_cells, num_quads, dim)"""
        return self._geometric()["physical_quad_points"]

    @property
    def physical_surface_quad_points(self):
        """per surface: (num_selected_faces, num_face_quads, dim)"""
        return self._geometric()["physical_surface_quad_points"]

    @property
    def selected_face_shape_grads(self):
        return self._geometric()["selected_face_shape_grads"]

    @property
    def selected_face_shape_vals(self):
        return self._geometric()["selected_face_shape_vals"]

    @property
    def nanson_scale(self):
        """per surface: (num_selected_faces, num_vars, num_face_quads)"""
        return self._geometric()["nanson_scale"]

    @property
    def JxW(self):
        """(num_cells, num_vars, num_quads)"""
        return self._geometric()["JxW"]

    @property
    def shape_grads(self):
        return self._geometric()["shape_grads"]

    @property
    def v_grads_JxW(self):
        return self._geometric()["v_grads_JxW"]

    # ---------------------------------------------------------------- layout
    def unflatten_fn_sol_list(self, dofs):
        """Global DOF vector -> [(num_total_nodes_i, vec_i), ...] (problem.py:120-122)."""
        out = []
        for o, fe in zip(self.offset, self.fes):
            out.append(dofs[o:o + fe.num_total_dofs].reshape(fe.num_total_nodes, fe.vec))
        return out

    def flatten_sol_list(self, sol_list):
        """[(num_total_nodes_i, vec_i), ...] (torch or NumPy) -> one device f64 DOF vector."""
        if torch.is_tensor(sol_list) or not isinstance(sol_list, (list, tuple)):
            return torch.as_tensor(sol_list, dtype=torch.float64).reshape(-1).to(self.device)
        return torch.cat([torch.as_tensor(s, dtype=torch.float64).reshape(-1).to(self.device) for s in sol_list])

    @property
    def nn(self):
        return self._enn

    @property
    def N(self):
        return self.num_total_dofs_all_vars

    # ---------------------------------------------------------- device state
    def _dev(self):
        if "points" not in self._bufs:
            if self.device.type != "cuda":
                raise RuntimeError("the jax_fem_amd solve path needs a GPU (torch.cuda is unavailable)")
            _lib.load()
            self._bufs["points"] = torch.as_tensor(self.fe.points, dtype=torch.float64, device=self.device).contiguous()
            pc = self._dof_cells if self._multi else self.fe.cells
            self._bufs["cells"] = torch.as_tensor(pc, dtype=torch.int32, device=self.device).contiguous()
            self._bufs["cells_var0"] = torch.as_tensor(self.fe.cells, dtype=torch.int32,
                                                       device=self.device).contiguous()
        return self._bufs["points"], self._bufs["cells"]

    def _var0_cells(self):
        """Node cells of the first variable on the device (its pointwise maps act on it)."""
        self._dev()
        return self._bufs["cells_var0"]

    @property
    def pattern(self):
        if self._pattern is None:
            _, cells = self._dev()
            h = ctypes.c_void_p()
            nent = self.N if self._multi else self.fe.num_total_nodes
            _lib.call("hf_pattern_build", _lib.ptr(cells), self.num_cells, self._enn, nent,
                      self._evec, _lib.stream(), ctypes.byref(h))
            self._pattern = Pattern(h, self)
        return self._pattern

    def _hfmat(self):
        """The material's C struct with the volume rule's Gauss points per direction."""
        mat = self.material.hf()
        mat.quad1d = (int(self.fe.quadrature_order) + 2) // 2
        return mat

    def _params(self):
        """The built laws' per-quadrature parameter (ncell, nq) f64 on the device, or
        None.  Inside a solve (hold_params) the tensor made at its start serves every
        evaluation: a NumPy field is copied to the device once per solve, not per
        residual."""
        held = getattr(self, "_held_params", None)
        if held is not None:
            return held[0]
        ivs = self.volume_internal_vars()
        if not ivs:
            return None
        if len(ivs) != 1 or ivs[0].dim() != 2:
            raise NotImplementedError("the built cell kernels take one scalar parameter per quadrature point")
        return ivs[0]

    def hold_params(self, on):
        """Freeze (on=True) or release the device copies of the internal variables
        for the duration of a solve; internal_vars must not change while they are held."""
        self._held_ivs = None
        self._held_raw = None
        self._held_params = None
        if on:
            plan = self._plan()
            self._held_raw = self.cell_internal_vars()
            if plan["law"] == "fast" or plan["fn"] is not None or plan["mass"] == "cell":
                self._held_ivs = self.volume_internal_vars()
            if plan["law"] == "fast":
                self._held_params = (self._params(),)

    def _buf(self, name, n, dtype):
        b = self._bufs.get(name)
        if b is None or b.numel() < n or b.dtype != dtype:
            b = torch.empty(n, dtype=dtype, device=self.device)
            self._bufs[name] = b
        return b[:n]

    # ------------------------------------------------------- Dirichlet data
    def bc_specs(self):
        """Every Dirichlet spec of every variable: (DOF offset, vec, node_inds, vec_inds, vals)."""
        return [(o, fe.vec, nodes, comps, vals) for o, fe in zip(self.offset, self.fes)
                for nodes, comps, vals in zip(fe.node_inds_list, fe.vec_inds_list, fe.vals_list)]

    def bc_arrays(self):
        """Unique BC DOFs (int32, device) and values (f64, device); later specs win (solver.py:298-300)."""
        specs = self.bc_specs()
        # the value arrays are made by FiniteElement (immutable in the reference, JAX
        # arrays): keyed by identity, with references kept so no id is reused;
        # torch tensors also by their version counter (an in-place edit bumps it)
        key = (tuple((id(v), v._version if torch.is_tensor(v) else -1) for _, _, _, _, v in specs),
               tuple(id(n) for _, _, n, _, _ in specs))
        if self._bc_key != key:
            # the DOF set depends on the node arrays and components only; a load step
            # that changes the values alone reuses it (location functions are cached
            # per function object, so their node arrays are the same objects)
            struct = [(nodes, int(np.asarray(comps).reshape(-1)[0]) if len(comps) else -1, len(comps), o)
                      for o, _, nodes, comps, _ in specs]
            same = self._bc_struct is not None and len(struct) == len(self._bc_struct) and all(
                a[0] is b[0] and a[1:] == b[1:] and np.all(np.asarray(sp[3]) == a[1])
                for a, b, sp in zip(struct, self._bc_struct, specs))
            if not same:
                ds = [np.zeros(0, dtype=np.int64)]
                for o, vec, nodes, comps, _ in specs:
                    ds.append(o + np.asarray(nodes, dtype=np.int64) * vec + np.asarray(comps, dtype=np.int64))
                all_d = np.concatenate(ds)
                # the last spec naming a DOF wins: first occurrence in reversed order; sorted DOFs
                d, first = np.unique(all_d[::-1], return_index=True)
                self._bc_first = first
                self._bc_dofs = torch.as_tensor(d.astype(np.int32), device=self.device)
                self._bc_struct = struct
            vs = [np.zeros(0)]
            for _, _, nodes, _, vals in specs:
                vals = vals.detach().cpu().numpy() if torch.is_tensor(vals) else vals
                vs.append(np.broadcast_to(np.asarray(vals, dtype=np.float64).reshape(-1), (len(nodes),)))
            v = np.concatenate(vs)[::-1][self._bc_first]
            self._bc_vals = torch.as_tensor(v, device=self.device)
            self._bc_key = key
            self._bc_refs = ([sp[4] for sp in specs], [sp[2] for sp in specs])   # ids in the key stay unique
        return self._bc_dofs, self._bc_vals

    # ------------------------------------------------------------ load vector
    def _external(self):
        """The u-independent mass and surface loads (problem.py:216-259) as one f64
        device vector, or None; the other terms are element terms (_element_terms)."""
        plan = self._plan()
        key = (_fingerprint(self.internal_vars_surfaces), tuple(plan["surfaces"]), plan["mass"])
        if self._ext_key == key:
            return self._fext
        self._ext_refs = self.internal_vars_surfaces   # keeps ids in the key unique
        fe = self.fe          # mass and surface maps act on the first variable (problem.py:216-259)
        vec = fe.vec
        f = np.zeros((fe.num_total_nodes, vec))
        have = False
        if plan["mass"] == "load":
            mmap = self.get_mass_map()
            x = np.asarray(fe.get_physical_quad_points()).reshape(-1, self.dim)
            m = self._eval_device(lambda xx: mmap(torch.zeros(vec, dtype=torch.float64, device=xx.device), xx),
                                  lambda xx: mmap(np.zeros(vec), xx), x)
            if m is None:
                m = _eval_points(lambda xx: mmap(np.zeros(vec), xx), x)
            m = m.reshape(self.num_cells, fe.num_quads, -1)
            r = np.einsum("cqi,qn,cq->cni", m, fe.shape_vals, fe.JxW)
            np.add.at(f, fe.cells.reshape(-1), r.reshape(-1, vec))
            have = True
        if hasattr(self, "get_surface_maps"):
            for k, (bi, smap) in enumerate(zip(self.boundary_inds_list, self.get_surface_maps())):
                if len(bi) == 0 or plan["surfaces"][k] != "load":
                    continue
                _, nanson = fe.get_face_shape_grads(bi)
                xs = fe.get_physical_surface_quad_points(bi).reshape(-1, self.dim)
                isv = [np.asarray(a.detach().cpu() if torch.is_tensor(a) else a).reshape(len(xs), -1)
                       for a in (self.internal_vars_surfaces[k] or ())]
                u0 = np.zeros((len(xs), vec))
                t = _eval_points(lambda xx, uu, *a: smap(uu, xx, *a), xs, (u0, *isv)).reshape(
                    len(bi), fe.num_face_quads, -1)
                vals = fe.face_shape_vals[bi[:, 1]]
                r = np.einsum("sqi,sqn,sq->sni", t, vals, nanson)
                np.add.at(f, fe.cells[bi[:, 0]].reshape(-1), r.reshape(-1, vec))
                have = True
        if have and self._multi:
            f = np.concatenate([f.reshape(-1), np.zeros(self.N - f.size)])
        self._fext = torch.as_tensor(f.reshape(-1), dtype=torch.float64, device=self.device) if have else None
        self._ext_key = key
        return self._fext

    def _eval_device(self, fn_t, fn_h, x, extra=()):
        """fn at every point, vmapped on the device (maps written with
        jax_fem_amd.numpy and its transforms), checked against host evaluations at a
        sample of points (1e-12 relative); None when the map does not run on tensors."""
        if self.device.type != "cuda" or len(x) == 0:
            return None
        from . import numpy as jnp
        from .generic import _vmap_chunks
        try:
            args = [torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=self.device)
                    for a in [x] + list(extra)]
            with jnp.on_device(self.device):
                out = _vmap_chunks(fn_t, args)
            out = out.detach().cpu().numpy()
            for i in sorted({(len(x) - 1) * k // 6 for k in range(7)}):
                ref = np.asarray(fn_h(x[i], *[np.asarray(e)[i] for e in extra]), dtype=np.float64)
                if out[i].shape != ref.shape or not np.allclose(out[i], ref, rtol=1e-12,
                                                                atol=1e-12 * max(1.0, float(np.abs(ref).max()))):
                    return None
            return np.ascontiguousarray(out)
        except Exception:
            return None

    # -------------------------------------------------------------- engines
    def _element_terms(self, u, jac, dtype=torch.float64):
        """Element residuals (ncell*nn*vec f64, a reused buffer) and, with jac, element
        Jacobians block-major (ncell, nn, nn, vec, vec) in dtype: the built cell
        kernels or the pointwise evaluation, plus the mass, universal and face terms."""
        plan = self._plan()
        pts, cells = self._dev()
        s = _lib.stream()
        tag = "64" if dtype == torch.float64 else "32"
        if plan["law"] == "fast":
            nd2 = self.nn * self.nn * self.vec * self.vec
            mat = self._hfmat()
            prm = self._params()
            er = self._buf("elem_res", self.num_cells * self.nn * self.vec, torch.float64)
            geo = self._box_geometry()
            if geo is not None:   # identical box cells: the points' gradients formed once
                _lib.call("hf_cell_residual_box_f64", ctypes.byref(mat), _lib.ptr(geo), _lib.ptr(cells),
                          self.num_cells, _lib.ptr(u), _lib.ptr(prm), _lib.ptr(er), s)
            else:
                _lib.call("hf_cell_residual_f64", self.nn, ctypes.byref(mat), _lib.ptr(pts), _lib.ptr(cells),
                          self.num_cells, _lib.ptr(u), _lib.ptr(prm), _lib.ptr(er), s)
            ej = None
            if jac:
                ej = self._buf("elem_jac" + tag, self.num_cells * nd2, dtype)
                _lib.call("hf_cell_jacobian_f" + tag, self.nn, ctypes.byref(mat), _lib.ptr(pts), _lib.ptr(cells),
                          self.num_cells, _lib.ptr(u), _lib.ptr(prm), _lib.ptr(ej), s)
            if plan["mass"] == "cell":
                rm, jm = self._pointwise().mass_terms(u, jac, dtype)
                er += rm
                if jac:
                    ej += jm
        elif self._multi:
            return self._pointwise().multi_terms(u, jac, dtype)
        else:
            er, ej = self._pointwise().cell_terms(u, jac, dtype)
            self._bufs["elem_res"] = er
        extra = []
        if plan["universal"]:
            r, K = self._pointwise().universal_terms(u, jac)
            extra.append((None, r, K))
        faces = [k for k, how in enumerate(plan["surfaces"]) if how == "face"]
        if faces:
            extra += self._pointwise().face_terms(u, jac, faces)
        if extra:
            from .generic import add_element_terms
            add_element_terms(er, ej, extra, self.nn, self.vec, self.num_cells)
        return er, ej

    def residual_vec(self, u):
        """Unconstrained residual R(u) (N,) f64 on the device: cells + loads (problem.py:439-445)."""
        u = u.to(device=self.device, dtype=torch.float64).contiguous().reshape(-1)
        er, _ = self._element_terms(u, False)
        s = _lib.stream()
        R = torch.empty(self.N, dtype=torch.float64, device=self.device)
        fext = self._external()
        box = self._gather_box()
        if box is not None:   # a box_mesh: the cells of each node from its coordinates (same sums, same order)
            (nx, ny, nz), corner = box
            _lib.call("hf_gather_residual_box_f64", self.pattern.handle, nx, ny, nz, corner, _lib.ptr(er),
                      _lib.ptr(fext), _lib.ptr(R), s)
        else:
            _lib.call("hf_gather_residual_f64", self.pattern.handle, _lib.ptr(er), _lib.ptr(fext), _lib.ptr(R), s)
        return R

    def _gather_box(self):
        """((nx, ny, nz), corner table) of a HEX8 vec-3 box_mesh, whose residual gathers by
        lattice coordinates; None otherwise (checked once)."""
        if not hasattr(self, "_gbox"):
            self._gbox = None
            g = getattr(self.fe.mesh, "node_grid", None)
            if self.fe.ele_type == "HEX8" and self.vec == 3 and g is not None and \
                    int(np.prod(g)) == self.fe.num_total_nodes:
                grid = tuple(int(v) for v in g)
                corner = self._box_corners(grid)
                if corner is not None:
                    self._gbox = (grid, corner)
        return self._gbox

    def jacobian_bsr(self, u, dtype=torch.float64, bc=True):
        """Tangent values (the library's sliced-ELL layout of node-pair blocks); Dirichlet rows -> identity rows."""
        u = u.to(device=self.device, dtype=torch.float64).contiguous().reshape(-1)
        p = self.pattern
        tag = "64" if dtype == torch.float64 else "32"
        s = _lib.stream()
        vals = torch.empty(p.nvals, dtype=dtype, device=self.device)
        _, ej = self._element_terms(u, True, dtype)
        if self._fused_ok(dtype):
            # HEX8: rows summed per node over its cells (same order and bits as hf_assemble)
            _lib.call("hf_assemble_hex8_f" + tag, p.handle, _lib.ptr(ej), _lib.ptr(vals), s)
        else:
            _lib.call("hf_assemble_f" + tag, p.handle, _lib.ptr(ej), _lib.ptr(vals), s)
        if bc:
            dofs, _ = self.bc_arrays()
            _lib.call("hf_bc_rows_f" + tag, p.handle, _lib.ptr(dofs), dofs.numel(), _lib.ptr(vals), s)
        return vals

    def cell_operator(self):
        """The tangent as sum_c scale[c] K when that form holds exactly: a built
        LinearElastic law alone (no other term adds to the tangent) on a box_mesh
        HEX8 lattice of identical cells, its parameter constant over each cell's
        points.  Returns (K, scale): K (576 float32, host) the 3x3 blocks (q, q') of
        the cell matrix at unit modulus, corners q = 4 dx + 2 dy + dz; scale (ncell
        float32, device) each cell's modulus (Material._E).  None otherwise, or when
        K lacks the mid-plane reflection symmetry the multigrid's matrix-free kernels
        use.  K is computed once per law and geometry by the FP64 cell kernel on one
        cell (hf_cell_jacobian_f64)."""
        from .materials import LinearElastic
        plan = self._plan()
        mat = self.material
        if not (isinstance(mat, LinearElastic) and plan["law"] == "fast" and plan["pure_fast"]) or \
                self.fe.ele_type != "HEX8" or self.vec != 3 or self.dim != 3:
            return None
        grid = getattr(self.fe.mesh, "node_grid", None)
        if grid is None or int(np.prod(grid)) != self.fe.num_total_nodes:
            return None
        grid = tuple(int(v) for v in grid)
        corner = self._box_corners(grid)
        if corner is None:
            return None
        key = ("cell_op", grid, float(mat.nu), int(self.fe.quadrature_order))
        hit = getattr(self, "_cell_op_K", None)
        if hit is None or hit[0] != key:
            self._cell_op_K = (key, self._cell_matrix(grid, corner, mat.nu))
        K = self._cell_op_K[1]
        if K is None:
            return None
        prm = self._params()
        # the per-cell moduli live in one buffer (a stable address: the multigrid's
        # captured iteration keeps it as a kernel argument), refreshed when the
        # parameter tensor or its contents change; the tensor is referenced by the
        # key, so its address cannot be reused while the key holds it
        law = (mat.param, float(mat.E), float(mat.E0), float(mat.Emin), float(mat.penal))
        key = (None, law) if prm is None else (prm.data_ptr(), prm._version, tuple(prm.shape), law)
        hit = getattr(self, "_cell_scale_key", None)
        scale = self._buf("cell_scale", self.num_cells, torch.float32)
        if hit is None or hit[0] != key or hit[3] is not scale:
            if prm is None:
                scale.fill_(float(mat._E(None)))
                ok = True
            else:
                ok = prm.dim() == 2 and bool((prm == prm[:, :1]).all())
                if ok:
                    scale.copy_(mat._E(prm[:, 0]))
            self._cell_scale_key = (key, ok, prm, scale)   # prm kept: its address stays its own
        if not self._cell_scale_key[1]:
            return None
        return K, scale

    def cell_stencil(self, mg):
        """The multigrid's level-0 stencil (as jacobian_stencil) from the cell operator
        set on `mg` (hf_mg_assemble_cellop_f32): no element Jacobians.  Returns
        (stencil, version)."""
        S = self._buf("mg_stencil", mg.stencil_size(), torch.float32)
        _lib.call("hf_mg_assemble_cellop_f32", mg.handle, _lib.ptr(S), _lib.stream())
        self._stencil_version = getattr(self, "_stencil_version", 0) + 1
        return S, self._stencil_version

    def _cell_matrix(self, grid, corner, nu):
        """The unit-modulus linear-elastic cell matrix in corner order (see
        cell_operator), or None if the lattice's cells differ, its block rows do not
        sum to zero or it is not reflection-symmetric (relative 1e-10)."""
        from .materials import LinearElastic
        nx, ny, nz = grid
        pts = np.asarray(self.fe.points, dtype=np.float64).reshape(nx, ny, nz, 3)
        axes = [pts[:, 0, 0, 0], pts[0, :, 0, 1], pts[0, 0, :, 2]]
        for d, a in enumerate(axes):
            h = np.diff(a)
            if not np.allclose(h, h[0], rtol=1e-12, atol=0.0):
                return None
        lat = np.stack(np.meshgrid(*axes, indexing="ij"), axis=3)
        if np.abs(lat - pts).max() > 1e-12 * max(1.0, np.abs(pts).max()):
            return None
        m1 = LinearElastic(E=1.0, nu=nu).hf()
        m1.quad1d = (int(self.fe.quadrature_order) + 2) // 2
        pts_d, cells_d = self._dev()
        u0 = torch.zeros(self.N, dtype=torch.float64, device=self.device)
        ej = torch.empty(64 * 9, dtype=torch.float64, device=self.device)
        _lib.call("hf_cell_jacobian_f64", 8, ctypes.byref(m1), _lib.ptr(pts_d), _lib.ptr(cells_d), 1,
                  _lib.ptr(u0), None, _lib.ptr(ej), _lib.stream())
        Kb = ej.cpu().numpy().reshape(8, 8, 3, 3)          # block-major (a, b, i, k), local node order
        t = [int(corner[q]) for q in range(8)]             # corner q -> local node
        Kq = Kb[np.ix_(t, t)]                               # (q, q', i, k)
        refl = lambda q: np.array([-1.0 if (q >> (2 - ax)) & 1 else 1.0 for ax in range(3)])
        scale = np.abs(Kq).max()
        if np.abs(Kq.sum(axis=1)).max() > 1e-10 * scale:      # translations carry no force
            return None
        for a in range(8):
            for b in range(8):
                Ra = refl(a)
                want = Ra[:, None] * Kq[0, a ^ b] * Ra[None, :]
                if np.abs(Kq[a, b] - want).max() > 1e-10 * scale:
                    return None
        return np.ascontiguousarray(Kq.reshape(-1).astype(np.float32))

    def jacobian_stencil(self, u, mg):
        """FP32 tangent of a HEX8 box lattice in the multigrid's stencil form: the
        element Jacobians' upper blocks (hf_hex8_jacobian_upper_f32) summed into each
        node's diagonal and 13 upper node blocks by hf_mg_assemble_hex8_f32, without
        the SELL matrix.
        Dirichlet rows stay in place (the multigrid applies the row elimination of
        apply_bc, solver.py:290-363).  The stencil lives in a buffer of this problem
        that the next call overwrites; returns (stencil, version)."""
        pts, cells = self._dev()
        u = u.to(device=self.device, dtype=torch.float64).contiguous().reshape(-1)
        s = _lib.stream()
        mat = self._hfmat()
        prm = self._params()
        ej = self._buf("elem_jac_upper32", self.num_cells * 64 * 9, torch.float32)
        _lib.call("hf_hex8_jacobian_upper_f32", ctypes.byref(mat), _lib.ptr(pts), _lib.ptr(cells),
                  self.num_cells, _lib.ptr(u), _lib.ptr(prm), _lib.ptr(ej), s)
        S = self._buf("mg_stencil", mg.stencil_size(), torch.float32)
        corner = self._box_corners(mg.grid)
        _lib.call("hf_mg_assemble_hex8_f32", mg.handle, self.pattern.handle, _lib.ptr(cells), self.num_cells,
                  corner, _lib.ptr(ej), _lib.ptr(S), s)
        self._stencil_version = getattr(self, "_stencil_version", 0) + 1
        return S, self._stencil_version

    box_residual = True   # HEX8 box lattices of identical cells: hf_cell_residual_box_f64

    def _box_geometry(self):
        """Device buffer of the per-point gradients and weights shared by every cell
        (hf_hex8_box_geometry) when the mesh is a HEX8 box lattice of identical cells:
        a lattice cell order with one corner order (_box_corners) and uniform spacing
        along each axis; None otherwise (checked once per mesh)."""
        if not (self.box_residual and self.nn == 8 and self.dim == 3):
            return None
        if not hasattr(self, "_box_geo"):
            self._box_geo = None
            g = getattr(self.fe.mesh, "node_grid", None)
            if g is not None and int(np.prod(g)) == self.fe.num_total_nodes and \
                    self._box_corners(tuple(int(v) for v in g)) is not None:
                nx, ny, nz = (int(v) for v in g)
                pts = np.asarray(self.fe.points, dtype=np.float64).reshape(nx, ny, nz, 3)
                axes = [pts[:, 0, 0, 0], pts[0, :, 0, 1], pts[0, 0, :, 2]]
                uniform = all(np.allclose(np.diff(a), np.diff(a)[0], rtol=1e-12, atol=0.0) for a in axes)
                lat = np.stack(np.meshgrid(*axes, indexing="ij"), axis=3)
                if uniform and np.abs(lat - pts).max() <= 1e-12 * max(1.0, np.abs(pts).max()):
                    pts_d, cells_d = self._dev()
                    nb = int(_lib.load().hf_hex8_box_geometry_bytes())
                    geo = torch.empty(nb // 8, dtype=torch.float64, device=self.device)
                    _lib.call("hf_hex8_box_geometry", _lib.ptr(pts_d), _lib.ptr(cells_d), self.num_cells,
                              _lib.ptr(geo), _lib.stream())
                    self._box_geo = geo
        return self._box_geo

    def _box_corners(self, grid):
        """(ctypes int32[8]) local index of corner (dx, dy, dz) at 4dx + 2dy + dz when every
        cell is the lattice cell (i, j, k) = (i*(ny-1) + j)*(nz-1) + k with one corner
        order (a box_mesh), checked here once; None otherwise."""
        key = ("box_corners", tuple(grid))
        if getattr(self, "_corner_key", None) != key:
            self._corner_key, self._corner = key, None
            nx, ny, nz = grid
            cells = np.asarray(self.fe.cells, dtype=np.int64)
            nc = (nx - 1) * (ny - 1) * (nz - 1)
            if cells.shape == (nc, 8) and nc > 0:
                cid = np.arange(nc)
                cx, cy, cz = cid // ((ny - 1) * (nz - 1)), (cid // (nz - 1)) % (ny - 1), cid % (nz - 1)
                d = cells - ((cx * ny + cy) * nz + cz)[:, None]
                step = {bx * ny * nz + by * nz + bz: 4 * bx + 2 * by + bz
                        for bx in (0, 1) for by in (0, 1) for bz in (0, 1)}
                if np.all(d == d[0][None, :]) and all(int(v) in step for v in d[0]):
                    bits = [step[int(v)] for v in d[0]]
                    if sorted(bits) == list(range(8)):
                        table = [0] * 8
                        for b, q in enumerate(bits):
                            table[q] = b
                        self._corner = (ctypes.c_int32 * 8)(*table)
        return self._corner

    fused_assembly = True   # HEX8 vec-3 patterns: hf_assemble_hex8 (per-node rows) instead of hf_assemble

    def _fused_ok(self, dtype):
        if not (self.fused_assembly and self.nn == 8 and self.vec == 3 and self.dim == 3):
            return False
        limit = 42 if dtype == torch.float64 else 85   # LDS: 32 rows x max_row x 9 values
        return self.pattern.max_row_blocks <= limit

    def element_jacobians(self, u, dtype=torch.float64):
        """Element matrices (ncell, nd, nd), jac[r,c] = dR_r/du_c (problem.py:262-266)."""
        u = u.to(device=self.device, dtype=torch.float64).contiguous().reshape(-1)
        nn, v = self._enn, self._evec
        _, ej = self._element_terms(u, True, dtype)
        # block-major (c, a, b, i, k) -> (c, a*v+i, b*v+k)
        return ej.reshape(self.num_cells, nn, nn, v, v).permute(0, 1, 3, 2, 4).reshape(
            self.num_cells, nn * v, nn * v).clone()

    # ------------------------------------------- reference-compatible methods
    def compute_residual(self, sol_list):
        return self.unflatten_fn_sol_list(self.residual_vec(self.flatten_sol_list(sol_list)))

    def newton_update(self, sol_list):
        u = self.flatten_sol_list(sol_list)
        self._newton_u = u.clone()
        self._A_cache = None
        return self.compute_residual(u)

    @property
    def V(self):
        """Host COO values of the last newton_update: cells, then zero face blocks (problem.py:453-458)."""
        K = self.element_jacobians(self._newton_u).reshape(-1).cpu().numpy()
        nd = self._enn * self._evec
        nface = sum(len(b) for b in self.boundary_inds_list) * nd * nd
        return np.concatenate([K, np.zeros(nface)])

    def _coo_inds(self):
        d = self._dof_cells
        return [d] + [d[b[:, 0]] for b in self.boundary_inds_list]

    @property
    def I(self):
        return np.concatenate([np.repeat(x[:, :, None], x.shape[1], axis=2).reshape(-1) for x in self._coo_inds()])

    @property
    def J(self):
        return np.concatenate([np.repeat(x[:, None, :], x.shape[1], axis=1).reshape(-1) for x in self._coo_inds()])


class Pattern:
    """Owner of an hf_pattern handle (destroyed with the Problem)."""

    def __init__(self, handle, problem):
        self.handle = handle
        nb, nnz, nc = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.call("hf_pattern_sizes", handle, ctypes.byref(nb), ctypes.byref(nnz), ctypes.byref(nc))
        self.nblocks, self.nnz, self.ncontrib = nb.value, nnz.value, nc.value
        self.nvals = int(_lib.load().hf_matrix_values(handle))
        self.max_row_blocks = int(_lib.load().hf_pattern_max_row_blocks(handle))
        self.N = problem.N
        self.vec = problem._evec
        self.device = problem.device

    def csr(self):
        """Scalar CSR pattern (rowptr, colind) as device int32 tensors."""
        rp = torch.empty(self.N + 1, dtype=torch.int32, device=self.device)
        ci = torch.empty(self.nnz, dtype=torch.int32, device=self.device)
        _lib.call("hf_pattern_csr", self.handle, _lib.ptr(rp), _lib.ptr(ci), _lib.stream())
        return rp, ci

    def __del__(self):
        try:
            if self.handle:
                torch.cuda.synchronize()
                _lib.load().hf_pattern_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
