"""Cell and face terms of maps that are only known pointwise (problem.py:189-325).

The reference differentiates every user map with jacfwd inside its cell loop.
Here the division of work is:

* HIP (libhipfem, generic.hip): the element contractions -- displacement
  gradients at the quadrature points, the residual sum_q JxW (P : grad N + m N)
  and the element Jacobian sum_q JxW (grad N . C . grad N + N Mt N) -- over all
  cells in one launch each, for any element the set-up tabulates;
* torch.func on the device: the pointwise law itself and its exact derivative
  (vmap over the quadrature points, jacfwd for C = dP/du_grad and Mt = dm/du),
  i.e. the user's own code, written with ``jax_fem_amd.numpy``.

Face terms (surface maps that depend on u, universal surface kernels) and
universal cell kernels act on the element buffers before the deterministic
gather / assembly: their element vectors and matrices are added to the cell's
(index_add over the faces' cells), which is what the reference's COO layout
(cells, then faces; problem.py:86-107, 426-460) sums to.
"""
import numpy as np
import torch

from . import _lib
from . import numpy as jnp

_CHUNK = 1 << 17     # quadrature points per torch.func evaluation (bounds the tangent's memory)


def _vmap_chunks(fn, args, batch_dims=None):
    """fn vmapped over the leading axis of every tensor in args, in chunks."""
    n = args[0].shape[0]
    outs = []
    for s in range(0, n, _CHUNK):
        e = min(n, s + _CHUNK)
        outs.append(torch.func.vmap(fn)(*[a[s:e] for a in args]))
    if isinstance(outs[0], tuple):
        return tuple(torch.cat([o[k] for o in outs]) for k in range(len(outs[0])))
    return torch.cat(outs)


class PointwiseTerms:
    """Device evaluation of a problem's pointwise maps and their element terms."""

    def __init__(self, problem):
        self.p = problem
        fe = problem.fe          # tensor, mass and surface maps act on the first variable
        dev = problem.device
        self.nn, self.nq, self.vec, self.dim = fe.num_nodes, fe.num_quads, fe.vec, problem.dim
        self.ncell = problem.num_cells
        self.ndof0 = fe.num_total_dofs
        grads, jxw = fe.get_shape_grads()
        self.sg = torch.as_tensor(np.ascontiguousarray(grads), dtype=torch.float64, device=dev)
        self.jxw = torch.as_tensor(np.ascontiguousarray(jxw), dtype=torch.float64, device=dev)
        self.sv = torch.as_tensor(np.ascontiguousarray(fe.shape_vals), dtype=torch.float64, device=dev)
        self.xq = torch.as_tensor(np.ascontiguousarray(fe.get_physical_quad_points()), dtype=torch.float64,
                                  device=dev)
        self.tensor_fn = None
        self.mass_fn = None
        self.universal = None
        self.faces = []      # per surface: dict of face data, or None

    # -- quadrature-point data ----------------------------------------------------
    def point_data(self, u, need_uq):
        """u_grad (ncell*nq, vec, dim) and u at the points (or None) of the first variable."""
        cells = self.p._var0_cells()
        u = u[:self.ndof0].contiguous()
        Q = self.ncell * self.nq
        ug = torch.empty(Q * self.vec * self.dim, dtype=torch.float64, device=u.device)
        uq = torch.empty(Q * self.vec, dtype=torch.float64, device=u.device) if need_uq else None
        _lib.call("hf_generic_ugrad_f64", _lib.ptr(cells), self.ncell, self.nn, self.nq, self.vec, self.dim,
                  _lib.ptr(u), _lib.ptr(self.sg), _lib.ptr(self.sv), _lib.ptr(ug), _lib.ptr(uq), _lib.stream())
        return ug.reshape(Q, self.vec, self.dim), (uq.reshape(Q, self.vec) if need_uq else None)

    def _ivs(self):
        """Volume internal variables, each (ncell*nq, ...) on the device."""
        out = []
        for a in self.p.volume_internal_vars():
            out.append(a.reshape((self.ncell * self.nq,) + tuple(a.shape[2:])))
        return out

    # -- cell terms -------------------------------------------------------------------
    def cell_terms(self, u, jac, dtype=torch.float64):
        """Element residuals (ncell, nn*vec) f64 and, with jac, element Jacobians
        block-major (ncell, nn, nn, vec, vec) in dtype from the tensor and mass maps."""
        dev = u.device
        ivs = self._ivs() if (self.tensor_fn is not None or self.mass_fn is not None) else []
        need_uq = self.mass_fn is not None
        ug, uq = self.point_data(u, need_uq)
        P = C = M = Mt = None
        with jnp.on_device(dev):
            if self.tensor_fn is not None:
                f = self.tensor_fn
                P = _vmap_chunks(f, [ug] + ivs)
                if jac:
                    C = _vmap_chunks(torch.func.jacfwd(f, argnums=0), [ug] + ivs)
            if self.mass_fn is not None:
                m = self.mass_fn
                xq = self.xq.reshape(-1, self.dim)
                M = _vmap_chunks(m, [uq, xq] + ivs)
                if jac:
                    Mt = _vmap_chunks(torch.func.jacfwd(m, argnums=0), [uq, xq] + ivs)
        er = torch.empty(self.ncell * self.nn * self.vec, dtype=torch.float64, device=dev)
        Pc = P.contiguous() if P is not None else None
        Mc = M.contiguous() if M is not None else None
        _lib.call("hf_generic_residual_f64", self.ncell, self.nn, self.nq, self.vec, self.dim, _lib.ptr(self.sg),
                  _lib.ptr(self.sv), _lib.ptr(self.jxw), _lib.ptr(Pc), _lib.ptr(Mc), _lib.ptr(er), _lib.stream())
        ej = None
        if jac and C is None and Mt is None:       # no tensor or mass map: the other terms fill it
            ej = torch.zeros(self.ncell * self.nn * self.nn * self.vec * self.vec, dtype=dtype, device=dev)
        elif jac:
            ej = torch.empty(self.ncell * self.nn * self.nn * self.vec * self.vec, dtype=dtype, device=dev)
            Cc = C.contiguous() if C is not None else None
            Mtc = Mt.contiguous() if Mt is not None else None
            tag = "64" if dtype == torch.float64 else "32"
            _lib.call("hf_generic_jacobian_f" + tag, self.ncell, self.nn, self.nq, self.vec, self.dim,
                      _lib.ptr(self.sg), _lib.ptr(self.sv), _lib.ptr(self.jxw), _lib.ptr(Cc), _lib.ptr(Mtc),
                      _lib.ptr(ej), _lib.stream())
        return er, ej

    def mass_terms(self, u, jac, dtype):
        """The mass map's element residuals / Jacobians alone (added to a built law's)."""
        tf = self.tensor_fn
        self.tensor_fn = None
        try:
            return self.cell_terms(u, jac, dtype)
        finally:
            self.tensor_fn = tf

    def _all_vars_geometry(self):
        """Shape gradients (ncell, nq, sum nn, dim), JxW (ncell, nvars, nq) and v_grads_JxW
        (ncell, nq, sum nn, 1, dim) of every variable, concatenated (problem.py:130-147)."""
        if getattr(self, "_geo_all", None) is None:
            dev = self.sg.device
            if self.p.num_vars == 1:
                sg, jxw = self.sg, self.jxw[:, None, :]
            else:
                gs, js = zip(*[fe.get_shape_grads() for fe in self.p.fes])
                sg = torch.as_tensor(np.ascontiguousarray(np.concatenate(gs, axis=2)), dtype=torch.float64,
                                     device=dev)
                jxw = torch.as_tensor(np.ascontiguousarray(np.stack(js, axis=1)), dtype=torch.float64, device=dev)
            nn_of = np.diff(self.p.num_nodes_cumsum)
            w = torch.cat([jxw[:, k, :, None, None].expand(-1, -1, int(n), 1) for k, n in enumerate(nn_of)], dim=2)
            vgj = (sg * w)[:, :, :, None, :]
            self._geo_all = (sg, jxw, vgj)
        return self._geo_all

    def _cell_sol_flat(self, u):
        """(ncell, nd): each cell's DOFs of every variable, in the reference's flat order."""
        _, dcells = self.p._dev()
        if self.p.num_vars == 1:
            return u.reshape(-1, self.vec)[dcells.long()].reshape(self.ncell, -1)
        return u[dcells.long()]

    def universal_terms(self, u, jac):
        """get_universal_kernel's element residuals (ncell, nd) and Jacobians (ncell, nd, nd)."""
        p = self.p
        cu = self._cell_sol_flat(u)
        sgc, jxw, vgj = self._all_vars_geometry()
        args = [cu, self.xq, sgc, jxw, vgj] + p.cell_internal_vars()
        k = self.universal
        with jnp.on_device(u.device):
            r = torch.func.vmap(k)(*args)
            K = torch.func.vmap(torch.func.jacfwd(k, argnums=0))(*args) if jac else None
        return r, K

    # -- faces -------------------------------------------------------------------------
    def face_terms(self, u, jac, surfaces):
        """Element vectors / matrices of the faces' u-dependent terms, per surface:
        list of (cells (nsel,), r (nsel, nd), K (nsel, nd, nd) or None)."""
        p = self.p
        _, cells = p._dev()
        dev = u.device
        out = []
        for k in surfaces:
            fd = self.faces[k]
            if fd is None:
                continue
            cu = u[:self.ndof0].reshape(-1, self.vec)[fd["cell_nodes"]]          # (nsel, nn, vec)
            isv = fd["isv"]
            with jnp.on_device(dev):
                if fd["universal"] is not None:
                    kern = fd["universal"]
                    if "sol_idx" in fd:       # several variables: the cells' flat DOFs, geometry of all
                        args = [u[fd["sol_idx"]], fd["x"], fd["vals_all"], fd["grads_all"], fd["nanson_all"]] + isv
                    else:
                        args = [cu.reshape(cu.shape[0], -1), fd["x"], fd["vals"], fd["grads"],
                                fd["nanson"][:, None, :]] + isv
                    r = torch.func.vmap(kern)(*args)
                    K = torch.func.vmap(torch.func.jacfwd(kern, argnums=0))(*args) if jac else None
                else:
                    smap = fd["map"]
                    ufq = torch.einsum("sqn,sni->sqi", fd["vals"], cu)        # (nsel, nfq, vec)
                    S = fd["nsel"] * fd["nfq"]
                    flat = [a.reshape((S,) + tuple(a.shape[2:])) for a in isv]
                    t = torch.func.vmap(smap)(ufq.reshape(S, self.vec), fd["x"].reshape(S, self.dim), *flat)
                    t = t.reshape(fd["nsel"], fd["nfq"], self.vec)
                    w = fd["vals"] * fd["nanson"][:, :, None]                # (nsel, nfq, nn)
                    r = torch.einsum("sqi,sqn->sni", t, w).reshape(fd["nsel"], -1)
                    K = None
                    if jac:
                        dt = torch.func.vmap(torch.func.jacfwd(smap, argnums=0))(
                            ufq.reshape(S, self.vec), fd["x"].reshape(S, self.dim), *flat)
                        dt = dt.reshape(fd["nsel"], fd["nfq"], self.vec, self.vec)
                        K = torch.einsum("sqik,sqa,sqb->saibk", dt, w, fd["vals"]).reshape(
                            fd["nsel"], self.nn * self.vec, self.nn * self.vec)
            out.append((fd["cells"], r, K))
        return out


    def multi_terms(self, u, jac, dtype):
        """Element residuals (ncell*nd f64) and Jacobians (ncell*nd*nd, dense per cell:
        the layout of a DOF-level pattern) of a problem with several variables: the
        first variable's tensor / mass / surface maps placed in its DOFs of each cell
        (problem.py:189-259 use cell_sol_list[0]), plus the universal kernels."""
        p = self.p
        nd = p._enn
        nd0 = self.nn * self.vec
        dev = u.device
        er = torch.zeros(self.ncell, nd, dtype=torch.float64, device=dev)
        ej = torch.zeros(self.ncell, nd, nd, dtype=dtype, device=dev) if jac else None
        if self.tensor_fn is not None or self.mass_fn is not None:
            r0, k0 = self.cell_terms(u, jac, dtype)
            er[:, :nd0] += r0.view(self.ncell, nd0)
            if jac:
                k0 = k0.view(self.ncell, self.nn, self.nn, self.vec, self.vec).permute(0, 1, 3, 2, 4)
                ej[:, :nd0, :nd0] += k0.reshape(self.ncell, nd0, nd0)
        terms = []
        if self.universal is not None:
            r, K = self.universal_terms(u, jac)
            terms.append((None, r, K))
        faces = [k for k, how in enumerate(p._plan()["surfaces"]) if how == "face"]
        if faces:
            for cells, r, K in self.face_terms(u, jac, faces):
                if r.shape[1] != nd:          # a first-variable surface map: its DOFs
                    r = torch.nn.functional.pad(r, (0, nd - r.shape[1]))
                    if K is not None:
                        K = torch.nn.functional.pad(K, (0, nd - K.shape[2], 0, nd - K.shape[1]))
                terms.append((cells, r, K))
        for cells, r, K in terms:
            if cells is None:
                er += r.to(er.dtype)
                if jac and K is not None:
                    ej += K.to(ej.dtype)
            else:
                er.index_add_(0, cells, r.to(er.dtype))
                if jac and K is not None:
                    ej.index_add_(0, cells, K.to(ej.dtype))
        return er.reshape(-1), (ej.reshape(-1) if jac else None)


def add_element_terms(er, ej, terms, nn, vec, ncell):
    """Add (cells, r (n, nd), K (n, nd, nd)) element terms into the residual buffer
    (ncell*nd) and the block-major Jacobian buffer (ncell, nn, nn, vec, vec)."""
    nd = nn * vec
    erv = er.view(ncell, nd)
    for cells, r, K in terms:
        if cells is None:
            erv += r.to(er.dtype)
        else:
            erv.index_add_(0, cells, r.to(er.dtype))
        if ej is not None and K is not None:
            Kb = K.reshape(-1, nn, vec, nn, vec).permute(0, 1, 3, 2, 4).reshape(-1, nn * nn * vec * vec)
            ejv = ej.view(ncell, nn * nn * vec * vec)
            if cells is None:
                ejv += Kb.to(ej.dtype)
            else:
                ejv.index_add_(0, cells, Kb.to(ej.dtype))
