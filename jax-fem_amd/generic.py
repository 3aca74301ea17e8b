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
        fe = problem.fe
        dev = problem.device
        self.nn, self.nq, self.vec, self.dim = fe.num_nodes, fe.num_quads, problem.vec, problem.dim
        self.ncell = problem.num_cells
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
        """u_grad (ncell*nq, vec, dim) and u at the points (or None)."""
        _, cells = self.p._dev()
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
        ivs = self._ivs()
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

    def universal_terms(self, u, jac):
        """get_universal_kernel's element residuals (ncell, nd) and Jacobians (ncell, nd, nd)."""
        p = self.p
        _, cells = p._dev()
        dev = u.device
        cu = u.reshape(-1, self.vec)[cells.long()].reshape(self.ncell, -1)
        sgc = self.sg
        jxw = self.jxw[:, None, :]                          # (ncell, 1, nq): one variable
        vgj = (sgc * self.jxw[:, :, None, None])[:, :, :, None, :]     # (ncell, nq, nn, 1, dim)
        ivs = [a for a in p.volume_internal_vars()]
        args = [cu, self.xq, sgc, jxw, vgj] + ivs
        k = self.universal
        with jnp.on_device(dev):
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
            cu = u.reshape(-1, self.vec)[fd["cell_nodes"]]          # (nsel, nn, vec)
            isv = fd["isv"]
            with jnp.on_device(dev):
                if fd["universal"] is not None:
                    kern = fd["universal"]
                    args = [cu.reshape(cu.shape[0], -1), fd["x"], fd["vals"], fd["grads"], fd["nanson"][:, None, :]] + isv
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
