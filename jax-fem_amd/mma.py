"""Topology optimisation with the method of moving asymptotes (drop-in for jax_fem/mma.py).

Same interface as the reference module:
  * ``optimize(fe, rho_ini, optimizationParams, objectiveHandle, consHandle,
    numConstraints)`` (mma.py:415-528): design variables in [0, 1], MMA with
    a0 = 1, a = 0, c = 1e4, d = 0 and the move limit
    ``optimizationParams['movelimit']``, ``optimizationParams['maxIters']``
    iterations; ``objectiveHandle(rho) -> (J, dJ)`` and
    ``consHandle(rho, iter) -> (vc (m,), dvc (m, ...))`` as in the reference;
  * the module switches ``density_filtering`` (off) and ``sensitivity_filtering``
    (on), as mma.py:23-24;
  * ``compute_filter_kd_tree(fe)`` (mma.py:27-56), ``applySensitivityFilter``
    and ``applyDensityFilter`` (mma.py:58-66); ``MMA`` with the reference's
    setters and ``mmasub`` (mma.py:68-205).

Where the work runs: the filter matrix is built once on the host (a k-d tree
over the cell centroids, as the reference) and held on the device as a
20-wide neighbour table; applying it, the MMA subproblem set-up and its
primal-dual Newton iterations run in libhipfem (mma.hip) in FP64 on the
device.  The design vector and the asymptotes stay on the device between
iterations; the handles receive and may return torch tensors (or NumPy
arrays), and ``optimize`` returns a device tensor of rho_ini's shape.
"""
import ctypes
import logging
import time

import numpy as np
import torch

from . import _lib

logger = logging.getLogger("jax_fem_amd")

density_filtering = False
sensitivity_filtering = True

_NEIGHBOURS = 20


def _device():
    return torch.device("cuda", torch.cuda.current_device())


class Filter(dict):
    """The filter matrix H (mma.py:27-56) as a k-wide neighbour table on the
    device: keys 'nbr' (n, k) int32, 'w' (n, k) f64, 'Hs' (n,) row sums, plus
    'rmin' and the host CSR matrix 'H_host' for inspection."""


def compute_filter_kd_tree(fe, num_nbs=_NEIGHBOURS):
    """Weights max(rmin - d, 0) from each flexible cell's centroid to its num_nbs
    nearest flexible centroids, rmin = 1.5 (sum JxW / num_cells)^(1/dim)
    (mma.py:27-56).  The k-d tree query runs for all cells at once."""
    from scipy.spatial import KDTree
    import scipy.sparse as sp
    flex = np.asarray(getattr(fe, "flex_inds", np.arange(fe.num_cells)))
    cent = np.asarray(fe.points)[np.asarray(fe.cells)].mean(axis=1)[flex]
    nflex = len(cent)
    avg = float(np.sum(np.asarray(fe.JxW))) / fe.num_cells
    rmin = 1.5 * avg ** (1.0 / fe.dim)
    k = min(num_nbs, nflex)
    dd, ii = KDTree(cent).query(cent, k)
    dd, ii = dd.reshape(nflex, k), ii.reshape(nflex, k)
    w = np.maximum(rmin - dd, 0.0)
    dev = _device()
    ft = Filter()
    ft["nbr"] = torch.as_tensor(ii.astype(np.int32), device=dev).contiguous()
    ft["w"] = torch.as_tensor(w, dtype=torch.float64, device=dev).contiguous()
    ft["Hs"] = torch.as_tensor(w.sum(axis=1), dtype=torch.float64, device=dev).contiguous()
    ft["rmin"] = rmin
    ft["H_host"] = sp.csr_matrix((w.reshape(-1), (np.repeat(np.arange(nflex), k), ii.reshape(-1))),
                                 shape=(nflex, nflex))
    return ft


def _dev64(a, shape=None):
    t = a.detach() if torch.is_tensor(a) else torch.as_tensor(np.asarray(a))
    t = t.to(device=_device(), dtype=torch.float64)
    return t.reshape(shape).contiguous() if shape is not None else t.contiguous()


def applySensitivityFilter(ft, rho, dJ, dvc):
    """dJ <- H (rho dJ / max(1e-3, rho) / Hs), each constraint gradient likewise
    (mma.py:58-61).  Returns device tensors shaped like dJ and dvc."""
    n, k = ft["nbr"].shape
    m = int(np.prod(np.shape(dvc)[:1]))
    g = torch.cat([_dev64(dJ, (1, n)), _dev64(dvc, (m, n))])
    out = torch.empty_like(g)
    _lib.call("hf_mma_sensitivity_filter_f64", _lib.ptr(ft["nbr"]), _lib.ptr(ft["w"]), _lib.ptr(ft["Hs"]), n, k,
              _lib.ptr(_dev64(rho, (n,))), _lib.ptr(g), m + 1, _lib.ptr(out), _lib.stream())
    return out[0].reshape(tuple(np.shape(dJ))), out[1:].reshape(tuple(np.shape(dvc)))


def applyDensityFilter(ft, rho):
    """rho <- H rho / Hs (mma.py:63-66)."""
    n, k = ft["nbr"].shape
    out = torch.empty(n, dtype=torch.float64, device=_device())
    _lib.call("hf_mma_density_filter_f64", _lib.ptr(ft["nbr"]), _lib.ptr(ft["w"]), _lib.ptr(ft["Hs"]), n, k,
              _lib.ptr(_dev64(rho, (n,))), _lib.ptr(out), _lib.stream())
    return out.reshape(tuple(np.shape(rho)))


class MMA:
    """The reference's MMA state holder (mma.py:68-112) over device vectors;
    ``mmasub(xval)`` solves one subproblem in libhipfem."""

    def __init__(self):
        self.epoch = 0
        self._handle = None
        self.info = None

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and _lib._lib is not None:
            _lib.load().hf_mma_destroy(h)

    def resetMMACounter(self):
        self.epoch = 0

    def registerMMAIter(self, xval, xold1, xold2):
        self.epoch += 1
        self.xval, self.xold1, self.xold2 = xval, xold1, xold2

    def setNumConstraints(self, numConstraints):
        self.numConstraints = numConstraints

    def setNumDesignVariables(self, numDesVar):
        self.numDesignVariables = numDesVar

    def setMinandMaxBoundsForDesignVariables(self, xmin, xmax):
        self.xmin, self.xmax = _dev64(xmin, (-1,)), _dev64(xmax, (-1,))

    def setObjectiveWithGradient(self, obj, objGrad):
        self.objective, self.objectiveGradient = obj, objGrad

    def setConstraintWithGradient(self, cons, consGrad):
        self.constraint, self.consGrad = cons, consGrad

    def setScalingParams(self, zconst, zscale, ylinscale, yquadscale):
        self.zconst = float(zconst)
        self.zscale = np.asarray(zscale, dtype=np.float64).reshape(-1)
        self.ylinscale = np.asarray(ylinscale, dtype=np.float64).reshape(-1)
        self.yquadscale = np.asarray(yquadscale, dtype=np.float64).reshape(-1)

    def setMoveLimit(self, movelim):
        self.moveLimit = float(movelim)

    def setLowerAndUpperAsymptotes(self, low, upp):
        self.lowAsymp, self.upAsymp = _dev64(low, (-1,)).clone(), _dev64(upp, (-1,)).clone()

    def getOptimalValues(self):
        return self.xmma, None, None

    def getAsymptoteValues(self):
        return self.lowAsymp, self.upAsymp

    def mmasub(self, xval):
        """One subproblem at xval (mma.py:114-205 + subsolv :207-410, m < n)."""
        m, n = int(self.numConstraints), int(self.numDesignVariables)
        if self._handle is None:
            h = ctypes.c_void_p()
            _lib.call("hf_mma_create", n, m, ctypes.byref(h))
            self._handle = h
        xv = _dev64(xval, (n,))
        xo1 = _dev64(self.xold1, (n,))
        xo2 = _dev64(self.xold2, (n,))
        df0 = _dev64(self.objectiveGradient, (n,))
        fval = _dev64(self.constraint, (m,))
        dfdx = _dev64(self.consGrad, (m, n))
        low = self.lowAsymp.contiguous()
        upp = self.upAsymp.contiguous()
        xmma = torch.empty(n, dtype=torch.float64, device=_device())
        arr = lambda a: (ctypes.c_double * m)(*[float(t) for t in a])
        info = _lib.HFMmaInfo()
        _lib.call("hf_mma_update_f64", self._handle, int(self.epoch), _lib.ptr(xv), _lib.ptr(xo1), _lib.ptr(xo2),
                  _lib.ptr(self.xmin), _lib.ptr(self.xmax), _lib.ptr(low), _lib.ptr(upp), _lib.ptr(df0),
                  _lib.ptr(fval), _lib.ptr(dfdx), self.zconst, arr(self.zscale), arr(self.ylinscale),
                  arr(self.yquadscale), self.moveLimit, _lib.ptr(xmma), ctypes.byref(info), _lib.stream())
        self.lowAsymp, self.upAsymp = low, upp
        self.xmma = xmma
        self.info = {"newton_steps": info.newton_steps, "trials": info.trials, "res_max": info.res_max}


def optimize(fe, rho_ini, optimizationParams, objectiveHandle, consHandle, numConstraints):
    """Topology optimisation loop (mma.py:415-528); returns rho (device tensor)."""
    ft = compute_filter_kd_tree(fe)
    dev = _device()
    rho = _dev64(rho_ini)
    shape = tuple(rho.shape)
    m = int(numConstraints)
    n = rho.numel()
    mma = MMA()
    mma.setNumConstraints(m)
    mma.setNumDesignVariables(n)
    mma.setMinandMaxBoundsForDesignVariables(torch.zeros(n, dtype=torch.float64, device=dev),
                                             torch.ones(n, dtype=torch.float64, device=dev))
    xval = rho.reshape(-1).clone()
    xold1, xold2 = xval.clone(), xval.clone()
    mma.registerMMAIter(xval, xold1, xold2)
    mma.setLowerAndUpperAsymptotes(torch.ones(n, dtype=torch.float64, device=dev),
                                   torch.ones(n, dtype=torch.float64, device=dev))
    mma.setScalingParams(1.0, np.zeros(m), 10000 * np.ones(m), np.zeros(m))
    mma.setMoveLimit(optimizationParams["movelimit"])
    history = []
    loop = 0
    while loop < optimizationParams["maxIters"]:
        loop += 1
        logger.info("MMA solver...")
        rho_physical = applyDensityFilter(ft, rho) if density_filtering else rho
        J, dJ = objectiveHandle(rho_physical)
        vc, dvc = consHandle(rho_physical, loop)
        if sensitivity_filtering:
            dJ, dvc = applySensitivityFilter(ft, rho, dJ, dvc)
        J = float(J.detach()) if torch.is_tensor(J) else float(np.asarray(J))
        start = time.time()
        mma.setObjectiveWithGradient(J, _dev64(dJ, (n,)))
        mma.setConstraintWithGradient(_dev64(vc, (m,)), _dev64(dvc, (m, n)))
        mma.mmasub(xval)
        xmma, _, _ = mma.getOptimalValues()
        xold2 = xold1.clone()
        xold1 = xval.clone()
        xval = xmma.clone()
        mma.registerMMAIter(xval, xold1, xold2)
        rho = xval.reshape(shape)
        logger.info("MMA took %s [s]" % (time.time() - start))
        vc_h = vc.detach().cpu().numpy() if torch.is_tensor(vc) else np.asarray(vc)
        logger.info("Iter %d; J %.5f; constraint %s\n\n\n" % (loop, J, vc_h))
        history.append({"J": J, "constraint": vc_h, "mma": dict(mma.info)})
    optimize.history = history
    return rho
