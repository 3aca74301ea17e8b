"""``jax_fem_amd.numpy``: the array namespace for user-written constitutive maps.

The reference's users write their pointwise maps (``get_tensor_map``,
``get_mass_map``, ``get_surface_maps``, ``get_universal_kernel``) with
``jax.numpy`` so that JAX can trace and differentiate them
(e.g. tests/benchmarks/hyperelasticity/test_hyper_elasticity.py:15-34).  Here
the same code runs with ``import jax_fem_amd.numpy as np``:

* called with NumPy arrays or Python numbers (location and value functions,
  post-processing, the probe that recognises a built law), every function is
  NumPy's own;
* called with torch tensors -- inside the transforms of ``jax_fem_amd``
  (``vmap``, ``grad``, ``jacfwd``: torch.func) when a Problem evaluates a map at
  its quadrature points on the GPU -- the functions are torch's, so the map is
  differentiated exactly (forward mode, as the reference's jacfwd).  Constants
  made inside (``eye``, ``zeros``, ``array`` of numbers) are then created on the
  device of the evaluation in FP64.
"""
import contextlib
import math
import threading

import numpy as _np
import torch

pi = math.pi
e = math.e
inf = math.inf
nan = math.nan
float64 = _np.float64
float32 = _np.float32
int32 = _np.int32
int64 = _np.int64
newaxis = None
ndarray = _np.ndarray

_ctx = threading.local()


def _device():
    return getattr(_ctx, "device", None)


@contextlib.contextmanager
def on_device(device):
    """Constants made by this namespace inside the block are torch tensors on `device`."""
    old = getattr(_ctx, "device", None)
    _ctx.device = device
    try:
        yield
    finally:
        _ctx.device = old


def _has_tensor(args):
    for a in args:
        if torch.is_tensor(a):
            return True
        if isinstance(a, (list, tuple)) and _has_tensor(a):
            return True
    return False


def _t(x):
    if torch.is_tensor(x):
        return x
    dev = _device() or "cpu"
    if isinstance(x, (list, tuple)) and _has_tensor(x):
        return torch.stack([_t(v) for v in x])
    a = _np.asarray(x)
    dt = torch.float64 if a.dtype.kind in "fc" or a.dtype == object else (
        torch.bool if a.dtype == bool else torch.int64)
    return torch.as_tensor(a, dtype=dt, device=dev)


def _torch_mode(*args):
    return _has_tensor(args) or _device() is not None


def _maybe_t(*args):
    return [_t(a) for a in args]


# ---- constructors -------------------------------------------------------------
def array(x, dtype=None):
    if _torch_mode(x):
        t = _t(x)
        return t.to(torch.float64) if dtype in (None, float, float64) and t.dtype.is_floating_point else t
    return _np.array(x, dtype=dtype)


asarray = array


def eye(n, m=None, dtype=None):
    if _device() is not None:
        return torch.eye(n, m if m is not None else n, dtype=torch.float64, device=_device())
    return _np.eye(n, m, dtype=dtype or float)


def zeros(shape, dtype=None):
    if _device() is not None:
        return torch.zeros(shape, dtype=torch.float64, device=_device())
    return _np.zeros(shape, dtype=dtype or float)


def ones(shape, dtype=None):
    if _device() is not None:
        return torch.ones(shape, dtype=torch.float64, device=_device())
    return _np.ones(shape, dtype=dtype or float)


def full(shape, value, dtype=None):
    if _device() is not None or torch.is_tensor(value):
        shape = tuple(shape) if isinstance(shape, (tuple, list)) else (shape,)
        dev = value.device if torch.is_tensor(value) else _device()
        return torch.zeros(shape, dtype=torch.float64, device=dev) + _t(value)
    return _np.full(shape, value, dtype=dtype)


def zeros_like(a):
    return torch.zeros_like(a) if torch.is_tensor(a) else _np.zeros_like(a)


def ones_like(a):
    return torch.ones_like(a) if torch.is_tensor(a) else _np.ones_like(a)


def arange(*args, **kw):
    if _device() is not None:
        return torch.arange(*args, device=_device(), **kw)
    return _np.arange(*args, **kw)


def linspace(a, b, n):
    if _device() is not None:
        return torch.linspace(a, b, n, dtype=torch.float64, device=_device())
    return _np.linspace(a, b, n)


# ---- elementwise ----------------------------------------------------------------
def _unary(npf, tf):
    def f(x):
        return tf(_t(x)) if _torch_mode(x) and (torch.is_tensor(x) or _device() is not None) else npf(x)
    f.__name__ = npf.__name__
    return f


sqrt = _unary(_np.sqrt, torch.sqrt)
exp = _unary(_np.exp, torch.exp)
log = _unary(_np.log, torch.log)
abs = _unary(_np.abs, torch.abs)
absolute = abs
sin = _unary(_np.sin, torch.sin)
cos = _unary(_np.cos, torch.cos)
tan = _unary(_np.tan, torch.tan)
tanh = _unary(_np.tanh, torch.tanh)
sinh = _unary(_np.sinh, torch.sinh)
cosh = _unary(_np.cosh, torch.cosh)
arctan = _unary(_np.arctan, torch.atan)
arcsin = _unary(_np.arcsin, torch.asin)
arccos = _unary(_np.arccos, torch.acos)
sign = _unary(_np.sign, torch.sign)
square = _unary(_np.square, torch.square)
floor = _unary(_np.floor, torch.floor)
ceil = _unary(_np.ceil, torch.ceil)


def power(a, b):
    if _torch_mode(a, b):
        return torch.pow(_t(a), b if not isinstance(b, (_np.ndarray, list)) else _t(b))
    return _np.power(a, b)


def maximum(a, b):
    if _torch_mode(a, b):
        a, b = _maybe_t(a, b)
        return torch.maximum(a, b)
    return _np.maximum(a, b)


def minimum(a, b):
    if _torch_mode(a, b):
        a, b = _maybe_t(a, b)
        return torch.minimum(a, b)
    return _np.minimum(a, b)


def where(cond, x, y):
    if _torch_mode(cond, x, y):
        c = _t(cond)
        xt = x if torch.is_tensor(x) else _t(x)
        yt = y if torch.is_tensor(y) else _t(y)
        return torch.where(c.bool() if c.dtype != torch.bool else c, xt, yt)
    return _np.where(cond, x, y)


def clip(a, lo, hi):
    if _torch_mode(a):
        return torch.clamp(_t(a), lo, hi)
    return _np.clip(a, lo, hi)


def isclose(a, b, rtol=1e-5, atol=1e-8):
    if _has_tensor((a, b)):
        a, b = _maybe_t(a, b)
        return torch.isclose(a.to(torch.float64), b.to(torch.float64), rtol=rtol, atol=atol)
    return _np.isclose(a, b, rtol=rtol, atol=atol)


def logical_and(a, b):
    return torch.logical_and(_t(a), _t(b)) if _has_tensor((a, b)) else _np.logical_and(a, b)


def logical_or(a, b):
    return torch.logical_or(_t(a), _t(b)) if _has_tensor((a, b)) else _np.logical_or(a, b)


# ---- reductions and shape ---------------------------------------------------------
def _axis(axis):
    return tuple(axis) if isinstance(axis, (list, tuple)) else axis


def sum(a, axis=None, keepdims=False):
    if torch.is_tensor(a):
        return a.sum() if axis is None else a.sum(dim=_axis(axis), keepdim=keepdims)
    return _np.sum(a, axis=axis, keepdims=keepdims)


def mean(a, axis=None, keepdims=False):
    if torch.is_tensor(a):
        return a.mean() if axis is None else a.mean(dim=_axis(axis), keepdim=keepdims)
    return _np.mean(a, axis=axis, keepdims=keepdims)


def max(a, axis=None, keepdims=False):
    if torch.is_tensor(a):
        return a.max() if axis is None else torch.amax(a, dim=_axis(axis), keepdim=keepdims)
    return _np.max(a, axis=axis, keepdims=keepdims)


def min(a, axis=None, keepdims=False):
    if torch.is_tensor(a):
        return a.min() if axis is None else torch.amin(a, dim=_axis(axis), keepdim=keepdims)
    return _np.min(a, axis=axis, keepdims=keepdims)


def prod(a, axis=None):
    if torch.is_tensor(a):
        return a.prod() if axis is None else a.prod(dim=axis)
    return _np.prod(a, axis=axis)


def trace(a, offset=0, axis1=0, axis2=1):
    if torch.is_tensor(a):
        return torch.diagonal(a, offset=offset, dim1=axis1, dim2=axis2).sum(-1)
    return _np.trace(a, offset, axis1, axis2)


def diag(a):
    return torch.diag(a) if torch.is_tensor(a) else _np.diag(a)


def transpose(a, axes=None):
    if torch.is_tensor(a):
        return a.permute(*(axes if axes is not None else reversed(range(a.dim()))))
    return _np.transpose(a, axes)


def swapaxes(a, i, j):
    return a.transpose(i, j) if torch.is_tensor(a) else _np.swapaxes(a, i, j)


def reshape(a, shape):
    return a.reshape(shape) if torch.is_tensor(a) else _np.reshape(a, shape)


def expand_dims(a, axis):
    return a.unsqueeze(axis) if torch.is_tensor(a) else _np.expand_dims(a, axis)


def squeeze(a, axis=None):
    if torch.is_tensor(a):
        return a.squeeze() if axis is None else a.squeeze(axis)
    return _np.squeeze(a, axis)


def stack(arrays, axis=0):
    if _has_tensor(arrays) or _device() is not None:
        return torch.stack([_t(a) for a in arrays], dim=axis)
    return _np.stack(arrays, axis=axis)


def concatenate(arrays, axis=0):
    if _has_tensor(arrays) or _device() is not None:
        return torch.cat([_t(a) for a in arrays], dim=axis)
    return _np.concatenate(arrays, axis=axis)


def hstack(arrays):
    if _has_tensor(arrays) or _device() is not None:
        return torch.hstack([_t(a) for a in arrays])
    return _np.hstack(arrays)


def vstack(arrays):
    if _has_tensor(arrays) or _device() is not None:
        return torch.vstack([_t(a) for a in arrays])
    return _np.vstack(arrays)


# ---- products ---------------------------------------------------------------------
def dot(a, b):
    if _torch_mode(a, b):
        a, b = _maybe_t(a, b)
        if a.dim() <= 1 and b.dim() <= 1:
            return (a * b).sum()
        return torch.matmul(a, b)
    return _np.dot(a, b)


def matmul(a, b):
    if _torch_mode(a, b):
        a, b = _maybe_t(a, b)
        return torch.matmul(a, b)
    return _np.matmul(a, b)


def outer(a, b):
    if _torch_mode(a, b):
        a, b = _maybe_t(a, b)
        return torch.outer(a.reshape(-1), b.reshape(-1))
    return _np.outer(a, b)


def einsum(spec, *ops):
    if _torch_mode(*ops):
        return torch.einsum(spec, *[_t(o) for o in ops])
    return _np.einsum(spec, *ops)


def tensordot(a, b, axes=2):
    if _torch_mode(a, b):
        a, b = _maybe_t(a, b)
        return torch.tensordot(a, b, dims=axes)
    return _np.tensordot(a, b, axes=axes)


def cross(a, b):
    if _torch_mode(a, b):
        a, b = _maybe_t(a, b)
        return torch.linalg.cross(a, b)
    return _np.cross(a, b)


def _det_small(a):
    """Determinant of (..., n, n), n <= 3, by cofactors: a polynomial, so every
    derivative is exact (torch.linalg.det's second derivative is NaN at the
    identity, where every strain energy is evaluated first)."""
    n = a.shape[-1]
    if n == 1:
        return a[..., 0, 0]
    if n == 2:
        return a[..., 0, 0] * a[..., 1, 1] - a[..., 0, 1] * a[..., 1, 0]
    if n == 3:
        return (a[..., 0, 0] * (a[..., 1, 1] * a[..., 2, 2] - a[..., 1, 2] * a[..., 2, 1])
                - a[..., 0, 1] * (a[..., 1, 0] * a[..., 2, 2] - a[..., 1, 2] * a[..., 2, 0])
                + a[..., 0, 2] * (a[..., 1, 0] * a[..., 2, 1] - a[..., 1, 1] * a[..., 2, 0]))
    return torch.linalg.det(a)


def _inv_small(a):
    """Inverse of (..., n, n), n <= 3, as adjugate / determinant (exact derivatives)."""
    n = a.shape[-1]
    if n == 1:
        return 1.0 / a
    d = _det_small(a)[..., None, None]
    if n == 2:
        adj = torch.stack([torch.stack([a[..., 1, 1], -a[..., 0, 1]], -1),
                           torch.stack([-a[..., 1, 0], a[..., 0, 0]], -1)], -2)
        return adj / d
    if n == 3:
        c = lambda i, j, k, l: a[..., i, j] * a[..., k, l]
        adj = torch.stack([
            torch.stack([c(1, 1, 2, 2) - c(1, 2, 2, 1), c(0, 2, 2, 1) - c(0, 1, 2, 2), c(0, 1, 1, 2) - c(0, 2, 1, 1)], -1),
            torch.stack([c(1, 2, 2, 0) - c(1, 0, 2, 2), c(0, 0, 2, 2) - c(0, 2, 2, 0), c(0, 2, 1, 0) - c(0, 0, 1, 2)], -1),
            torch.stack([c(1, 0, 2, 1) - c(1, 1, 2, 0), c(0, 1, 2, 0) - c(0, 0, 2, 1), c(0, 0, 1, 1) - c(0, 1, 1, 0)], -1),
        ], -2)
        return adj / d
    return torch.linalg.inv(a)


class _Linalg:
    @staticmethod
    def det(a):
        return _det_small(a) if torch.is_tensor(a) else _np.linalg.det(a)

    @staticmethod
    def inv(a):
        return _inv_small(a) if torch.is_tensor(a) else _np.linalg.inv(a)

    @staticmethod
    def norm(a, ord=None, axis=None):
        if torch.is_tensor(a):
            if axis is None and ord is None:
                return torch.sqrt((a * a).sum())
            return torch.linalg.norm(a, ord=ord, dim=axis)
        return _np.linalg.norm(a, ord=ord, axis=axis)

    @staticmethod
    def eigh(a):
        return torch.linalg.eigh(a) if torch.is_tensor(a) else _np.linalg.eigh(a)

    @staticmethod
    def solve(a, b):
        return torch.linalg.solve(a, b) if _has_tensor((a, b)) else _np.linalg.solve(a, b)


linalg = _Linalg()
