"""Transforms for user-written pointwise maps: ``grad``, ``vmap``, ``jacfwd``.

The reference's users differentiate and vectorise their maps with ``jax.grad``
and ``jax.vmap`` (e.g. ``P_fn = jax.grad(psi)`` in
tests/benchmarks/hyperelasticity/test_hyper_elasticity.py:15-34, and
``jax.vmap(jax.vmap(stress_return_map))`` in
tests/benchmarks/plasticity/test_plasticity.py:66-78).  These are the same
transforms here:

* on torch tensors they are torch.func's (exact forward- and reverse-mode
  derivatives, batched evaluation), run with ``jax_fem_amd.numpy`` creating its
  constants on the tensors' device;
* a function that torch.func cannot transform but torch autograd can
  differentiate -- an objective through ``ad_wrapper``'s implicit solve, as
  ``jax.grad(J)`` in the reference's inverse-problem demos -- is differentiated
  by reverse-mode autograd on the eager graph (``grad``) or row by row
  (``jacfwd``);
* on NumPy arrays (post-processing, the Problem's probe) ``grad`` and ``jacfwd``
  run the above on CPU tensors and return NumPy arrays when the map is written
  with ``jax_fem_amd.numpy``; a map written with NumPy itself is differentiated
  by the complex step, f'(x) = Im f(x + i h) / h with h = 1e-30 (Squire and
  Trapp 1998; exact to rounding for analytic maps such as the hyperelastic
  energies); ``vmap`` maps over the leading axis.
"""
import numpy as np
import torch

from . import numpy as jnp

_H = 1e-30


def _first_tensor(args):
    for a in args:
        if torch.is_tensor(a):
            return a
        if isinstance(a, (list, tuple)):
            t = _first_tensor(a)
            if t is not None:
                return t
    return None


def _to_device(args, dev):
    out = []
    for a in args:
        if isinstance(a, np.ndarray) or isinstance(a, (float, int)) and not isinstance(a, bool):
            a = torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
        out.append(a)
    return out


def _host(out):
    """A transform's result on NumPy inputs: NumPy arrays that meet torch tensors
    on their device (fe.HostField)."""
    from .fe import _field
    if isinstance(out, tuple):
        return tuple(_host(o) for o in out)
    if torch.is_tensor(out):
        out = out.detach().cpu().numpy()
    return _field(np.asarray(out))


def vmap(f, in_axes=0, out_axes=0):
    def run(*args):
        t = _first_tensor(args)
        if t is None:
            try:          # maps written with jax_fem_amd.numpy: one batched evaluation
                with jnp.on_device(torch.device("cpu")):
                    out = torch.func.vmap(f, in_dims=in_axes, out_dims=out_axes)(
                        *_to_device(args, torch.device("cpu")))
                return _host(out)
            except Exception:
                pass
            axes = in_axes if isinstance(in_axes, (list, tuple)) else [in_axes] * len(args)
            n = next(np.shape(a)[ax] for a, ax in zip(args, axes) if ax is not None)
            outs = [f(*[a if ax is None else np.take(a, k, axis=ax) for a, ax in zip(args, axes)]) for k in range(n)]
            if isinstance(outs[0], tuple):
                return _host(tuple(np.stack([o[j] for o in outs], axis=out_axes) for j in range(len(outs[0]))))
            return _host(np.stack([np.asarray(o) for o in outs], axis=out_axes))
        dev = t.device
        with jnp.on_device(dev):
            return torch.func.vmap(f, in_dims=in_axes, out_dims=out_axes)(*_to_device(args, dev))
    return run


def _complex_step(f, args, argnums, vector_out):
    x = np.asarray(args[argnums], dtype=np.float64)
    flat = x.reshape(-1)
    cols = []
    for k in range(flat.size):
        xc = flat.astype(np.complex128)
        xc[k] += 1j * _H
        a = list(args)
        a[argnums] = xc.reshape(x.shape)
        cols.append(np.imag(np.asarray(f(*a))) / _H)
    if not vector_out:
        return np.array(cols).reshape(x.shape)
    y = np.stack(cols, axis=-1)                      # (out..., in)
    return y.reshape(y.shape[:-1] + x.shape)


def _eager(f, args, argnums, vector_out, with_value=False):
    """Derivative of f with respect to args[argnums] by torch autograd on the eager
    graph (reverse mode; a Jacobian row by row).  Raises TypeError when f does not
    produce a tensor connected to its argument."""
    a = list(args)
    x = a[argnums].detach().clone().requires_grad_(True)
    a[argnums] = x
    with torch.enable_grad():
        y = f(*a)
    if not torch.is_tensor(y) or y.grad_fn is None:
        raise TypeError("the function's output is not a differentiable tensor of its argument")
    if not vector_out:
        if y.numel() != 1:
            raise TypeError("grad needs a scalar-valued function")
        g = torch.autograd.grad(y, x)[0]
        return (y.detach(), g) if with_value else g
    flat = y.reshape(-1)
    rows = [torch.autograd.grad(flat[k], x, retain_graph=k + 1 < flat.numel(), allow_unused=True)[0]
            for k in range(flat.numel())]
    rows = [torch.zeros_like(x) if r is None else r for r in rows]
    return torch.stack(rows).reshape(tuple(y.shape) + tuple(x.shape))


def _torch_call(transform, f, args, argnums, vector_out):
    """torch.func's transform; where it refuses the function (an autograd.Function
    without functorch support, as the implicit solve), eager autograd."""
    try:
        return transform(f, argnums=argnums)(*args)
    except Exception as err:
        if not isinstance(argnums, int):
            raise
        try:
            return _eager(f, args, argnums, vector_out)
        except Exception:
            raise err


def value_and_grad(f, argnums=0):
    """(f(x), grad f(x)) in one evaluation, as jax.value_and_grad."""
    def both(*args):
        try:
            g, v = torch.func.grad_and_value(f, argnums=argnums)(*args)
            return v, g
        except Exception as err:
            try:
                return _eager(f, args, argnums, False, with_value=True)
            except Exception:
                raise err

    def run(*args):
        t = _first_tensor(args)
        if t is None:
            try:
                with jnp.on_device(torch.device("cpu")):
                    v, g = both(*_to_device(args, torch.device("cpu")))
                return _host(v), _host(g)
            except Exception:
                return _host(np.asarray(f(*args))), _host(_complex_step(f, args, argnums, False))
        with jnp.on_device(t.device):
            return both(*_to_device(args, t.device))
    return run


def _numpy_call(transform, f, args, argnums, vector_out):
    """A transform on NumPy arguments: torch (torch.func, else eager autograd) on CPU
    tensors when the function runs on them (maps written with jax_fem_amd.numpy,
    objectives through ad_wrapper), else the complex step on the NumPy map."""
    try:
        with jnp.on_device(torch.device("cpu")):
            out = _torch_call(transform, f, _to_device(args, torch.device("cpu")), argnums, vector_out)
        return _host(out)
    except Exception:
        return _host(_complex_step(f, args, argnums, vector_out))


def grad(f, argnums=0):
    def run(*args):
        t = _first_tensor(args)
        if t is None:
            return _numpy_call(torch.func.grad, f, args, argnums, False)
        with jnp.on_device(t.device):
            return _torch_call(torch.func.grad, f, _to_device(args, t.device), argnums, False)
    return run


def jacfwd(f, argnums=0):
    def run(*args):
        t = _first_tensor(args)
        if t is None:
            return _numpy_call(torch.func.jacfwd, f, args, argnums, True)
        with jnp.on_device(t.device):
            return _torch_call(torch.func.jacfwd, f, _to_device(args, t.device), argnums, True)
    return run


jacrev = jacfwd


def hessian(f, argnums=0):
    """Second derivative (jax.hessian): torch.func.hessian on tensors; on NumPy
    arrays torch.func on CPU tensors, else central differences (step 1e-6) of the
    complex-step gradient."""
    def run(*args):
        t = _first_tensor(args)
        if t is not None:
            with jnp.on_device(t.device):
                return torch.func.hessian(f, argnums=argnums)(*_to_device(args, t.device))
        try:
            with jnp.on_device(torch.device("cpu")):
                out = torch.func.hessian(f, argnums=argnums)(*_to_device(args, torch.device("cpu")))
            return _host(out)
        except Exception:
            x = np.asarray(args[argnums], dtype=np.float64)
            g = lambda z: _complex_step(f, [z if k == argnums else a for k, a in enumerate(args)], argnums,
                                        True)
            h = 1e-6
            cols = []
            for k in range(x.size):
                e = np.zeros(x.size)
                e[k] = h
                cols.append((g(x + e.reshape(x.shape)) - g(x - e.reshape(x.shape))) / (2 * h))
            H = np.stack(cols, axis=-1)
            return _host(H.reshape(H.shape[:-1] + x.shape))
    return run
