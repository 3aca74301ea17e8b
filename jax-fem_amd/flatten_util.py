"""``ravel_pytree`` for user kernels (the role of jax.flatten_util.ravel_pytree in the
reference's universal kernels, e.g. applications/stokes/example.py:60-62): a nest of
lists / tuples of arrays -> (one flat vector, the function that rebuilds the nest).
Torch tensors stay torch (on their device, also under torch.func transforms);
NumPy arrays stay NumPy."""
import numpy as np
import torch


def _leaves(tree):
    if isinstance(tree, (list, tuple)):
        out = []
        for t in tree:
            out += _leaves(t)
        return out
    return [tree]


def _rebuild(tree, it):
    if isinstance(tree, (list, tuple)):
        return type(tree)(_rebuild(t, it) for t in tree)
    return next(it)


def ravel_pytree(tree):
    leaves = _leaves(tree)
    use_torch = any(torch.is_tensor(l) for l in leaves)
    if use_torch:
        dev = next(l.device for l in leaves if torch.is_tensor(l))
        ts = [l if torch.is_tensor(l) else torch.as_tensor(np.asarray(l), dtype=torch.float64, device=dev)
              for l in leaves]
        flat = torch.cat([t.reshape(-1) for t in ts]) if ts else torch.zeros(0, dtype=torch.float64, device=dev)
        shapes = [tuple(t.shape) for t in ts]
    else:
        arrs = [np.asarray(l) for l in leaves]
        flat = np.concatenate([a.reshape(-1) for a in arrs]) if arrs else np.zeros(0)
        shapes = [a.shape for a in arrs]

    def unravel(v):
        parts, k = [], 0
        for shp in shapes:
            n = int(np.prod(shp)) if shp else 1
            parts.append(v[k:k + n].reshape(shp))
            k += n
        return _rebuild(tree, iter(parts))
    return flat, unravel
