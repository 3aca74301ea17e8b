"""CPU restatement of the learned preconditioner's network (oracle, FP64).

The network has no reference counterpart; this module is the checker for the
HIP forward/backward kernels (jax-fem_amd/csrc/lpc.hip):
    f  = [r sqrt(dinv) / s, g],  s = sqrt(sum(r^2 dinv) / N)
    h1 = tanh(conv1(f)), h2 = tanh(conv2(h1)), y = conv3(h2)      (3x3x3, zero padding)
    z  = dinv r + s sqrt(dinv) mask y
with W[tap][ci][co], tap = ((di+1)*3 + (dj+1))*3 + (dk+1) acting on the node
(i+di, j+dj, k+dk).  Convolutions are torch.nn.functional.conv3d on the CPU in
float64 (a library routine), so gradients come from torch autograd; the test
suite pins the tap / channel convention with a direct loop over one node.
"""
import numpy as np
import torch


def to_torch_weight(W):
    """W (27, ci, co) -> torch conv3d weight (co, ci, 3, 3, 3)."""
    W = torch.as_tensor(np.asarray(W), dtype=torch.float64)
    ci, co = W.shape[1], W.shape[2]
    return W.reshape(3, 3, 3, ci, co).permute(4, 3, 0, 1, 2).contiguous()


def conv(x, W, b, grid):
    """x (B, nnode, ci) -> (B, nnode, co), channels-last node order n = (i*ny + j)*nz + k."""
    nx, ny, nz = grid
    B, _, ci = x.shape
    xt = x.reshape(B, nx, ny, nz, ci).permute(0, 4, 1, 2, 3)
    y = torch.nn.functional.conv3d(xt, W, b, padding=1)
    return y.permute(0, 2, 3, 4, 1).reshape(B, nx * ny * nz, -1)


def forward(params, R, dinv, g, mask, grid):
    """params: [W1, b1, W2, b2, W3, b3] as torch float64 tensors (W in torch layout)."""
    W1, b1, W2, b2, W3, b3 = params
    B, N = R.shape
    s = torch.sqrt((R * R * dinv).sum(1) / N)
    nn_ = N // 3
    f = torch.cat([(R * torch.sqrt(dinv) / s[:, None]).reshape(B, nn_, 3),
                   g.reshape(1, nn_, 1).expand(B, nn_, 1)], dim=2)
    h1 = torch.tanh(conv(f, W1, b1, grid))
    h2 = torch.tanh(conv(h1, W2, b2, grid))
    y = conv(h2, W3, b3, grid).reshape(B, N)
    return dinv * R + s[:, None] * torch.sqrt(dinv) * mask * y


def conv_direct_node(x, W, b, grid, n):
    """One output node of a layer by the defining sum (pins tap/channel order)."""
    nx, ny, nz = grid
    i, j, k = n // (ny * nz), (n // nz) % ny, n % nz
    out = np.array(b, dtype=np.float64).copy()
    for tap in range(27):
        di, dj, dk = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        ii, jj, kk = i + di, j + dj, k + dk
        if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
            m = (ii * ny + jj) * nz + kk
            out += np.asarray(x[m]) @ np.asarray(W[tap])
    return out
