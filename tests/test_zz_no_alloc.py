"""The allocation-free Krylov entry point (hf_krylov_solve_noalloc_f64) against
hf_krylov_solve_f64 with a caller workspace: same bits, same iteration count
and status, on two calls in a row (the second on a smaller system), PCG and
BiCGSTAB; sizes past the static scratch are rejected."""
import ctypes
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def _system(jfa, n):
    """FP64 row-eliminated elasticity tangent of an n x n x n box and a right-hand side
    that is zero on the Dirichlet rows."""
    class P(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=10.0, nu=0.3)

    left = lambda p: np.isclose(p[0], 0.0, atol=1e-7)
    prob = P(jfa.box_mesh(n, n, n), vec=3, dim=3, dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.0] * 3])
    from importlib import import_module
    S = import_module("jax_fem_amd.solver")
    A = S.TangentMatrix(prob, prob.jacobian_bsr(torch.zeros(prob.N, dtype=torch.float64, device=DEV)))
    g = torch.Generator(device="cpu").manual_seed(n)
    b = torch.randn(prob.N, generator=g, dtype=torch.float64).to(DEV)
    d, _ = prob.bc_arrays()
    b[d.long()] = 0.0
    return prob, A, b


def _solve(jfa, A, b, method, noalloc):
    L = jfa._lib
    x = torch.zeros_like(b)
    info = L.HFKrylovInfo()
    dinv = A.dinv()
    args = (A.pattern.handle, L.ptr(A.vals), L.ptr(dinv), L.ptr(b), L.ptr(x), 0, 5000, 1e-10, 1e-10, 8)
    if noalloc:
        L.call("hf_krylov_solve_noalloc_f64", method, *args, ctypes.byref(info), L.stream())
    else:
        ws = torch.empty(int(L.load().hf_krylov_workspace_size(method, A.shape[0], 8)), dtype=torch.uint8, device=DEV)
        L.call("hf_krylov_solve_f64", method, *args, L.ptr(ws), ctypes.byref(info), L.stream())
    torch.cuda.synchronize()
    return x, (info.iterations, info.status, info.res_norm)


@pytest.mark.parametrize("method", [0, 1])
def test_noalloc_matches_workspace_solve(jfa, method):
    L = jfa._lib
    L.call("hf_init")
    for n in (6, 4):    # second call on a smaller system: the scratch keeps the first call's contents
        say("noalloc krylov method %d, %d^3 cells" % (method, n))
        prob, A, b = _system(jfa, n)
        x0, i0 = _solve(jfa, A, b, method, False)
        x1, i1 = _solve(jfa, A, b, method, True)
        assert i0[1] == 0 and i0 == i1, (i0, i1)
        assert torch.equal(x0, x1)
        r = torch.linalg.norm(A.matvec(x1) - b) / torch.linalg.norm(b)
        assert r < 1e-9


def test_noalloc_rejects_oversized_system(jfa):
    """A system whose BiCGSTAB workspace exceeds the static array (70^3 cells,
    1,073,733 DOF) is refused by the host-side size check before any launch."""
    L = jfa._lib
    L.call("hf_init")

    class P(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=10.0, nu=0.3)

    prob = P(jfa.box_mesh(70, 70, 70), vec=3, dim=3)
    assert int(L.load().hf_krylov_workspace_size(1, prob.N, 8)) > 64 << 20
    say("noalloc size check, N =", prob.N)
    h = prob.pattern.handle
    buf = torch.zeros(prob.N, dtype=torch.float64, device=DEV)
    info = L.HFKrylovInfo()
    rc = L.load().hf_krylov_solve_noalloc_f64(1, h, L.ptr(buf), L.ptr(buf), L.ptr(buf), L.ptr(buf), 0, 10, 1e-10,
                                               1e-10, 8, ctypes.byref(info), L.stream())
    assert rc == -1
    assert "static scratch" in L.last_error()
