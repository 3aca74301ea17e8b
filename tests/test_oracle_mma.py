"""The MMA oracle (oracle/mma.py) against mathematics and library routines:
the subproblem's KKT system at the returned point, a separable problem with a
closed-form optimum, Svanberg's two-constraint toy problem against scipy's
SLSQP, and the filter weights against brute force over all centroid pairs."""
import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import fe as ofe
from oracle import mma


def test_subproblem_meets_its_kkt_conditions():
    rng = np.random.default_rng(0)
    n, m = 400, 2
    x = rng.uniform(0.2, 0.8, n)
    xmin, xmax = np.zeros(n), np.ones(n)
    df0 = rng.standard_normal(n)
    dfdx = rng.standard_normal((m, n))
    fval = np.array([0.1, -0.2])
    for it in (1, 5):
        xold1 = np.clip(x + 0.01 * rng.standard_normal(n), 0, 1)
        xold2 = np.clip(x + 0.01 * rng.standard_normal(n), 0, 1)
        low, upp = x - 0.3, x + 0.3
        xm, low2, upp2, sol = mma.mmasub(it, x, xold1, xold2, xmin, xmax, low, upp, df0, fval, dfdx, 1.0,
                                         np.zeros(m), 1e4 * np.ones(m), np.zeros(m), 0.2)
        assert np.all(low2 < x) and np.all(x < upp2)
        # bounds: alfa <= x <= beta, inside the move limit
        assert np.all(xm >= np.maximum(xmin, x - 0.2) - 1e-12) and np.all(xm <= np.minimum(xmax, x + 0.2) + 1e-12)
        ux2, xl2 = (upp2 - x) ** 2, (x - low2) ** 2
        box = xmax - xmin
        alfa = np.maximum(np.maximum(low2 + 0.1 * (x - low2), x - 0.2 * box), xmin)
        beta = np.minimum(np.minimum(upp2 - 0.1 * (upp2 - x), x + 0.2 * box), xmax)
        reg = 0.001 * np.abs(df0) + 1e-5
        p0 = (np.maximum(df0, 0) + reg) * ux2
        q0 = (np.maximum(-df0, 0) + reg) * xl2
        regm = 0.001 * np.abs(dfdx) + 1e-5
        P = (np.maximum(dfdx, 0) + regm) * ux2
        Q = (np.maximum(-dfdx, 0) + regm) * xl2
        r = mma.kkt_residual(sol, low2, upp2, alfa, beta, p0, q0, P, Q, 1.0, np.zeros(m), sol["b"],
                             1e4 * np.ones(m), np.zeros(m), epsi=1e-7)
        assert r <= 0.9e-7 * 1.0000001
        # the approximations are exact to first order at x: the subproblem's gradient matches df0
        dpsi = p0 / ux2 - q0 / xl2
        np.testing.assert_allclose(dpsi, df0 + 0.0, rtol=0, atol=2.1e-3 * np.abs(df0).max() + 2e-5)


def test_separable_problem_reaches_closed_form():
    """min sum c_i / x_i  s.t.  mean(x) <= v: x_i = v sqrt(c_i) / mean(sqrt(c))."""
    rng = np.random.default_rng(1)
    n, v = 200, 0.4
    c = rng.uniform(1.0, 2.0, n)
    obj = lambda x: (float(np.sum(c / x)), -c / x ** 2)
    con = lambda x, it: (np.array([np.mean(x) / v - 1.0]), np.full((1, n), 1.0 / (n * v)))
    x, hist = mma.optimize_loop(np.full(n, 0.5), 0.2, 60, obj, con, sensitivity=False)
    exact = v * np.sqrt(c) / np.mean(np.sqrt(c))
    np.testing.assert_allclose(x, exact, rtol=2e-5)
    assert hist[-1] <= hist[5]


def test_svanberg_toy_problem_against_slsqp():
    """min x1^2+x2^2+x3^2 s.t. |x-(5,2,1)|^2 <= 9, |x-(3,4,3)|^2 <= 9, 0 <= x <= 5."""
    c1, c2 = np.array([5.0, 2.0, 1.0]), np.array([3.0, 4.0, 3.0])
    x = np.array([4.0, 3.0, 2.0])
    xold1, xold2 = x.copy(), x.copy()
    low, upp = np.zeros(3), np.full(3, 5.0)
    xmin, xmax = np.zeros(3), np.full(3, 5.0)
    for it in range(1, 41):
        fval = np.array([np.sum((x - c1) ** 2) - 9, np.sum((x - c2) ** 2) - 9])
        dfdx = np.stack([2 * (x - c1), 2 * (x - c2)])
        xm, low, upp, _ = mma.mmasub(it, x, xold1, xold2, xmin, xmax, low, upp, 2 * x, fval, dfdx, 1.0,
                                     np.zeros(2), 1e4 * np.ones(2), np.zeros(2), 1.0)
        xold2, xold1, x = xold1, x, xm
    ref = minimize(lambda z: z @ z, np.array([4.0, 3.0, 2.0]), jac=lambda z: 2 * z, method="SLSQP",
                   bounds=[(0, 5)] * 3, tol=1e-14,
                   constraints=[{"type": "ineq", "fun": lambda z, cc=cc: 9 - np.sum((z - cc) ** 2),
                                 "jac": lambda z, cc=cc: -2 * (z - cc)} for cc in (c1, c2)])
    np.testing.assert_allclose(x, ref.x, atol=1e-6)


@pytest.mark.parametrize("shape", [(6, 5), (4, 3, 3)])
def test_filter_weights_brute_force(shape):
    if len(shape) == 2:
        pts, cells = ofe.rectangle_mesh(*shape, 1.2, 1.0)
        vol = 1.2
    else:
        pts, cells = ofe.box_mesh(*shape, 1.0, 0.75, 0.75)
        vol = 0.5625
    dim = len(shape)
    flex = np.arange(len(cells))[::2] if dim == 2 else None
    H, Hs, rmin = mma.filter_matrix(pts, cells, vol, dim, flex)
    cent = pts[cells].mean(1)
    if flex is not None:
        cent = cent[flex]
    D = np.linalg.norm(cent[:, None] - cent[None], axis=-1)
    W = np.maximum(rmin - D, 0.0)
    # 20 nearest neighbours cover every centroid closer than rmin on these meshes
    assert np.all((D < rmin).sum(1) <= 20)
    np.testing.assert_allclose(H.toarray(), W, atol=1e-15)
    np.testing.assert_allclose(Hs, W.sum(1), atol=1e-14)
    rho = np.linspace(0.0005, 1.0, len(cent))[:, None]
    g = np.cos(np.arange(len(cent)))[:, None]
    dJ, dvc = mma.sensitivity_filter(H, Hs, rho, g, np.stack([g, 2 * g]))
    scale = rho[:, 0] / np.maximum(1e-3, rho[:, 0]) / Hs
    np.testing.assert_allclose(dJ[:, 0], W @ (scale * g[:, 0]), atol=1e-14)
    np.testing.assert_allclose(dvc[1, :, 0], 2 * dJ[:, 0], atol=1e-14)
    np.testing.assert_allclose(mma.density_filter(H, Hs, rho)[:, 0], W @ rho[:, 0] / W.sum(1), atol=1e-14)
