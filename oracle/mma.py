"""MMA optimiser and filters for topology optimisation (oracle, FP64, CPU).

Test infrastructure: a NumPy restatement that the GPU optimiser is checked
against.  Restates ``jax_fem/mma.py``:
  * ``compute_filter_kd_tree`` (mma.py:27-56): weights max(rmin - d, 0) to the
    20 nearest cell centroids (k-d tree), rmin = 1.5 (V / ncell)^(1/dim), H as a
    sparse matrix with rows = cells, Hs its row sums;
  * ``applySensitivityFilter`` / ``applyDensityFilter`` (mma.py:58-66):
    dJ <- H (rho dJ / max(1e-3, rho) / Hs), rho <- H rho / Hs;
  * ``MMA.mmasub`` (mma.py:114-205): moving asymptotes (initial 0.5 of the box,
    then x1.2 / x0.7 by the sign of the last two steps, clamped to
    [0.01, 10] box widths), move limits, the convex approximation's p0, q0, P, Q
    with the 0.001 / 1e-5 regularisation, b = P/(U-x) + Q/(x-L) - f;
  * ``subsolv`` (mma.py:207-410): Svanberg's primal-dual interior-point Newton
    method for the subproblem, branch m < n (the reduced (m+1)x(m+1) system),
    epsi from 1 down to 1e-7 by factors of 10, at most 200 Newton steps per epsi,
    step to 1/1.01 of the distance to the bounds, residual-norm backtracking
    (halving, at most 50 times);
  * ``optimize`` (mma.py:415-528): x in [0, 1], a0 = 1, a = 0, c = 1e4, d = 0,
    sensitivity filtering on, density filtering off.
Svanberg, "The method of moving asymptotes", IJNME 24 (1987); the subproblem
solver as in his MMA notes (2007).

Pinned by: the subproblem's KKT conditions at the returned point (every
residual of the interior-point system <= 0.9 epsimin), a separable problem with
a closed-form optimum (min sum c_i/x_i s.t. mean x <= v: x_i proportional to
sqrt(c_i)), Svanberg's two-constraint toy problem against scipy's SLSQP
(library routine), and the filter weights against brute force over all pairs
of centroids.
"""
import numpy as np
import scipy.sparse as sp
from scipy.spatial import KDTree


# -- filters (mma.py:27-66) ----------------------------------------------------
def filter_matrix(points, cells, jxw_total, dim, flex_inds=None, k=20):
    """(H (csr, nflex x nflex), Hs (nflex,)) over the centroids of flex_inds."""
    cent = points[cells].mean(axis=1)
    if flex_inds is not None:
        cent = cent[flex_inds]
    nflex = len(cent)
    rmin = 1.5 * (jxw_total / len(cells)) ** (1.0 / dim)
    k = min(k, nflex)
    dd, ii = KDTree(cent).query(cent, k)
    dd, ii = dd.reshape(nflex, k), ii.reshape(nflex, k)
    w = np.maximum(rmin - dd, 0.0)
    H = sp.csr_matrix((w.reshape(-1), (np.repeat(np.arange(nflex), k), ii.reshape(-1))), shape=(nflex, nflex))
    return H, np.asarray(H.sum(axis=1)).reshape(-1), rmin


def sensitivity_filter(H, Hs, rho, dJ, dvc):
    """rho, dJ (n,1); dvc (m,n,1)."""
    r = rho.reshape(-1)
    scale = r / np.maximum(1e-3, r) / Hs
    dJf = H @ (scale * dJ.reshape(-1))
    dvcf = np.stack([H @ (scale * g.reshape(-1)) for g in dvc])
    return dJf.reshape(dJ.shape), dvcf.reshape(dvc.shape)


def density_filter(H, Hs, rho):
    return (H @ rho.reshape(-1) / Hs).reshape(rho.shape)


# -- the subproblem (mma.py:114-205) -------------------------------------------
def mmasub(it, xval, xold1, xold2, xmin, xmax, low, upp, df0dx, fval, dfdx, a0, a, c, d, move):
    """One MMA subproblem.  Vectors are 1-D (n,); fval, a, c, d (m,); dfdx (m, n).
    Returns (xmma, low, upp, info) with info the subproblem's final state."""
    box = xmax - xmin
    if it <= 2:
        low = xval - 0.5 * box
        upp = xval + 0.5 * box
    else:
        trend = (xval - xold1) * (xold1 - xold2)
        factor = np.where(trend > 0, 1.2, np.where(trend < 0, 0.7, 1.0))
        low = xval - factor * (xold1 - low)
        upp = xval + factor * (upp - xold1)
        low = np.minimum(np.maximum(low, xval - 10 * box), xval - 0.01 * box)
        upp = np.maximum(np.minimum(upp, xval + 10 * box), xval + 0.01 * box)
    alfa = np.maximum(np.maximum(low + 0.1 * (xval - low), xval - move * box), xmin)
    beta = np.minimum(np.minimum(upp - 0.1 * (upp - xval), xval + move * box), xmax)
    inv_box = 1.0 / np.maximum(box, 1e-5)
    ux2 = (upp - xval) ** 2
    xl2 = (xval - low) ** 2
    pos, neg = np.maximum(df0dx, 0.0), np.maximum(-df0dx, 0.0)
    reg = 0.001 * (pos + neg) + 1e-5 * inv_box
    p0 = (pos + reg) * ux2
    q0 = (neg + reg) * xl2
    Pp, Qn = np.maximum(dfdx, 0.0), np.maximum(-dfdx, 0.0)
    reg = 0.001 * (Pp + Qn) + 1e-5 * inv_box[None, :]
    P = (Pp + reg) * ux2[None, :]
    Q = (Qn + reg) * xl2[None, :]
    b = P @ (1.0 / (upp - xval)) + Q @ (1.0 / (xval - low)) - fval
    sol = subsolv(low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d)
    return sol["x"], low, upp, sol


def _residual(v, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d, epsi):
    x, y, z, lam, xsi, eta, mu, zet, s = (v[k] for k in ("x", "y", "z", "lam", "xsi", "eta", "mu", "zet", "s"))
    ux, xl = upp - x, x - low
    plam = p0 + P.T @ lam
    qlam = q0 + Q.T @ lam
    gvec = P @ (1.0 / ux) + Q @ (1.0 / xl)
    parts = [plam / ux ** 2 - qlam / xl ** 2 - xsi + eta,          # d/dx of the Lagrangian
             c + d * y - mu - lam,                                 # d/dy
             np.array([a0 - zet - a @ lam]),                       # d/dz
             gvec - a * z - y + s - b,                             # constraints
             xsi * (x - alfa) - epsi, eta * (beta - x) - epsi,     # perturbed complementarity
             mu * y - epsi, np.array([zet * z - epsi]), lam * s - epsi]
    r = np.concatenate(parts)
    return np.sqrt(r @ r), np.max(np.abs(r))


def subsolv(low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d, epsimin=1e-7):
    """Primal-dual Newton for the MMA subproblem (mma.py:207-410), m < n."""
    m, n = P.shape
    assert m < n
    v = {"x": 0.5 * (alfa + beta), "y": np.ones(m), "z": 1.0, "lam": np.ones(m), "s": np.ones(m),
         "zet": 1.0, "mu": np.maximum(1.0, 0.5 * c)}
    v["xsi"] = np.maximum(1.0 / (v["x"] - alfa), 1.0)
    v["eta"] = np.maximum(1.0 / (beta - v["x"]), 1.0)
    args = (low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d)
    epsi = 1.0
    newton_steps = 0
    trials = 0
    while epsi > epsimin:
        rnorm, rmax = _residual(v, *args, epsi)
        k = 0
        while rmax > 0.9 * epsi and k < 200:
            k += 1
            newton_steps += 1
            x, y, z, lam, xsi, eta, mu, zet, s = (v[q] for q in ("x", "y", "z", "lam", "xsi", "eta", "mu", "zet", "s"))
            ux, xl = upp - x, x - low
            plam = p0 + P.T @ lam
            qlam = q0 + Q.T @ lam
            gvec = P @ (1.0 / ux) + Q @ (1.0 / xl)
            GG = P / ux[None, :] ** 2 - Q / xl[None, :] ** 2
            delx = plam / ux ** 2 - qlam / xl ** 2 - epsi / (x - alfa) + epsi / (beta - x)
            dely = c + d * y - lam - epsi / y
            delz = a0 - a @ lam - epsi / z
            dellam = gvec - a * z - y - b + epsi / lam
            diagx = 2.0 * (plam / ux ** 3 + qlam / xl ** 3) + xsi / (x - alfa) + eta / (beta - x)
            diagy = d + mu / y
            diaglamyi = s / lam + 1.0 / diagy
            # reduced system in (dlam, dz)
            AA = np.zeros((m + 1, m + 1))
            AA[:m, :m] = np.diag(diaglamyi) + (GG / diagx[None, :]) @ GG.T
            AA[:m, m] = a
            AA[m, :m] = a
            AA[m, m] = -zet / z
            rhs = np.concatenate([dellam + dely / diagy - GG @ (delx / diagx), [delz]])
            sol = np.linalg.solve(AA, rhs)
            dlam, dz = sol[:m], sol[m]
            dx = -delx / diagx - (GG.T @ dlam) / diagx
            dy = -dely / diagy + dlam / diagy
            dxsi = -xsi + epsi / (x - alfa) - xsi * dx / (x - alfa)
            deta = -eta + epsi / (beta - x) + eta * dx / (beta - x)
            dmu = -mu + epsi / y - mu * dy / y
            dzet = -zet + epsi / z - zet * dz / z
            ds = -s + epsi / lam - s * dlam / lam
            dv = {"x": dx, "y": dy, "z": dz, "lam": dlam, "xsi": dxsi, "eta": deta, "mu": dmu, "zet": dzet, "s": ds}
            # largest step keeping the positive variables and the bounds 1% inside
            ratios = [-1.01 * dv[q] / v[q] for q in ("y", "lam", "xsi", "eta", "mu", "s")]
            ratios += [np.array([-1.01 * dz / z, -1.01 * dzet / zet]),
                       -1.01 * dx / (x - alfa), 1.01 * dx / (beta - x)]
            steg = 1.0 / max(1.0, max(float(np.max(r)) for r in ratios))
            old = dict(v)
            new_norm = 2.0 * rnorm
            t = 0
            while new_norm > rnorm and t < 50:
                t += 1
                trials += 1
                v = {q: old[q] + steg * dv[q] for q in old}
                new_norm, new_max = _residual(v, *args, epsi)
                steg /= 2.0
            rnorm, rmax = new_norm, new_max
        epsi *= 0.1
    v["newton_steps"] = newton_steps
    v["trials"] = trials
    v["b"] = b
    return v


def kkt_residual(sol, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d, epsi=1e-7):
    """Max-norm residual of the subproblem's perturbed KKT system at sol."""
    return _residual(sol, low, upp, alfa, beta, p0, q0, P, Q, a0, a, b, c, d, epsi)[1]


def optimize_loop(rho_ini, move, max_iters, objective, constraint, H=None, Hs=None, sensitivity=True,
                  density=False):
    """optimize (mma.py:415-528) with objective(rho) -> (J, dJ (n,)), constraint(rho, it)
    -> (vc (m,), dvc (m, n)).  Returns (rho, history of J)."""
    n = rho_ini.size
    xval = rho_ini.reshape(-1).astype(np.float64).copy()
    xold1, xold2 = xval.copy(), xval.copy()
    low, upp = np.ones(n), np.ones(n)
    xmin, xmax = np.zeros(n), np.ones(n)
    hist = []
    epoch = 1
    rho = xval.copy()
    for loop in range(1, max_iters + 1):
        phys = density_filter(H, Hs, rho[:, None]).reshape(-1) if density else rho
        J, dJ = objective(phys)
        vc, dvc = constraint(phys, loop)
        vc, dvc = np.asarray(vc, dtype=np.float64).reshape(-1), np.asarray(dvc, dtype=np.float64)
        m = len(vc)
        dvc = dvc.reshape(m, n)
        if sensitivity:
            dJf, dvcf = sensitivity_filter(H, Hs, rho[:, None], np.asarray(dJ).reshape(-1, 1), dvc[:, :, None])
            dJ, dvc = dJf.reshape(-1), dvcf.reshape(m, n)
        xmma, low, upp, _ = mmasub(epoch, xval, xold1, xold2, xmin, xmax, low, upp, np.asarray(dJ).reshape(-1),
                                   vc, dvc, 1.0, np.zeros(m), 1e4 * np.ones(m), np.zeros(m), move)
        xold2, xold1, xval = xold1, xval, xmma.copy()
        epoch += 1
        rho = xval.copy()
        hist.append(float(J))
    return rho, hist
