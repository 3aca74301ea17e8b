"""User maps written the reference's way (plain functions, jax.numpy -> jax_fem_amd.numpy,
jax.grad -> jax_fem_amd.grad) through the C ABI.

* the hyperelasticity fixture's problem defined as
  tests/benchmarks/hyperelasticity/test_hyper_elasticity.py:15-34 does (an energy,
  its grad, plain NumPy), the mesh re-wrapped as Mesh(m.points, m.cells_dict[...]):
  recognised as the built Neo-Hookean law, FEniCSx solution at decimal 5 and the
  top traction 8014.825068329136 at decimal 5 (post-processing written as the
  reference's compute_traction, with jax_fem_amd.vmap);
* a law evaluated pointwise (probing off): cell residuals and element Jacobians
  against the FP64 oracle at 1e-11, on a distorted mesh, Neo-Hookean and linear
  elasticity with a per-point modulus; Newton solutions against the built path;
  parameter gradients against the built path's hf_cell_param_vjp;
* J2 plasticity with two internal variables per point (the reference's
  test_plasticity.py:16-161 problem): FEniCSx solution at decimal 5 and the
  average sigma_zz of the 5 load steps at decimal 3;
* a u-dependent mass map (reaction k u) and a u-dependent surface map (Robin
  h u): for u = c the residual sums to c (k |volume| + h |face|) and the Jacobian
  is symmetric with entries summing to k |volume| + h |face| (integrals of the
  partition of unity);
* a universal kernel written like the reference's laplace kernel equals the
  built linear-elastic kernels (residual, Jacobian).
Tolerances: FP64 1e-11 relative to the oracle / built kernels (summation order);
solutions 1e-9.
"""
import os
import sys

import numpy as onp
import pytest
import torch

from conftest import GOLDEN
from oracle import fe as ofe
from oracle import kernels as ok
from oracle import solver as osv

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def fixture(name):
    return onp.load(os.path.join(GOLDEN, "fenicsx_%s.npz" % name), allow_pickle=False)


def test_hyperelasticity_fixture_reference_style(jfa, tmp_path):
    say("hyperelasticity fixture, reference-style problem")
    import numpy as np          # plain NumPy in place of jax.numpy

    class HyperElasticity(jfa.Problem):
        def get_tensor_map(self):
            def psi(F):
                E = 1e3
                nu = 0.3
                mu = E / (2. * (1. + nu))
                kappa = E / (3. * (1. - 2. * nu))
                J = np.linalg.det(F)
                Jinv = J ** (-2. / 3.)
                I1 = np.trace(F.T @ F)
                energy = (mu / 2.) * (Jinv * I1 - 3.) + (kappa / 2.) * (J - 1.) ** 2.
                return energy
            P_fn = jfa.grad(psi)

            def first_PK_stress(u_grad):
                I = np.eye(self.dim)
                F = u_grad + I
                P = P_fn(F)
                return P
            return first_PK_stress

        def compute_traction(self, location_fn, sol):
            stress = self.get_tensor_map()
            fe = self.fes[0]
            bi = fe.get_boundary_conditions_inds([location_fn])[0]
            grads, nanson = fe.get_face_shape_grads(bi)
            cs = np.asarray(sol)[fe.cells][bi[:, 0]]
            u_grads = np.sum(cs[:, None, :, :, None] * grads[:, :, :, None, :], axis=2)
            sigmas = jfa.vmap(stress)(u_grads.reshape(-1, 3, 3)).reshape(u_grads.shape)
            normals = np.array([0., 0., 1.]).reshape((3, 1))
            traction = (sigmas @ normals[None, None, :, :])[:, :, :, 0]
            return np.sum(traction * nanson[:, :, None], axis=(0, 1))

    d = fixture("hyperelasticity")
    vtu = str(tmp_path / "fenicsx" / "sol.vtu")      # the FEniCSx gold as a .vtu, read back
    jfa.write_vtu(vtu, d["points"], d["cells"], "hexahedron", {"sol": d["sol"]})
    fenicsx_vtu = jfa.read_vtu(vtu)
    mesh = jfa.Mesh(fenicsx_vtu.points, fenicsx_vtu.cells_dict[jfa.get_meshio_cell_type("HEX8")])
    top = lambda point: np.isclose(point[2], 10., atol=1e-5)
    bottom = lambda point: np.isclose(point[2], 0., atol=1e-5)
    zero = lambda point: 0.
    one = lambda point: 1.
    info = [[bottom, bottom, bottom, top, top, top], [0, 1, 2, 0, 1, 2], [zero] * 5 + [one]]
    problem = HyperElasticity(mesh, vec=3, dim=3, dirichlet_bc_info=info)
    sol = jfa.solver(problem)[0].cpu().numpy()
    assert problem._plan()["law"] == "fast" and type(problem.material).__name__ == "NeoHookean"
    onp.testing.assert_array_almost_equal(fenicsx_vtu.point_data["sol"].reshape(sol.shape), sol, decimal=5)
    tz = problem.compute_traction(top, sol)[2]
    say("  traction %.9f (FEniCSx %.9f)" % (tz, float(d["traction"])))
    onp.testing.assert_almost_equal(float(d["traction"]), tz, decimal=5)


def _distorted_box(n=(3, 2, 2), seed=0):
    pts, cells = ofe.box_mesh(*n)
    rng = onp.random.default_rng(seed)
    interior = onp.all((pts > 1e-9) & (pts < 1 - 1e-9), axis=1)
    pts = pts.copy()
    pts[interior] += 0.08 * rng.uniform(-1, 1, (interior.sum(), 3)) / max(n)
    return pts, cells


def _neo_map(jfa, E, nu):
    np = jfa.numpy
    mu, kappa = E / (2 * (1 + nu)), E / (3 * (1 - 2 * nu))

    def psi(F, rho):
        J = np.linalg.det(F)
        return rho * (mu / 2 * (J ** (-2. / 3.) * np.trace(F.T @ F) - 3.) + kappa / 2 * (J - 1.) ** 2)
    P = jfa.grad(psi)
    return lambda u_grad, rho: P(u_grad + np.eye(3), rho)


def _lin_map(jfa, E, nu):
    np = jfa.numpy
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))

    def stress(u_grad, rho):
        eps = 0.5 * (u_grad + u_grad.T)
        return rho * (lam * np.trace(eps) * np.eye(3) + 2 * mu * eps)
    return stress


@pytest.mark.parametrize("kind", ["neo_hookean", "linear_elastic"])
def test_pointwise_law_vs_oracle_and_built(jfa, kind):
    say("pointwise law vs oracle", kind)
    pts, cells = _distorted_box()
    fn = (_neo_map if kind == "neo_hookean" else _lin_map)(jfa, 10.0, 0.3)
    built = (jfa.NeoHookean if kind == "neo_hookean" else jfa.LinearElastic)(E=10.0, nu=0.3, param="scale_E")
    left = lambda p: onp.isclose(p[0], 0.0, atol=1e-5)
    right = lambda p: onp.isclose(p[0], 1.0, atol=1e-5)
    info = [[left] * 3 + [right], [0, 1, 2, 0], [lambda p: 0.0] * 3 + [lambda p: 0.05]]

    class Pointwise(jfa.Problem):
        probe_laws = False

        def get_tensor_map(self):
            return fn

    class Built(jfa.Problem):
        def get_tensor_map(self):
            return built

    rng = onp.random.default_rng(2)
    rho = rng.uniform(0.5, 1.5, (len(cells), 8))
    pg = Pointwise(jfa.Mesh(pts, cells), vec=3, dim=3, dirichlet_bc_info=info)
    pb = Built(jfa.Mesh(pts, cells), vec=3, dim=3, dirichlet_bc_info=info)
    pg.internal_vars = [rho]
    pb.internal_vars = [rho]
    assert pg._plan()["law"] == "generic"
    op = osv.OracleProblem(pts, cells, 3, ok.Material(kind, E=10.0, nu=0.3, param="scale_E"), params=rho)
    u = 0.05 * rng.standard_normal(op.N)
    ut = torch.tensor(u, device=DEV)
    R = pg.residual_vec(ut).cpu().numpy()
    Ro = op.compute_residual(u)
    onp.testing.assert_allclose(R, Ro, rtol=1e-11, atol=1e-11 * onp.abs(Ro).max())
    K = pg.element_jacobians(ut).cpu().numpy()
    Ko = op.element_jacobians(u)
    onp.testing.assert_allclose(K, Ko, rtol=1e-11, atol=1e-11 * onp.abs(Ko).max())
    K32 = pg.element_jacobians(ut, torch.float32).cpu().numpy()
    onp.testing.assert_allclose(K32, Ko, rtol=2e-5, atol=2e-5 * onp.abs(Ko).max())
    # Newton: the pointwise path against the built one (FP64 and mixed)
    for prec in ("fp64", "mixed"):
        o = {"newton": {"precision": prec, "tol": 0.0, "rel_tol": 1e-10,
                        "linear": {"pcg": {"tol": 1e-12, "atol": 0.0}} if prec == "fp64" else {"pcg": {}}}}
        ug = jfa.solver(pg, o)[0]
        ub = jfa.solver(pb, o)[0]
        rel = (torch.linalg.norm(ug - ub) / torch.linalg.norm(ub)).item()
        say("  newton %s: rel %.1e, iterations %d" % (prec, rel, pg.solve_info["newton_iters"]))
        assert rel <= 1e-9
    # parameter gradient of <lam, R> through the pointwise VJP vs the built kernel's
    S = __import__("jax_fem_amd.solver", fromlist=["x"])
    lam = torch.tensor(rng.standard_normal(op.N), device=DEV)
    d, _ = pb.bc_arrays()
    lam[d.long()] = 0.0
    gb = S.param_vjp(pb, ut, lam)
    gg = S.param_vjp_pointwise(pg, ut, lam, [0])[0]
    torch.testing.assert_close(gg, gb, rtol=1e-10, atol=1e-12 * gb.abs().max().item())


class _Plasticity:
    """test_plasticity.py:16-97 with jax_fem_amd in place of jax."""

    @staticmethod
    def build(jfa, mesh, info):
        np = jfa.numpy

        class Plasticity(jfa.Problem):
            def custom_init(self):
                self.fe = self.fes[0]
                self.epsilons_old = np.zeros((len(self.fe.cells), self.fe.num_quads, self.fe.vec, self.dim))
                self.sigmas_old = np.zeros_like(self.epsilons_old)
                self.internal_vars = [self.sigmas_old, self.epsilons_old]

            def get_tensor_map(self):
                _, stress_return_map = self.get_maps()
                return stress_return_map

            def get_maps(self):
                def safe_sqrt(x):
                    safe_x = np.where(x > 0., np.sqrt(x), 0.)
                    return safe_x

                def safe_divide(x, y):
                    return np.where(y == 0., 0., x / y)

                def strain(u_grad):
                    epsilon = 0.5 * (u_grad + u_grad.T)
                    return epsilon

                def stress(epsilon):
                    E = 70.e3
                    nu = 0.3
                    mu = E / (2. * (1. + nu))
                    lmbda = E * nu / ((1 + nu) * (1 - 2 * nu))
                    sigma = lmbda * np.trace(epsilon) * np.eye(self.dim) + 2 * mu * epsilon
                    return sigma

                def stress_return_map(u_grad, sigma_old, epsilon_old):
                    sig0 = 250.
                    epsilon_crt = strain(u_grad)
                    epsilon_inc = epsilon_crt - epsilon_old
                    sigma_trial = stress(epsilon_inc) + sigma_old
                    s_dev = sigma_trial - 1. / self.dim * np.trace(sigma_trial) * np.eye(self.dim)
                    s_norm = safe_sqrt(3. / 2. * np.sum(s_dev * s_dev))
                    f_yield = s_norm - sig0
                    f_yield_plus = np.where(f_yield > 0., f_yield, 0.)
                    sigma = sigma_trial - safe_divide(f_yield_plus * s_dev, s_norm)
                    return sigma
                return strain, stress_return_map

            def stress_strain_fns(self):
                strain, stress_return_map = self.get_maps()
                vmap_strain = jfa.vmap(jfa.vmap(strain))
                vmap_stress_return_map = jfa.vmap(jfa.vmap(stress_return_map))
                return vmap_strain, vmap_stress_return_map

            def update_stress_strain(self, sol):
                u_grads = self.fe.sol_to_grad(sol)
                vmap_strain, vmap_stress_rm = self.stress_strain_fns()
                self.sigmas_old = vmap_stress_rm(u_grads, self.sigmas_old, self.epsilons_old)
                self.epsilons_old = vmap_strain(u_grads)
                self.internal_vars = [self.sigmas_old, self.epsilons_old]

            def compute_avg_stress(self):
                JxW = torch.as_tensor(self.fe.JxW, device=self.sigmas_old.device)
                sigma = np.sum(self.sigmas_old.reshape(-1, self.fe.vec, self.dim) * JxW.reshape(-1)[:, None, None], 0)
                vol = np.sum(JxW)
                avg_sigma = sigma / vol
                return avg_sigma

        return Plasticity(mesh, vec=3, dim=3, dirichlet_bc_info=info)


def test_plasticity_fixture(jfa):
    say("J2 plasticity fixture")
    d = fixture("plasticity")
    mesh = jfa.Mesh(d["points"], d["cells"])
    H = 10.
    top = lambda point: onp.isclose(point[2], H, atol=1e-5)
    bottom = lambda point: onp.isclose(point[2], 0., atol=1e-5)
    zero = lambda point: 0.

    def get_dirichlet_top(disp):
        return lambda point: disp

    disps = d["disps"]
    info = [[bottom, bottom, bottom, top, top, top], [0, 1, 2, 0, 1, 2], [zero] * 5 + [get_dirichlet_top(disps[0])]]
    problem = _Plasticity.build(jfa, mesh, info)
    avg = []
    sol = None
    for i, disp in enumerate(disps):
        info[-1][-1] = get_dirichlet_top(disp)
        problem.fe.update_Dirichlet_boundary_conditions(info)
        sol = jfa.solver(problem)[0]
        problem.update_stress_strain(sol)
        avg.append(float(problem.compute_avg_stress()[2, 2]))
        say("  step %d disp %.3f: newton %d, avg sigma_zz %.6f" % (i, disp, problem.solve_info["newton_iters"],
                                                                   avg[-1]))
    assert problem._plan()["law"] == "generic"
    onp.testing.assert_array_almost_equal(d["sol"], sol.cpu().numpy(), decimal=5)
    onp.testing.assert_array_almost_equal(d["avg_stresses"], onp.array(avg), decimal=3)


def test_u_dependent_mass_and_surface_terms(jfa):
    say("u-dependent mass (reaction) and surface (Robin) maps")
    k_r, h_r = 3.0, 0.7
    pts, cells = _distorted_box((3, 3, 2), seed=4)
    right = lambda p: onp.isclose(p[0], 1.0, atol=1e-5)

    class Robin(jfa.Problem):
        def get_tensor_map(self):
            return jfa.Poisson()

        def get_mass_map(self):
            return lambda u, x: k_r * u

        def get_surface_maps(self):
            return [lambda u, x: h_r * u]

    prob = Robin(jfa.Mesh(pts, cells), vec=1, dim=3, location_fns=[right])
    plan = prob._plan()
    assert plan["mass"] == "cell" and plan["surfaces"] == ["face"] and plan["law"] == "fast"
    vol = float(prob.fe.JxW.sum())
    bi = prob.boundary_inds_list[0]
    area = float(prob.fe.get_face_shape_grads(bi)[1].sum())
    c = 1.7
    R = prob.residual_vec(torch.full((prob.N,), c, dtype=torch.float64, device=DEV))
    want = c * (k_r * vol + h_r * area)
    assert abs(R.sum().item() / want - 1) < 1e-12, (R.sum().item(), want)
    S = __import__("jax_fem_amd.solver", fromlist=["x"])
    A = S.TangentMatrix(prob, prob.jacobian_bsr(torch.zeros(prob.N, dtype=torch.float64, device=DEV), bc=False))
    Ah = A.to_scipy()
    assert abs(Ah.sum() / (k_r * vol + h_r * area) - 1) < 1e-12
    assert abs(Ah - Ah.T).max() <= 1e-13 * abs(Ah).max()
    # a solve with a source: the nonlinear solver converges to the linear system's solution
    prob2 = Robin(jfa.Mesh(pts, cells), vec=1, dim=3, location_fns=[right],
                  dirichlet_bc_info=[[lambda p: onp.isclose(p[0], 0.0, atol=1e-5)], [0], [lambda p: 1.0]])
    u = jfa.solver(prob2)[0].reshape(-1)
    Rs = S.apply_bc_vec(prob2.residual_vec(u), u, prob2)
    assert torch.linalg.norm(Rs).item() <= 1e-9


def test_universal_kernel_equals_built_laplace(jfa):
    say("universal kernel vs built linear elasticity")
    np = jfa.numpy
    E, nu = 10.0, 0.3
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    pts, cells = _distorted_box(seed=5)

    class Universal(jfa.Problem):
        def get_universal_kernel(self):
            def kernel(cell_sol_flat, x, cell_shape_grads, cell_JxW, cell_v_grads_JxW, *iv):
                cell_sol = cell_sol_flat.reshape(-1, 3)
                u_grads = np.sum(cell_sol[None, :, :, None] * cell_shape_grads[:, :, None, :], axis=1)
                eps = 0.5 * (u_grads + np.transpose(u_grads, (0, 2, 1)))
                tr = eps[:, 0, 0] + eps[:, 1, 1] + eps[:, 2, 2]
                sigma = lam * tr[:, None, None] * np.eye(3)[None] + 2 * mu * eps
                val = np.sum(sigma[:, None, :, :] * cell_v_grads_JxW, axis=(0, -1))
                return val.reshape(-1)
            return kernel

    class Built(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=E, nu=nu)

    pu = Universal(jfa.Mesh(pts, cells), vec=3, dim=3)
    pb = Built(jfa.Mesh(pts, cells), vec=3, dim=3)
    u = torch.tensor(onp.random.default_rng(3).standard_normal(pu.N) * 0.05, device=DEV)
    Ru, Rb = pu.residual_vec(u), pb.residual_vec(u)
    torch.testing.assert_close(Ru, Rb, rtol=1e-11, atol=1e-11 * Rb.abs().max().item())
    Ku, Kb = pu.element_jacobians(u), pb.element_jacobians(u)
    torch.testing.assert_close(Ku, Kb, rtol=1e-11, atol=1e-11 * Kb.abs().max().item())
