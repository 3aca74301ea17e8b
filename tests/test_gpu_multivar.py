"""Problems with several variables through the C ABI (problem.py:55-128).

* applications/stokes/example_2vars.py written as the reference writes it: 3-D
  linear elasticity split into two variables (vec = [1, 2], a universal kernel on
  the flat cell solution, jax.flatten_util.ravel_pytree -> jax_fem_amd.flatten_util)
  on a HEX8 box; the solution equals the one-variable problem's (the built linear
  law) to 1e-10, and the FP64 oracle's to 1e-9;
* Taylor-Hood Stokes as applications/stokes/example.py writes its universal kernel
  (TRI6 velocity, TRI3 pressure, inner(grad u, grad v) + div(v) p + q div(u)), with
  a manufactured solution (stream function sin^2(pi x) sin^2(pi y), p = cos(pi x)
  cos(pi y)) and its body force as a mass map on the velocity, solved with the
  spsolve_solver option (a direct LU on the device): the velocity's L2 error falls
  with order 3 (within 0.2), the pressure's with at least the optimal order 2;
* the adjoint through ad_wrapper with two variables: a lid-driven cavity whose
  universal kernel takes a per-cell viscosity scale (a cell scalar); the gradient
  of sum u^2 + sum p^2 with respect to it against central differences.
"""
import sys

import numpy as onp
import pytest
import torch

from oracle import kernels as ok
from oracle import solver as osv

pytestmark = pytest.mark.gpu
DEV = "cuda"


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def test_two_variable_elasticity_equals_one_variable(jfa):
    say("two-variable linear elasticity (example_2vars) vs one variable")
    np = jfa.numpy

    class LinearElasticity(jfa.Problem):
        def get_universal_kernel(self):
            def stress(u_grad):
                E = 70e3
                nu = 0.3
                mu = E / (2. * (1. + nu))
                lmbda = E * nu / ((1 + nu) * (1 - 2 * nu))
                epsilon = 0.5 * (u_grad + u_grad.T)
                sigma = lmbda * np.trace(epsilon) * np.eye(self.dim) + 2 * mu * epsilon
                return sigma

            def universal_kernel(cell_sol_flat, x, cell_shape_grads, cell_JxW, cell_v_grads_JxW, *cell_internal_vars):
                cell_sol_list = self.unflatten_fn_dof(cell_sol_flat)
                cell_shape_grads = [cell_shape_grads[:, self.num_nodes_cumsum[i]: self.num_nodes_cumsum[i + 1], :]
                                    for i in range(self.num_vars)][0]
                cell_sol = np.concatenate(cell_sol_list, axis=1)
                cell_v_grads_JxW = [cell_v_grads_JxW[:, self.num_nodes_cumsum[i]: self.num_nodes_cumsum[i + 1], :, :]
                                    for i in range(self.num_vars)][0]
                vec = self.fes[0].vec + self.fes[1].vec
                u_grads = cell_sol[None, :, :, None] * cell_shape_grads[:, :, None, :]
                u_grads = np.sum(u_grads, axis=1)
                u_grads_reshape = u_grads.reshape(-1, vec, self.dim)
                u_physics = jfa.vmap(stress)(u_grads_reshape, *cell_internal_vars).reshape(u_grads.shape)
                val = np.sum(u_physics[:, None, :, :] * cell_v_grads_JxW, axis=(0, -1))
                val = [val[:, :1], val[:, 1:]]
                return jfa.flatten_util.ravel_pytree(val)[0]
            return universal_kernel

    ele_type = 'HEX8'
    meshio_mesh = jfa.box_mesh_gmsh(5, 3, 3, 1., 1., 1., None, ele_type=ele_type)
    mesh = jfa.Mesh(meshio_mesh.points, meshio_mesh.cells_dict['hexahedron'])

    def left(point):
        return np.isclose(point[0], 0., atol=1e-5)

    def right(point):
        return np.isclose(point[0], 1., atol=1e-5)

    def zero_dirichlet_val(point):
        return 0.

    def dirichlet_val(point):
        return 0.1

    dirichlet_bc_info1 = [[left, right], [0, 0], [zero_dirichlet_val, dirichlet_val]]
    dirichlet_bc_info2 = [[left, left, right, right], [0, 1, 0, 1], [zero_dirichlet_val] * 4]
    problem = LinearElasticity([mesh] * 2, vec=[1, 2], dim=3, ele_type=[ele_type] * 2, quadrature_order=[None, None],
                               dirichlet_bc_info=[dirichlet_bc_info1, dirichlet_bc_info2])
    assert problem.num_vars == 2 and problem.N == 3 * len(mesh.points)
    sol_list = jfa.solver(problem, {"newton": {"tol": 0.0, "rel_tol": 1e-10,
                                               "linear": {"bicgstab": {"tol": 1e-13, "atol": 0.0}}}})
    assert sol_list[0].shape == (len(mesh.points), 1) and sol_list[1].shape == (len(mesh.points), 2)
    u2 = torch.cat(sol_list, dim=1).cpu().numpy()

    class One(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=70e3, nu=0.3)

    one = One(mesh, vec=3, dim=3, dirichlet_bc_info=[[left, left, left, right, right, right], [0, 1, 2, 0, 1, 2],
                                                     [zero_dirichlet_val] * 3 + [dirichlet_val] + [zero_dirichlet_val] * 2])
    u1 = jfa.solver(one, {"newton": {"tol": 0.0, "rel_tol": 1e-10,
                                     "linear": {"bicgstab": {"tol": 1e-13, "atol": 0.0}}}})[0].cpu().numpy()
    say("  |u2 - u1| / |u1| = %.2e" % (onp.linalg.norm(u2 - u1) / onp.linalg.norm(u1)))
    onp.testing.assert_allclose(u2, u1, rtol=0, atol=1e-10 * onp.abs(u1).max())
    # the element matrix of the flat layout equals the one-variable one, permuted
    K2 = problem.element_jacobians(torch.zeros(problem.N, dtype=torch.float64, device=DEV)).cpu().numpy()
    K1 = one.element_jacobians(torch.zeros(one.N, dtype=torch.float64, device=DEV)).cpu().numpy()
    perm = onp.concatenate([3 * onp.arange(8), (3 * onp.arange(8)[:, None] + onp.array([1, 2])[None]).reshape(-1)])
    onp.testing.assert_allclose(K2, K1[:, perm][:, :, perm], rtol=0, atol=1e-9 * onp.abs(K1).max())
    oleft = lambda p: onp.isclose(p[0], 0., atol=1e-5)
    oright = lambda p: onp.isclose(p[0], 1., atol=1e-5)
    op = osv.OracleProblem(onp.asarray(mesh.points), onp.asarray(mesh.cells), 3,
                           ok.Material("linear_elastic", E=70e3, nu=0.3),
                           dirichlet_bc_info=[[oleft] * 3 + [oright] * 3, [0, 1, 2, 0, 1, 2],
                                              [lambda p: 0.0] * 3 + [lambda p: 0.1] + [lambda p: 0.0] * 2])
    uo, _ = osv.newton_solve(op, method="spsolve", tol=0.0, rel_tol=1e-10)
    onp.testing.assert_allclose(u2.reshape(-1), uo, rtol=0, atol=1e-9 * onp.abs(uo).max())


def _stokes(jfa, N, lid=False):
    """Taylor-Hood Stokes on the unit square, N x N cells split in two triangles.
    lid=False: the manufactured solution, body force as a mass map; lid=True: a
    lid-driven cavity (u_x = 1 on y = 1) with a per-cell viscosity scale passed as
    internal_vars (a cell scalar in the universal kernel, as dt in
    applications/thermal_mechanical_full/example.py:30-48)."""
    np = jfa.numpy
    pi = onp.pi

    def u_exact(x):
        return np.array([pi * np.sin(pi * x[0]) ** 2 * np.sin(2 * pi * x[1]),
                         -pi * np.sin(2 * pi * x[0]) * np.sin(pi * x[1]) ** 2])

    def p_exact(x):
        return np.array([np.cos(pi * x[0]) * np.cos(pi * x[1])])

    class StokesFlow(jfa.Problem):
        def custom_init(self):
            self.fe_u = self.fes[0]
            self.fe_p = self.fes[1]

        def get_universal_kernel(self):
            def universal_kernel(cell_sol_flat, x, cell_shape_grads, cell_JxW, cell_v_grads_JxW, *cell_internal_vars):
                cell_sol_list = self.unflatten_fn_dof(cell_sol_flat)
                cell_sol_u, cell_sol_p = cell_sol_list
                cell_shape_grads_list = [cell_shape_grads[:, self.num_nodes_cumsum[i]: self.num_nodes_cumsum[i + 1], :]
                                         for i in range(self.num_vars)]
                cell_shape_grads_u, cell_shape_grads_p = cell_shape_grads_list
                cell_v_grads_JxW_list = [cell_v_grads_JxW[:, self.num_nodes_cumsum[i]: self.num_nodes_cumsum[i + 1], :, :]
                                         for i in range(self.num_vars)]
                cell_v_grads_JxW_u, cell_v_grads_JxW_p = cell_v_grads_JxW_list
                cell_JxW_u, cell_JxW_p = cell_JxW[0], cell_JxW[1]
                u_grads = cell_sol_u[None, :, :, None] * cell_shape_grads_u[:, :, None, :]
                u_grads = np.sum(u_grads, axis=1)
                val1 = np.sum(u_grads[:, None, :, :] * cell_v_grads_JxW_u, axis=(0, -1))
                if cell_internal_vars:
                    val1 = cell_internal_vars[0] * val1
                p = np.sum(cell_sol_p[None, :, :] * self.fe_p.shape_vals[:, :, None], axis=1)[:, 0]
                div_v = cell_v_grads_JxW_u[:, :, 0, :]
                val2 = np.sum(p[:, None, None] * div_v, axis=0)
                div_u = u_grads[:, 0, 0] + u_grads[:, 1, 1]
                val3 = np.sum(div_u[:, None] * self.fe_p.shape_vals * cell_JxW_p[:, None], axis=0)[:, None]
                weak_form = [val1 + val2, val3]
                return jfa.flatten_util.ravel_pytree(weak_form)[0]
            return universal_kernel

        if not lid:
            def get_mass_map(self):
                # -lap u - grad p = f, with f such that (u_exact, p_exact) solves the problem
                def mass_map(u, x):
                    H = jfa.hessian(u_exact)(x)
                    lap = np.stack([H[0, 0, 0] + H[0, 1, 1], H[1, 0, 0] + H[1, 1, 1]])
                    gp = jfa.jacrev(p_exact)(x)[0]
                    return lap + gp
                return mass_map

        def set_params(self, nu):
            self.internal_vars = [nu]

    mu = jfa.rectangle_mesh_tri(N, N, 1., 1., "TRI6")
    mp = jfa.rectangle_mesh_tri(N, N, 1., 1., "TRI3")
    mesh_u, mesh_p = jfa.Mesh(mu.points, mu.cells), jfa.Mesh(mp.points, mp.cells)
    on_b = lambda p: np.isclose(p[0], 0., atol=1e-5) | np.isclose(p[0], 1., atol=1e-5) | \
        np.isclose(p[1], 0., atol=1e-5) | np.isclose(p[1], 1., atol=1e-5)
    top = lambda p: np.isclose(p[1], 1., atol=1e-5)
    corner = lambda p: np.isclose(p[0], 0., atol=1e-5) & np.isclose(p[1], 0., atol=1e-5)
    zero = lambda p: 0.
    if lid:
        info_u = [[on_b, on_b, top], [0, 1, 0], [zero, zero, lambda p: 1.0]]     # the later spec wins
        info_p = [[corner], [0], [zero]]
    else:
        info_u = [[on_b, on_b], [0, 1], [zero, zero]]
        info_p = [[corner], [0], [lambda p: 1.0]]
    problem = StokesFlow([mesh_u, mesh_p], vec=[2, 1], dim=2, ele_type=["TRI6", "TRI3"], quadrature_order=[2, 2],
                         dirichlet_bc_info=[info_u, info_p])
    return problem, u_exact, p_exact


def test_taylor_hood_stokes_convergence(jfa):
    say("Taylor-Hood Stokes, manufactured solution")
    np = jfa.numpy
    errs = []
    for N in (8, 16):
        problem, u_exact, p_exact = _stokes(jfa, N)
        sol_list = jfa.solver(problem, solver_options={'spsolve_solver': {}})
        u, p = sol_list
        assert u.shape == (len(problem.fes[0].points), 2) and p.shape == (len(problem.fes[1].points), 1)
        fu, fp = problem.fes
        uq = fu.convert_from_dof_to_quad(u)
        pq = fp.convert_from_dof_to_quad(p)
        xq = fu.get_physical_quad_points()
        ue = jfa.vmap(jfa.vmap(u_exact))(xq)
        pe = jfa.vmap(jfa.vmap(p_exact))(xq)
        eu = float(np.sqrt(np.sum((uq - ue) ** 2 * fu.JxW[:, :, None])))
        ep = float(np.sqrt(np.sum((pq - pe) ** 2 * fp.JxW[:, :, None])))
        errs.append((eu, ep))
        say("  N = %d: |u - u_e| %.4e, |p - p_e| %.4e (%d DOF)" % (N, eu, ep, problem.N))
    ru, rp = onp.log2(errs[0][0] / errs[1][0]), onp.log2(errs[0][1] / errs[1][1])
    say("  orders: velocity %.3f, pressure %.3f" % (ru, rp))
    # P2 velocity: order 3; P1 pressure: at least the optimal order 2 (2.77 here: the
    # pressure of Taylor-Hood pairs superconverges on uniform meshes at these sizes)
    assert abs(ru - 3) < 0.2 and rp > 1.8


def test_two_variable_adjoint_vs_fd(jfa):
    say("two-variable adjoint (lid-driven cavity, per-cell viscosity) vs central differences")
    problem, _, _ = _stokes(jfa, 4, lid=True)
    fwd = jfa.ad_wrapper(problem, solver_options={'spsolve_solver': {}}, adjoint_solver_options={'spsolve_solver': {}})
    rng = onp.random.default_rng(0)
    nu0 = torch.tensor(rng.uniform(0.8, 1.2, problem.num_cells), dtype=torch.float64, device=DEV)

    def J(nu):
        u, p = fwd(nu)
        return torch.sum(u ** 2) + torch.sum(p ** 2)
    nu = nu0.clone().requires_grad_(True)
    J(nu).backward()
    g = nu.grad.cpu().numpy()
    h = 1e-6
    for c in rng.choice(problem.num_cells, 4, replace=False):
        e = torch.zeros_like(nu0)
        e[c] = h
        fd = (float(J(nu0 + e)) - float(J(nu0 - e))) / (2 * h)
        say("  cell %d: fd %.8e adjoint %.8e" % (c, fd, g[c]))
        assert abs(fd - g[c]) <= 1e-5 * max(abs(g[c]), 1e-8)
