"""The learned network as the multigrid's fine-level post-smoother (hf_mg_set_learned).

* one V-cycle with the learned post-smoother equals the cycle stopped after the
  fine coarse correction plus the network applied once (hf_lpc_apply_f32, the
  inference path tested against the FP32 and FP64 network in test_gpu_learned_pc)
  to the coarse-corrected residual: the step is what it says it is;
* the device input field equals the host one (LearnedPC.gfield);
* held out (a log-normal field the network was not trained on, 64x64x32 cells):
  the trained smoother needs fewer flexible-CG iterations and less wall time than
  the same cycle with the untrained network (last layer zero: a plain Jacobi step,
  the method without it), and the whole Newton solve beats the plain FP64
  Jacobi-PCG drop-in by at least 3x in wall time, to the same FP64 residual, with
  no fallback; its solution agrees with the plain one to 1e-6;
* the same two checks on HEX27 node lattices (the SELL tangent as the fine level,
  weights trained on 16^3-cell HEX27 linear elasticity, held out on Neo-Hookean).
"""
import os
import sys
import time

import numpy as np
import pytest
import torch

from conftest import ROOT

pytestmark = pytest.mark.gpu
DEV = "cuda"
WEIGHTS = os.path.join(ROOT, "jax-fem_amd", "weights", "lpc_mg_smoother.npz")


def say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


def hetero(jfa, cells, seed):
    lpc = __import__("jax_fem_amd.learned_pc", fromlist=["x"])
    nx, ny, nz = cells
    E = lpc.heterogeneous_modulus(cells, seed)

    class Hetero(jfa.Problem):
        def get_tensor_map(self):
            return jfa.LinearElastic(E=1.0, nu=0.3, param="scale_E")

        def get_surface_maps(self):
            return [lambda u, x: np.array([0.0, 0.0, 10.0])]

    left = lambda p: np.isclose(p[0], 0.0, atol=1e-7)
    right = lambda p: np.isclose(p[0], 1.0, atol=1e-7)
    prob = Hetero(jfa.box_mesh(nx, ny, nz, 1.0, ny / nx, nz / nx), vec=3, dim=3,
                  dirichlet_bc_info=[[left] * 3, [0, 1, 2], [lambda p: 0.0] * 3], location_fns=[right])
    prob.internal_vars = [np.repeat(E[:, None], 8, axis=1)]
    return prob


def test_learned_step_is_network_on_coarse_corrected_residual(jfa):
    say("learned mg smoother: one cycle")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    S = __import__("jax_fem_amd.solver", fromlist=["x"])
    lpc = __import__("jax_fem_amd.learned_pc", fromlist=["x"])
    model = lpc.LearnedPC.load(WEIGHTS)
    prob = hetero(jfa, (12, 10, 9), seed=77)
    prob.hold_params(True)
    gd = lpc.LearnedPC.gfield_device(prob).double()
    gh = lpc.LearnedPC.gfield(prob).double()
    assert torch.allclose(gd, gh, rtol=0, atol=1e-6), float((gd - gh).abs().max())
    A = S.TangentMatrix(prob, prob.jacobian_bsr(torch.zeros(prob.N, dtype=torch.float64, device=DEV),
                                                torch.float32))
    mask = MG.dirichlet_mask(prob)
    mg = MG.Multigrid(MG.lattice(prob))
    mg.setup(A, mask)
    g = torch.Generator(device=DEV).manual_seed(3)
    r = torch.randn(prob.N, generator=g, device=DEV) * mask
    e = mg.apply(r, skip_fine_post=True)                 # pre-smoothing + coarse correction
    res = (r - A.matvec(e)) * mask
    ref = e + model.apply(A, res, prob)
    mg.set_learned(model, prob, "replace")
    z = mg.apply(r)
    prob.hold_params(False)
    err = (z - ref).abs().max().item()
    say("  max |cycle - (e + M(res))| = %.2e of %.2e" % (err, ref.abs().max().item()))
    assert err <= 1e-4 * ref.abs().max().item()
    mg.set_learned(None)
    z0 = mg.apply(r)                                     # off again: the Chebyshev post-smoothing
    assert (z0 - z).abs().max().item() > 1e-3 * ref.abs().max().item()


def _timed(jfa, prob, opts, reps=2):
    jfa.solver(prob, opts)                               # first call: set-up, capture
    best = None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u = jfa.solver(prob, opts)[0]
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    return best, u.reshape(-1), dict(prob.solve_info)


def test_trained_smoother_gain_held_out(jfa):
    say("learned mg smoother: held-out gain")
    lpc = __import__("jax_fem_amd.learned_pc", fromlist=["x"])
    trained = lpc.LearnedPC.load(WEIGHTS)
    untrained = lpc.LearnedPC(zero_last=True)
    prob = hetero(jfa, (64, 64, 32), seed=9001)

    def mg(model):
        return {"newton": {"precision": "mixed", "tol": 0.0, "eta_first": 1e-4,
                           "linear": {"mg": {"smoother": "learned", "model": model}}}}
    t_l, u_l, i_l = _timed(jfa, prob, mg(trained))
    t_z, _, i_z = _timed(jfa, prob, mg(untrained))
    plain = {"newton": {"precision": "fp64", "tol": 0.0, "forcing": "fixed",
                        "linear": {"pcg": {"tol": 1e-10, "atol": 1e-10}}}}
    t_p, u_p, i_p = _timed(jfa, prob, plain, reps=1)
    rel = (torch.linalg.norm(u_l - u_p) / torch.linalg.norm(u_p)).item()
    say("  trained: %.4f s, %d its | untrained: %.4f s, %d its | plain: %.4f s, %d its | rel-L2 %.1e"
        % (t_l, i_l["krylov_iters"], t_z, i_z["krylov_iters"], t_p, i_p["krylov_iters"], rel))
    assert i_l["fallbacks"] == 0 and i_l["rel"] <= 1e-8
    assert i_l["krylov_iters"] < i_z["krylov_iters"]
    assert t_l < t_z
    assert t_p / t_l >= 3.0
    assert rel <= 1e-6


# --- HEX27 lattices: the fine level is the SELL tangent -----------------------------------
WEIGHTS27 = os.path.join(ROOT, "jax-fem_amd", "weights", "lpc_mg_smoother_hex27.npz")


def test_learned_step_hex27_sell_fine_level(jfa):
    """On a HEX27 node lattice the learned step reads the residual r - A e formed with
    the SELL tangent's product: the cycle equals the coarse-corrected iterate plus the
    network (hf_lpc_apply_f32) on that residual."""
    say("learned mg smoother: one cycle, HEX27")
    MG = __import__("jax_fem_amd.multigrid", fromlist=["x"])
    S = __import__("jax_fem_amd.solver", fromlist=["x"])
    lpc = __import__("jax_fem_amd.learned_pc", fromlist=["x"])
    model = lpc.LearnedPC.load(WEIGHTS27)
    prob = lpc.elasticity_instance(jfa, (6, 5, 4), lpc.heterogeneous_modulus((6, 5, 4), 77), ele_type="HEX27")
    assert lpc.LearnedPC.grid(prob) == (13, 11, 9)
    prob.hold_params(True)
    A = S.TangentMatrix(prob, prob.jacobian_bsr(torch.zeros(prob.N, dtype=torch.float64, device=DEV),
                                                torch.float32))
    mask = MG.dirichlet_mask(prob)
    mg = MG.Multigrid(MG.lattice(prob), fine_sell=True, fine_f16=False)   # FP32 products, as A.matvec
    mg.setup(A, mask)
    g = torch.Generator(device=DEV).manual_seed(3)
    r = torch.randn(prob.N, generator=g, device=DEV) * mask
    e = mg.apply(r, skip_fine_post=True)
    res = (r - A.matvec(e)) * mask
    ref = e + model.apply(A, res, prob)
    mg.set_learned(model, prob, "replace")
    z = mg.apply(r)
    prob.hold_params(False)
    err = (z - ref).abs().max().item()
    say("  max |cycle - (e + M(res))| = %.2e of %.2e" % (err, ref.abs().max().item()))
    assert err <= 1e-4 * ref.abs().max().item()
    mg.set_learned(None)
    z0 = mg.apply(r)
    assert (z0 - z).abs().max().item() > 1e-3 * ref.abs().max().item()


def test_trained_smoother_gain_held_out_hex27(jfa):
    """Held out: Neo-Hookean HEX27 24^3 cells (the network was trained on linear
    elasticity with log-normal moduli on 16^3 cells), u_x = 0.03 in 3 load steps:
    the trained smoother needs fewer flexible-CG iterations and less wall time than
    the untrained one (a plain Jacobi step), beats the plain FP64 Jacobi-PCG drop-in
    by at least 3x with no fallback, and agrees with its solution to 1e-6."""
    say("learned mg smoother: held-out gain, HEX27")
    sys.path.insert(0, ROOT)
    import bench
    lpc = __import__("jax_fem_amd.learned_pc", fromlist=["x"])
    trained = lpc.LearnedPC.load(WEIGHTS27)
    untrained = lpc.LearnedPC(zero_last=True)
    prob, bc = bench.build_hex27(jfa, 24)

    def run(opts, reps):
        bench.run_sequence(jfa, prob, bc, opts, nload=3, dmax=0.03)     # set-up and capture
        best, sol, infos = None, None, None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sol, infos = bench.run_sequence(jfa, prob, bc, opts, nload=3, dmax=0.03)
            t = time.perf_counter() - t0
            best = t if best is None else min(best, t)
        return best, sol[0].reshape(-1), infos

    def mg(model):
        return {"newton": {"precision": "mixed", "tol": 0.0, "eta_first": 1e-4,
                           "linear": {"mg": {"smoother": "learned", "model": model}}}}
    t_l, u_l, i_l = run(mg(trained), 2)
    t_z, _, i_z = run(mg(untrained), 2)
    plain = {"newton": {"precision": "fp64", "tol": 0.0, "forcing": "fixed",
                        "linear": {"pcg": {"tol": 1e-10, "atol": 1e-10}}}}
    t_p, u_p, i_p = run(plain, 1)
    k_l = sum(i["krylov_iters"] for i in i_l)
    k_z = sum(i["krylov_iters"] for i in i_z)
    rel = (torch.linalg.norm(u_l - u_p) / torch.linalg.norm(u_p)).item()
    say("  trained: %.4f s, %d its | untrained: %.4f s, %d its | plain: %.4f s, %d its | rel-L2 %.1e"
        % (t_l, k_l, t_z, k_z, t_p, sum(i["krylov_iters"] for i in i_p), rel))
    assert all(i["fallbacks"] == 0 and i["rel"] <= 1e-8 for i in i_l)
    assert k_l < k_z
    assert t_l < t_z
    assert t_p / t_l >= 3.0
    assert rel <= 1e-6
