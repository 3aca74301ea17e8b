"""Train the learned preconditioner on the GPU with the library's own kernels.

    python jax-fem_amd/train_lpc.py --n 32 --instances 2000 --epochs 1 \
        --out jax-fem_amd/weights/lpc_elastic.npz

Training set: heterogeneous linear elasticity on a 32^3 HEX8 box, Young's
modulus a log-normal Gaussian random field per cell (mean log 0, sigma 0.5,
correlation length 0.1, clipped to [0.05, 20]), nu = 0.3, u = 0 on x = 0;
instance i uses seed i (1..instances).  --domains d1,d2,... samples the field
on the box [0, d]^3 of instance i (d cycling over the list), i.e. at grid spacing
d/32: d = 1 is the unit cube, d = 0.25 the spacing of a 128-cell-wide mesh of
the unit cube (correlation length 12.8 cells instead of 3.2).  Each step takes one instance's FP32
tangent, draws `batch` Jacobi-PCG residuals (iterations 0..k_max) of a random
right-hand side and takes an Adam step on the energy loss (learned_pc.train).
Held-out seeds (9001, ...) measure iterations of learned FCG vs Jacobi-PCG.
--element HEX27 trains on the quadratic node lattice (n^3 cells, (2n+1)^3 nodes,
the SELL tangent as the multigrid's fine level) with --sampler mg.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, the package directory is sys.path[0]: its numpy.py (the
# jax.numpy stand-in) would shadow NumPy itself
if sys.path and os.path.abspath(sys.path[0] or ".") == HERE:
    sys.path.pop(0)


def load_pkg():
    if "jax_fem_amd" in sys.modules:
        return sys.modules["jax_fem_amd"]
    spec = importlib.util.spec_from_file_location("jax_fem_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["jax_fem_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--instances", type=int, default=2000)
    ap.add_argument("--epochs", type=float, default=1.0)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--kmax", type=int, default=150)
    ap.add_argument("--eval", type=int, default=3)
    ap.add_argument("--init", default=None, help="start from these weights")
    ap.add_argument("--domains", default="1.0", help="comma-separated field box sizes, cycled over instances")
    ap.add_argument("--self-frac", type=float, default=0.5)
    ap.add_argument("--rhs", default="noise", help="comma-separated right-hand-side kinds: noise, face, smooth")
    ap.add_argument("--sampler", default="krylov", choices=("krylov", "mg"),
                    help="krylov: Jacobi-PCG / own-FCG residuals (the preconditioner); mg: residuals after a "
                         "multigrid coarse correction (a fine-level post-smoother)")
    ap.add_argument("--element", default="HEX8", choices=("HEX8", "HEX27"))
    ap.add_argument("--cache", type=int, default=256, help="tangent matrices kept on the device")
    ap.add_argument("--out", default=os.path.join(HERE, "weights", "lpc_elastic.npz"))
    args = ap.parse_args()

    import numpy as np
    import torch
    jfa = load_pkg()
    from jax_fem_amd import learned_pc as L
    from jax_fem_amd.solver import TangentMatrix, krylov

    torch.cuda.init()
    n = args.n
    domains = [float(x) for x in args.domains.split(",")]

    def field(seed):
        d = domains[(seed - 1) % len(domains)]
        return L.heterogeneous_modulus((n, n, n), seed, domain=(d, d, d))

    E0 = field(1)
    prob = L.elasticity_instance(jfa, n, E0, ele_type=args.element)
    u0 = torch.zeros(prob.N, dtype=torch.float64, device="cuda")
    cache = {}

    def matrix(i):
        seed = 1 + (i % args.instances)
        if seed not in cache:
            if len(cache) >= max(args.cache, 1):
                cache.pop(next(iter(cache)))
            E = field(seed)
            L.set_modulus(prob, E)
            cache[seed] = (TangentMatrix(prob, prob.jacobian_bsr(u0, torch.float32)), E)
        A, E = cache[seed]
        L.set_modulus(prob, E)
        return A, prob, seed

    model = L.LearnedPC.load(args.init) if args.init else L.LearnedPC(seed=0)
    steps = int(args.epochs * args.instances)
    sampler = None
    if args.sampler == "mg":
        from jax_fem_amd.multigrid import Multigrid, dirichlet_mask
        mg = Multigrid(L.LearnedPC.grid(prob), fine_sell=args.element == "HEX27", fine_f16=args.element == "HEX27")

        def sampler(model, A, p, ks, sd, g, kind):
            if not mg.is_current(A):
                mg.setup(A, dirichlet_mask(p))
            return L.mg_smoother_residuals(mg, A, p, ks, sd, kind)
    t0 = time.perf_counter()
    hist = L.train(model, matrix, steps, batch=args.batch, lr=args.lr, k_max=args.kmax,
                   log=lambda s: print("[train]", s, flush=True), log_every=max(1, steps // 40),
                   self_frac=args.self_frac, rhs_kinds=tuple(args.rhs.split(",")), sampler=sampler)
    t_train = time.perf_counter() - t0
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    model.save(args.out)

    # held-out evaluation
    res = []
    for k in range(args.eval * len(domains)):
        seed = 9001 + k
        E = field(seed)
        L.set_modulus(prob, E)
        A = TangentMatrix(prob, prob.jacobian_bsr(u0, torch.float32))
        b = torch.randn(prob.N, device="cuda", generator=torch.Generator(device="cuda").manual_seed(seed))
        b = b * L.LearnedPC.mask(prob)
        _, ij = krylov(A, b, None, "pcg", 1e-6, 0.0, 20000)
        _, il, fb = model.solve(A, b, None, prob, rtol=1e-6, maxiter=20000)
        res.append({"seed": seed, "domain": domains[(seed - 1) % len(domains)], "jacobi_pcg": ij.iterations,
                    "learned_fcg": il.iterations, "fallback": fb, "fcg_part": model.last_fcg})
        print("[eval]", res[-1], flush=True)
    print(json.dumps({"train_seconds": t_train, "steps": steps, "final_loss": hist[-1], "eval": res}))


if __name__ == "__main__":
    main()
