"""jax_fem_amd: JAX-FEM's implicit solve path on one AMD MI355X (gfx950).

Import name ``jax_fem_amd`` (the directory is ``jax-fem_amd``).  Public API
mirrors the reference package: ``Problem`` (jax_fem/problem.py),
``solver`` / ``ad_wrapper`` / ``implicit_vjp`` / ``get_A`` / ``linear_solver``
(jax_fem/solver.py), ``Mesh`` / ``box_mesh`` (jax_fem/generate_mesh.py) and
``FiniteElement`` (jax_fem/fe.py); ``jax_fem_amd.numpy`` with ``grad`` / ``vmap``
/ ``jacfwd`` take the place of jax.numpy and jax's transforms in user maps.  Compute runs in libhipfem.so through a C
ABI (include/hipfem.h); there is no CPU fallback.
"""
import logging

import torch  # noqa: F401  (load torch's HIP runtime before libhipfem)

from . import _lib
from ._lib import HipFemError
from .fe import FiniteElement
from .materials import LinearElastic, NeoHookean, Poisson, Material
from .mesh import (Mesh, box_mesh, box_mesh_hex27, box_mesh_tet, box_mesh_gmsh, rectangle_mesh, rectangle_mesh_tri,
                   get_meshio_cell_type)
from .problem import Problem
from .utils import save_sol, read_vtu, write_vtu, modify_vtu_file
from . import numpy
from .autodiff import grad, vmap, jacfwd, jacrev, value_and_grad, hessian
from . import mma
from . import flatten_util
from .learned_pc import LearnedPC
from .multigrid import Multigrid
from .solver import (solver, ad_wrapper, implicit_vjp, get_A, linear_solver, apply_bc_vec, assign_bc, copy_bc,
                     TangentMatrix, krylov)

logger = logging.getLogger("jax_fem_amd")
if not logger.handlers:
    _h = logging.StreamHandler()
    _h.setFormatter(logging.Formatter("[%(asctime)s][%(levelname)s] jax_fem_amd: %(message)s"))
    logger.addHandler(_h)
    logger.setLevel(logging.INFO)

__all__ = ["HipFemError", "Problem", "solver", "ad_wrapper", "implicit_vjp", "get_A", "linear_solver", "apply_bc_vec",
           "assign_bc", "copy_bc", "Mesh", "box_mesh", "box_mesh_hex27", "box_mesh_tet", "box_mesh_gmsh", "rectangle_mesh",
           "rectangle_mesh_tri", "get_meshio_cell_type",
           "FiniteElement", "LinearElastic", "NeoHookean", "Poisson", "Material", "TangentMatrix", "krylov",
           "LearnedPC", "Multigrid", "numpy", "grad", "vmap", "jacfwd", "jacrev", "value_and_grad", "hessian", "mma", "flatten_util",
           "save_sol", "read_vtu", "write_vtu", "modify_vtu_file"]
