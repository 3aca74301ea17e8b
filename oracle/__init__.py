"""FP64 CPU oracle of the reference's implicit solve path (TEST INFRASTRUCTURE).

This package is a from-scratch NumPy/SciPy restatement of JAX-FEM's algorithm
for the implicit solve path: reference-cell basis and quadrature, geometry,
cell residuals and element Jacobians, COO->CSR assembly with row elimination,
Jacobi-BiCGSTAB with jax's stop rule, Newton, and the adjoint (implicit VJP).
Every function cites the reference file:line it restates.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import it, and only as the checker or the timed CPU baseline.  The product
(``jax-fem_amd``) never imports it; the product path fails loudly when its HIP
library is missing instead of falling back to anything in here.

Pinning (which check fixes which function) is listed per module docstring.
"""
