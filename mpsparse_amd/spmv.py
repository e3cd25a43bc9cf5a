"""Drop-in for the reference's ``spmv`` module (iterative/BiCGSTAB/spmv.cpp:892-905).

``csr_tensor(keys, values, row_ptr, col_ind, out_vals, num_rows, num_cols)`` sorts
the packed COO keys on the GPU, keeps the CSR resident on the device and, unlike
the reference, also writes ``row_ptr``/``col_ind``/``out_vals`` when they have the
right sizes. Methods: ``mv(x, b)``, ``cgd(b, x, ...)``, ``bicgstab(b, x, ...)``,
``richardson(b, x, omega, iters)``. Solver keywords (all optional): ``tol`` (relative
residual, default 1e-6), ``maxiter`` (10000), ``check_every``, ``criterion``
(``"relres"`` or ``"reference"`` = sum(r^2) < tol tested every 50 steps),
``use_x0``, ``precond`` (``"none"``/``"jacobi"``, CG), ``values``
(``"f32"``/``"bf16"``/``"f16"``). Solvers return an info dict.

``release_workspace()`` returns the cached temporaries of matrix builds (a
scratch arena per device and stream that the library keeps for the next
build) to the system.
"""
from . import load_ext

_C = load_ext()
csr_tensor = _C.csr_tensor
release_workspace = _C.release_workspace

__all__ = ["csr_tensor", "release_workspace"]
