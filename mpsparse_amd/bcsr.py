"""Drop-in for the reference's ``bcsr`` module (iterative/BiCGSTAB/bcsr.cpp:953-965).

``bcsr_tensor(keys, values, row_ptr, col_ind, num_rows, num_cols, block_size=4)``
takes the BSR arrays (``values`` = row-major blocks, ``row_ptr`` = block-row
pointer, ``col_ind`` = block columns; ``keys`` unused), as the reference's
``MetalBCSRSolver`` passes them (iterative/BiCGSTAB/bcsr_benchmark.py:87-95);
with an empty ``row_ptr`` it converts packed COO ``keys``/``values`` on the GPU
instead. Block sizes 4, 8 and 16. Methods ``mv``, ``cgd``, ``bicgstab``,
``richardson``, ``export``; vectors may be logical-size or padded to whole blocks.
"""
from . import load_ext

_C = load_ext()
bcsr_tensor = _C.bcsr_tensor

__all__ = ["bcsr_tensor"]
