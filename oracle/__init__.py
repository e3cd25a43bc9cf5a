"""CPU oracle for the mpsparse_amd hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import anything from this package, and only as the checker (or as the timed
CPU baseline). The product library never calls it and has no CPU fallback.

Modules (each function cites the reference file:line it restates):

* :mod:`oracle.coo_csr` -- packed-key COO->CSR: numpy stable argsort form, the
  C restatement of the reference's 4-bit LSD radix sort (``lsd_coo_csr.c``),
  and a pure-Python brute force for tiny inputs.
* :mod:`oracle.spmv`    -- CSR / BCSR SpMV in FP64 and sequential FP32.
* :mod:`oracle.krylov`  -- CG, BiCGSTAB, Richardson (FP64 or FP32), plus the
  scipy CPU baseline solves.

How each function is pinned (golden vectors, closed forms, independent
formulations) is listed in the tests under ``tests/test_oracle_*.py``.
"""
