"""Drop-in for the reference's standalone ``coo_csr`` module
(spmv_kernel/coo_csr/coo_csr/coo_csr/coo_to_csr.cpp:251-283): writes the CSR of the
packed COO keys into ``row_ptr`` (int32[n+1]), ``col_ind`` (int32[nnz]) and
``out_vals`` (float32[nnz]); duplicate keys are kept in input order."""
from . import load_ext

_C = load_ext()
coo_csr = _C.coo_csr

__all__ = ["coo_csr"]
