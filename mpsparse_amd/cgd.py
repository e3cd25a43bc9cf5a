"""Drop-in for the reference's CG-flavoured ``spmv`` module (iterative/cgd/spmv.cpp:751-764),
whose ``iter_solve(b, x)`` runs CG."""
from . import load_ext

_C = load_ext()


class csr_tensor(_C.csr_tensor):
    def iter_solve(self, b, x, **kw):
        """CG on A x = b (x is the output)."""
        return self.cgd(b, x, **kw)


__all__ = ["csr_tensor"]
