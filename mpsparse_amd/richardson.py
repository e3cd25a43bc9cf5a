"""Drop-in for the reference's Richardson ``spmv`` module (iterative/richardson/spmv.cpp:499-512),
whose ``iter_solve(b, x)`` runs 1000 steps of x <- x + 0.01 (b - A x) from x = 0
(spmv.cpp:376-425)."""
from . import load_ext

_C = load_ext()


class csr_tensor(_C.csr_tensor):
    def iter_solve(self, b, x, omega=0.01, iters=1000, **kw):
        """Fixed-step Richardson iteration (x is the output)."""
        return self.richardson(b, x, omega=omega, iters=iters, **kw)


__all__ = ["csr_tensor"]
