"""mpsparse_amd -- MI355X-native sparse iterative solvers behind the mpsparse API.

The reference's PyTorch extension modules map to submodules here:

* ``mpsparse_amd.spmv``        -- ``csr_tensor`` with ``.mv``, ``.cgd``, ``.bicgstab``
  (iterative/BiCGSTAB/spmv.cpp:892-905), plus ``.richardson``;
* ``mpsparse_amd.cgd``         -- ``csr_tensor.iter_solve`` = CG (iterative/cgd/spmv.cpp:751-764);
* ``mpsparse_amd.richardson``  -- ``csr_tensor.iter_solve`` = Richardson, omega 0.01,
  1000 steps (iterative/richardson/spmv.cpp:499-512);
* ``mpsparse_amd.coo_csr``     -- ``coo_csr(...)`` (spmv_kernel/coo_csr/coo_csr/coo_csr/coo_to_csr.cpp:282-283).

All compute runs in ``libmpsparse_amd.so`` (HIP kernels for gfx950 behind the C ABI
in ``include/mpsparse_amd.h``). Importing a compute module without the built
extension raises ImportError; there is no CPU fallback.
"""
from .problems import pack_keys, unpack_keys  # noqa: F401

__all__ = ["pack_keys", "unpack_keys", "load_ext"]


_exit_hook = []


def _release_at_exit(ext):
    """Return the library's cached device blocks and idle scratch while the
    interpreter and the HIP runtime are both fully alive (handles destroyed
    later, during finalisation, still work; the library itself makes no HIP
    call once exit() has begun)."""
    try:
        ext.release_workspace()
    except Exception:  # no device, or nothing was ever allocated
        pass


def load_ext():
    """Import the compiled extension (``mpsparse_amd._C``), failing loudly."""
    try:
        from . import _C
    except ImportError as e:  # pragma: no cover - exercised only without a build
        raise ImportError(
            "mpsparse_amd native extension is not built: run `python -m mpsparse_amd.build` "
            f"(hipcc for gfx950). Original error: {e}") from e
    if not _exit_hook:
        import atexit

        atexit.register(_release_at_exit, _C)
        _exit_hook.append(True)
    return _C
