"""Batched Jacobi-PCG: many independent SPD systems with up to 4 right-hand
sides each, solved together on the GPU (the reference builds one csr_tensor
per system, iterative/BiCGSTAB/spmv.cpp:892-905; this is the batching addition).

``batch_tensor(A, sys_off)`` takes a csr_tensor whose diagonal blocks
[sys_off[s], sys_off[s+1]) are the systems (entries coupling two systems are
rejected) and keeps A alive. ``.pcg(b, x, tol=1e-6)`` solves every (system,
rhs) pair with Jacobi-preconditioned CG: b and x are (num_rows, nrhs) float32
(nrhs <= 4; the right-hand sides of a row are adjacent), and the result dict
carries per-pair ``status`` (1 converged, 2 breakdown, 3 maxiter,
4 stagnated), ``iterations`` and ``relres`` (the true relative residual of
the returned x), plus the autotuned tile schedule and the device time.
"""
from __future__ import annotations

import numpy as np

from . import load_ext
from .problems import batched_diffusion_sides, diffusion_keys_torch

_C = load_ext()
batch_tensor = _C.batch_tensor

__all__ = ["batch_tensor", "batched_diffusion"]


def batched_diffusion(num_systems=512, lo=32, hi=512, sigma=1.5, seed=0, device="cuda"):
    """Block-diagonal batch of 2D 5-point log-normal diffusion systems.

    Sides are uniform integers in [lo, hi] (problems.batched_diffusion_sides,
    numpy seed `seed`); system s has conductivities exp(sigma N(0,1)) drawn on
    the device from one torch generator seeded with `seed`, in system order.
    Returns (keys int64, vals float32, sys_off int64 numpy[num_systems + 1],
    sides) with keys already offset to the block-diagonal position."""
    import torch

    sides = batched_diffusion_sides(num_systems, lo, hi, seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    keys, vals = [], []
    off = np.zeros(num_systems + 1, dtype=np.int64)
    for s, side in enumerate(sides):
        k, v, n = diffusion_keys_torch((1, int(side), int(side)), sigma, gen, device)
        o = int(off[s])
        keys.append(k + ((o << 32) | o))
        vals.append(v)
        off[s + 1] = o + n
    return torch.cat(keys), torch.cat(vals), off, sides
