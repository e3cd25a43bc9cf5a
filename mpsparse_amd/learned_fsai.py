"""Learned FSAI preconditioner: an edge MLP predicts G on A's lower pattern,
M = G^T G, applied inside CG as two SpMVs (all kernels in libmpsparse_amd).

Network: per lower-pattern edge (i, j <= i) the 8 features of
``fsai_tensor.features()`` (normalised a_ij, a_ii, a_jj, row nnz, row |sum|,
diagonal flag) -> 64 -> 64 -> 1 with ReLU; BF16 matrix-core operands with FP32
accumulation (forward and backward are hand-written HIP kernels). The output
o_e becomes g_ij = o_e / sqrt(a_jj), so o = 1 on the diagonal and 0 elsewhere
is Jacobi. A non-positive g_ii makes the solver fall back to plain CG.

Training (``train``): log-normal diffusion matrices (2D 5-point / 3D 7-point,
sigma in [0.5, 2.5], n between 16k and 1M) are generated on the GPU; for each
one, residuals of short plain-CG runs (plus one white-noise vector) serve as
probes w, and Adam minimises sum_w ||G A G^T w - w||^2 / ||w||^2.
"""
from __future__ import annotations

import math
import os
import time

import numpy as np

from . import load_ext

_C = load_ext()
fsai_tensor = _C.fsai_tensor
NPARAMS = int(_C.mlp_num_params)
OW1, OB1 = 0, 512
OW2 = OB1 + 64
OB2 = OW2 + 64 * 64
OW3 = OB2 + 64
OB3 = OW3 + 64

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "learned_fsai_mlp.npz")


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation; output bias 1 and tiny output weights start the
    network near o = 1 everywhere (diagonal ~ Jacobi)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAMS, dtype=np.float64)
    P[OW1:OB1] = rng.standard_normal(512) * math.sqrt(2.0 / 5.0)
    P[OW2:OB2] = rng.standard_normal(4096) * math.sqrt(2.0 / 64.0)
    P[OW3:OB3] = rng.standard_normal(64) * 0.01
    P[OB3] = 1.0
    return P.astype(np.float32)


def load_params(path: str = DEFAULT_WEIGHTS, device="cuda"):
    import torch

    with np.load(path, allow_pickle=False) as z:
        P = z["params"].astype(np.float32)
    if P.size != NPARAMS:
        raise ValueError(f"{path}: {P.size} parameters, expected {NPARAMS}")
    return torch.from_numpy(P).to(device)


def save_params(P, path: str = DEFAULT_WEIGHTS, meta: dict | None = None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    arr = P.detach().cpu().numpy().astype(np.float32) if hasattr(P, "detach") else np.asarray(P, np.float32)
    np.savez(path, params=arr, meta=np.array(repr(meta or {})))


def build(A, params):
    """fsai_tensor for csr_tensor A with G from the network; returns (F, nonpositive_diag)."""
    F = fsai_tensor(A)
    nonpos = F.set_network(params)
    return F, nonpos


def solve(A, F, b, x, **kw):
    """FSAI-preconditioned CG (falls back to plain CG on a non-positive diagonal of G)."""
    return _C.pcg_fsai(A, F, b, x, **kw)


# ---------------------------------------------------------------------------
# training data on the GPU
def diffusion_keys_torch(shape, sigma, gen, device="cuda"):
    """Packed keys + FP32 values of a log-normal diffusion matrix (the 2D / 3D
    finite-volume operator of problems.diffusion_3d), generated with torch on
    the device. shape = (nz, ny, nx); nz = 1 gives the 5-point 2D operator."""
    import torch

    nz, ny, nx = shape
    n = nx * ny * nz
    k = torch.exp(sigma * torch.randn((nz, ny, nx), generator=gen, device=device, dtype=torch.float64))
    idx = torch.arange(n, device=device, dtype=torch.int64).reshape(nz, ny, nx)
    cols = torch.full((nz, ny, nx, 7), -1, dtype=torch.int64, device=device)
    vals = torch.zeros((nz, ny, nx, 7), dtype=torch.float64, device=device)
    diag = torch.zeros((nz, ny, nx), dtype=torch.float64, device=device)

    def harm(a, b):
        return 2.0 * a * b / (a + b)

    cx = harm(k[:, :, 1:], k[:, :, :-1])
    diag[:, :, 1:] += cx
    diag[:, :, :-1] += cx
    diag[:, :, 0] += 2.0 * k[:, :, 0]
    diag[:, :, -1] += 2.0 * k[:, :, -1]
    cols[:, :, 1:, 2] = idx[:, :, :-1]
    vals[:, :, 1:, 2] = -cx
    cols[:, :, :-1, 4] = idx[:, :, 1:]
    vals[:, :, :-1, 4] = -cx
    cy = harm(k[:, 1:, :], k[:, :-1, :])
    diag[:, 1:, :] += cy
    diag[:, :-1, :] += cy
    diag[:, 0, :] += 2.0 * k[:, 0, :]
    diag[:, -1, :] += 2.0 * k[:, -1, :]
    cols[:, 1:, :, 1] = idx[:, :-1, :]
    vals[:, 1:, :, 1] = -cy
    cols[:, :-1, :, 5] = idx[:, 1:, :]
    vals[:, :-1, :, 5] = -cy
    if nz > 1:
        cz = harm(k[1:, :, :], k[:-1, :, :])
        diag[1:, :, :] += cz
        diag[:-1, :, :] += cz
        cols[1:, :, :, 0] = idx[:-1, :, :]
        vals[1:, :, :, 0] = -cz
        cols[:-1, :, :, 6] = idx[1:, :, :]
        vals[:-1, :, :, 6] = -cz
        diag[0, :, :] += 2.0 * k[0, :, :]
        diag[-1, :, :] += 2.0 * k[-1, :, :]
    cols[..., 3] = idx
    vals[..., 3] = diag
    cols = cols.reshape(n, 7)
    vals = vals.reshape(n, 7)
    rows = torch.arange(n, device=device, dtype=torch.int64)[:, None].expand(n, 7)
    m = cols >= 0
    keys = (rows[m] << 32) | cols[m]
    return keys.contiguous(), vals[m].float().contiguous(), n


def sample_shape(rng: np.random.Generator):
    """2D side in [128, 1000] or 3D side in [26, 100] (n between ~16k and 1M)."""
    if rng.random() < 0.5:
        s = int(rng.integers(128, 1001))
        return (1, s, s)
    s = int(rng.integers(26, 101))
    return (s, s, s)


def harvest_probes(A, n, gen, device="cuda", iters=(5, 20, 80)):
    """Residuals of plain-CG runs stopped after `iters` steps, plus white noise,
    normalised to unit norm: (k, n) float32."""
    import torch

    W = []
    b = torch.randn(n, generator=gen, device=device)
    for it in iters:
        x = torch.zeros(n, device=device)
        A.cgd(b, x, tol=0.0, maxiter=int(it), verify=False)
        ax = torch.empty(n, device=device)
        A.mv(x, ax)
        W.append(b - ax)
    W.append(torch.randn(n, generator=gen, device=device))
    W = torch.stack(W)
    return (W / torch.linalg.vector_norm(W, dim=1, keepdim=True)).contiguous()


def train(params=None, matrices=2000, updates_per_matrix=2, lr=2e-3, seed=0, device="cuda", log=print,
          sigma_range=(0.5, 2.5)):
    """Train the edge MLP on freshly generated matrices; returns (params, history)."""
    import torch

    from . import spmv

    if params is None:
        params = torch.from_numpy(init_params(seed)).to(device)
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    grad = torch.zeros_like(params)
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    hist = []
    step = 0
    t0 = time.time()
    empty = torch.empty(0)
    for k in range(matrices):
        shape = sample_shape(rng)
        sigma = float(rng.uniform(*sigma_range))
        keys, vals, n = diffusion_keys_torch(shape, sigma, gen, device)
        A = spmv.csr_tensor(keys, vals, empty, empty, empty, n, n)
        del keys, vals
        F = fsai_tensor(A)
        W = harvest_probes(A, n, gen, device)
        for _ in range(updates_per_matrix):
            loss = F.train_step(params, W, grad) / W.shape[0]
            step += 1
            lr_t = lr * 0.5 * (1 + math.cos(math.pi * min(1.0, k / max(1, matrices))))
            _C.mlp_adam(params, grad, m, v, lr=lr_t, step=step)
            hist.append((k, n, sigma, float(loss)))
        if k % 50 == 0 or k == matrices - 1:
            recent = [h[3] for h in hist[-100:]]
            log(f"[train] matrix {k}/{matrices} n={n} sigma={sigma:.2f} loss={np.mean(recent):.4f} "
                f"elapsed {time.time() - t0:.0f}s")
        del F, A, W
    return params, hist
