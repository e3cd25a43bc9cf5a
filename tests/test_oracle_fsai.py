"""Pins the FSAI and edge-MLP oracle (CPU only)."""
import numpy as np
import pytest
import scipy.sparse as sp

from mpsparse_amd.problems import diffusion_2d
from oracle import fsai as OF


def _small():
    M = diffusion_2d(7, sigma=1.0, seed=3)
    A = M.to_scipy()
    A.sort_indices()
    return A


def test_classical_fsai_defining_property():
    A = _small()
    r, c, _ = OF.lower_pattern(A.indptr, A.indices)
    g = OF.classical_fsai(A)
    G = OF.G_matrix(A.shape[0], r, c, g)
    GA = (G @ A).toarray()
    for i in range(A.shape[0]):
        J = c[r == i]
        row = GA[i, J]
        gii = g[(r == i) & (c == i)][0]
        assert np.allclose(row[:-1], 0.0, atol=1e-12)
        assert abs(row[-1] * gii - 1.0) < 1e-12
    # GAG^T has unit diagonal (the FSAI normalisation)
    assert np.allclose((G @ A @ G.T).diagonal(), 1.0)


def test_classical_fsai_of_a_diagonal_is_inverse_sqrt():
    d = np.array([4.0, 1.0, 9.0, 2.0])
    A = sp.diags(d).tocsr()
    g = OF.classical_fsai(A)
    assert np.allclose(g, 1 / np.sqrt(d))


def test_probe_gradient_matches_finite_differences():
    A = _small()
    r, c, _ = OF.lower_pattern(A.indptr, A.indices)
    rng = np.random.default_rng(0)
    g = OF.classical_fsai(A) * (1 + 0.1 * rng.standard_normal(r.size))
    W = rng.standard_normal((2, A.shape[0]))
    L, grad = OF.probe_loss_grad(A, r, c, g, W)
    for k in rng.choice(r.size, 12, replace=False):
        h = 1e-6
        gp, gm = g.copy(), g.copy()
        gp[k] += h
        gm[k] -= h
        fd = (OF.probe_loss_grad(A, r, c, gp, W)[0] - OF.probe_loss_grad(A, r, c, gm, W)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(1.0, abs(fd))


def test_mlp_backward_matches_autograd():
    import torch

    rng = np.random.default_rng(1)
    P = rng.standard_normal(OF.NPARAMS) * 0.3
    F = rng.standard_normal((50, 8))
    go = rng.standard_normal(50)
    g = OF.mlp_backward(P, F, go, bf16=False)
    Pt = torch.tensor(P, dtype=torch.float64, requires_grad=True)
    W1 = Pt[OF.OW1:OF.OB1].reshape(64, 8)
    b1 = Pt[OF.OB1:OF.OW2]
    W2 = Pt[OF.OW2:OF.OB2].reshape(64, 64)
    b2 = Pt[OF.OB2:OF.OW3]
    w3 = Pt[OF.OW3:OF.OB3]
    b3 = Pt[OF.OB3]
    Ft = torch.tensor(F)
    out = torch.relu(torch.relu(Ft @ W1.T + b1) @ W2.T + b2) @ w3 + b3
    (out * torch.tensor(go)).sum().backward()
    assert np.allclose(g, Pt.grad.numpy(), rtol=1e-10, atol=1e-10)
    o, _, _ = OF.mlp_forward(P, F, bf16=False)
    assert np.allclose(o, out.detach().numpy())


def test_features_on_the_diagonal():
    A = _small()
    f, cs = OF.features(A.indptr, A.indices, A.data)
    r, c, _ = OF.lower_pattern(A.indptr, A.indices)
    dg = r == c
    assert np.allclose(f[dg, 0], 1.0) and np.allclose(f[dg, 1], 0.0) and np.all(f[dg, 4] == 1)
    assert np.all(f[~dg, 4] == 0) and np.all(f[~dg, 0] < 0)
    assert np.allclose(cs, 1 / np.sqrt(A.diagonal()[c]))


def test_row_coupling_features_give_classical_fsai_on_stencils():
    """7-point stencil: lower neighbours are mutually uncoupled, so classical FSAI
    has the closed form g_ii sqrt(a_ii) = 1/sqrt(1 - s_low), g_ij sqrt(a_jj) =
    -f0 / sqrt(1 - s_low) in the features f0 and s_low (f[5]); s_all (f[6]) is
    the same sum over the whole row."""
    from mpsparse_amd.problems import diffusion_3d

    M = diffusion_3d(5, 4, 3, sigma=1.5, seed=2)
    S = M.to_scipy()
    f, cs = OF.features(M.row_ptr, M.col_ind, M.vals)
    r, c, _ = OF.lower_pattern(M.row_ptr, M.col_ind)
    g = OF.classical_fsai(S)
    out = g / cs  # g_ij sqrt(a_jj)
    closed = np.where(r == c, 1.0, -f[:, 0]) / np.sqrt(1.0 - f[:, 5])
    assert np.allclose(out, closed, rtol=1e-12, atol=1e-14)
    d = S.diagonal()
    Sd = S.tocoo()
    off = Sd.row != Sd.col
    s_all = np.bincount(Sd.row[off], weights=Sd.data[off] ** 2 / (d[Sd.row[off]] * d[Sd.col[off]]),
                        minlength=S.shape[0])
    assert np.allclose(f[:, 6], s_all[r], rtol=1e-12)
    assert np.all(f[:, 6] >= f[:, 5])
    # t (f[7]) is classical FSAI's diagonal output; off the diagonal -f0 t
    assert np.allclose(out, np.where(r == c, 1.0, -f[:, 0]) * f[:, 7], rtol=1e-12)


def test_stated5_features_are_the_first_five():
    A = _small()
    f8, cs8 = OF.features(A.indptr, A.indices, A.data)
    f5, cs5 = OF.features(A.indptr, A.indices, A.data, feature_set="stated5")
    assert np.array_equal(f5[:, :5], f8[:, :5]) and np.all(f5[:, 5:] == 0) and np.array_equal(cs5, cs8)
    with pytest.raises(ValueError):
        OF.features(A.indptr, A.indices, A.data, feature_set="other")


@pytest.mark.parametrize("case", ["stencil", "coupled"])
def test_rownorm_of_classical_shape_is_classical(case):
    """Classical FSAI has diag(G A G^T) = 1, so any positive row rescaling of it
    normalises back to it exactly; on a stencil the shape o_ii = 1,
    o_ij = -a_ij / sqrt(a_ii a_jj) (what a 5-input network can compute from
    its own edge) gives classical FSAI."""
    from mpsparse_amd.problems import diffusion_3d

    if case == "stencil":
        M = diffusion_3d(5, 4, 3, sigma=1.5, seed=2)
        S = M.to_scipy()
    else:  # a 9-point-like pattern whose lower neighbours are coupled
        S = _small()
        S = (S @ S).tocsr()
        S.sort_indices()
    r, c, _ = OF.lower_pattern(S.indptr, S.indices)
    g = OF.classical_fsai(S)
    rng = np.random.default_rng(4)
    scale = np.exp(rng.standard_normal(S.shape[0]))[r]
    G, s = OF.rownorm(S, r, c, g * scale)
    assert np.allclose(G, g, rtol=1e-12, atol=1e-14)
    Gm = OF.G_matrix(S.shape[0], r, c, G)
    assert np.allclose((Gm @ S @ Gm.T).diagonal(), 1.0)
    if case == "stencil":
        f, cs = OF.features(S.indptr, S.indices, S.data, feature_set="stated5")
        shape = np.where(r == c, 1.0, -f[:, 0]) * cs
        G2, _ = OF.rownorm(S, r, c, shape)
        assert np.allclose(G2, g, rtol=1e-12, atol=1e-14)


def test_rownorm_grad_matches_finite_differences():
    A = _small()
    A = (A @ A).tocsr()  # coupled lower neighbours exercise every term of N_i
    A.sort_indices()
    r, c, _ = OF.lower_pattern(A.indptr, A.indices)
    rng = np.random.default_rng(5)
    gt = OF.classical_fsai(A) * (1 + 0.2 * rng.standard_normal(r.size))
    W = rng.standard_normal((2, A.shape[0]))

    def loss(x):
        G, _ = OF.rownorm(A, r, c, x)
        return OF.probe_loss_grad(A, r, c, G, W)

    L, gG = loss(gt)
    grad = OF.rownorm_grad(A, r, c, gt, gG)
    for k in rng.choice(r.size, 16, replace=False):
        h = 1e-6 * max(1.0, abs(gt[k]))
        xp, xm = gt.copy(), gt.copy()
        xp[k] += h
        xm[k] -= h
        fd = (loss(xp)[0] - loss(xm)[0]) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, grad[k])
