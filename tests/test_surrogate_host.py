"""Host-side pieces of the surrogate: the training box, input scaling, the
FP64 numpy references of the POD steps (the GPU POD's checkers) and the
price-space rotation, checked against closed forms."""
import numpy as np

import pod_ref
from fpe_option_amd import surrogate as S


def test_box_and_scaling():
    P = S.sample_params(1000, 5)
    assert P.shape == (1000, 7) and np.all(S.in_box(P))
    X = S.normalise(P)
    assert np.all(X >= -1) and np.all(X <= 1)
    np.testing.assert_allclose(S.normalise(S.BOX_LO[None]), -1.0)
    np.testing.assert_allclose(S.normalise(S.BOX_HI[None]), 1.0)
    Q = P.copy()
    Q[3, 2] = 0.9  # sigma above the box
    assert not S.in_box(Q)[3] and S.in_box(Q).sum() == 999


def test_pod_recovers_a_low_rank_subspace():
    rng = np.random.default_rng(0)
    U = np.linalg.qr(rng.standard_normal((300, 3)))[0]
    Q = 5.0 + (rng.standard_normal((500, 3)) * [3.0, 2.0, 1.0]) @ U.T
    mu, Phi, ev = pod_ref.pod(Q, k=3)
    np.testing.assert_allclose(mu, Q.mean(axis=0), atol=1e-12)
    # projector onto the recovered modes equals the projector onto U
    np.testing.assert_allclose(Phi @ Phi.T, U @ U.T, atol=1e-10)
    assert ev[0] > ev[1] > ev[2] > 0
    R = (Q - mu) - ((Q - mu) @ Phi) @ Phi.T
    assert np.max(np.abs(R)) < 1e-10


def test_pricing_inner_product_pod():
    # snapshots whose marginal images span 3 dimensions, plus components the
    # marginal operator annihilates: the basis must resolve the images exactly
    # and its image must be orthonormal
    rng = np.random.default_rng(1)
    n_dof, n_q = 60, 25
    G = rng.standard_normal((n_q, n_dof))
    null = np.linalg.svd(G)[2][n_q:]            # rows spanning ker G
    A = rng.standard_normal((3, n_dof))
    Q = 2.0 + rng.standard_normal((200, 3)) @ A + 50.0 * rng.standard_normal((200, n_dof - n_q)) @ null
    mu, Phi, P, w0, sv = pod_ref.fit_pod(Q, G, k=3)
    np.testing.assert_allclose(P.T @ P, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(G @ Phi, P, atol=1e-9)
    np.testing.assert_allclose(w0, G @ Q.mean(axis=0), atol=1e-10)
    W = Q @ G.T
    c = (W - w0) @ P
    np.testing.assert_allclose(c @ P.T + w0, W, atol=1e-9)
    # Euclidean q-space modes of the same snapshots chase the annihilated
    # components and miss the images
    mu_e, Phi_e, _ = pod_ref.pod(Q, k=3)
    We = ((Q - mu_e) @ Phi_e @ Phi_e.T + mu_e) @ G.T
    assert np.max(np.abs(We - W)) > 1e3 * np.max(np.abs(c @ P.T + w0 - W))


def test_price_space_basis():
    """Squared coordinate differences in the price-space basis are the summed
    squared payoff-weighted marginal differences over the strike grid (calls
    and puts), and the basis reproduces the same marginals as the POD one."""
    rng = np.random.default_rng(3)
    s_q = np.sort(rng.uniform(0.0, 180.0, 90))
    P = np.linalg.qr(rng.standard_normal((90, 12)))[0]
    Pb, R = S.price_basis(P, s_q, svd=pod_ref.svd)
    c1, c2 = rng.standard_normal(12), rng.standard_normal(12)
    A = S.payoff_matrix(s_q)
    np.testing.assert_allclose(np.linalg.norm(A @ P @ (c1 - c2)), np.linalg.norm(R @ (c1 - c2)), rtol=1e-10)
    np.testing.assert_allclose(Pb @ (R @ c1), P @ c1, atol=1e-10)
    assert A.shape == (2 * len(S.STRIKES), 90)
    # payoffs: call (s - K)+ rows first, put (K - s)+ rows after
    np.testing.assert_allclose(A[0], np.maximum(s_q - S.STRIKES[0], 0.0))
    np.testing.assert_allclose(A[len(S.STRIKES)], np.maximum(S.STRIKES[0] - s_q, 0.0))
