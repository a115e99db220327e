"""The GPU POD (fpe_pod_fit) and its one-sided Jacobi SVD (fpe_svd_jacobi)
against numpy's LAPACK SVD (tests/pod_ref.py), on synthetic matrices with
known structure and on real q(T) snapshots of the surrogate's grid."""
import time

import numpy as np
import pytest

import pod_ref
from fpe_option_amd import surrogate as S

pytestmark = pytest.mark.gpu


def _graded(rng, m, n, lo):
    U = np.linalg.qr(rng.standard_normal((m, n)))[0]
    V = np.linalg.qr(rng.standard_normal((n, n)))[0]
    return (U * np.logspace(0, np.log10(lo), n)) @ V.T


@pytest.mark.parametrize("case", ["gauss_300x40", "wide_5x7", "payoff_like_122x64", "one_column", "graded_500x50"])
def test_svd_jacobi_matches_lapack(engine, case):
    rng = np.random.default_rng(11)
    A = {"gauss_300x40": lambda: rng.standard_normal((300, 40)),
         "wide_5x7": lambda: rng.standard_normal((5, 7)),
         "payoff_like_122x64": lambda: S.payoff_matrix(np.linspace(0, 180, 90)) @ np.linalg.qr(
             rng.standard_normal((90, 64)))[0],
         "one_column": lambda: rng.standard_normal((1000, 1)),
         "graded_500x50": lambda: _graded(rng, 500, 50, 1e-10)}[case]()
    sv, Vt = S.svd_jacobi(A)
    sv_r, Vt_r = pod_ref.svd(A)
    r = min(A.shape)
    assert sv.shape == (r,) and Vt.shape == (r, A.shape[1])
    print(f"{case}: max |dsv| / sv0 = {np.max(np.abs(sv - sv_r[:r])) / sv_r[0]:.2e}")
    np.testing.assert_allclose(sv, sv_r[:r], rtol=0, atol=1e-13 * sv_r[0])
    assert np.all(np.diff(sv) <= 0)
    # right singular vectors (up to sign) where the singular values are separated
    gap = np.minimum(np.abs(np.diff(np.r_[np.inf, sv_r[:r]])), np.abs(np.diff(np.r_[sv_r[:r], -np.inf])))
    sep = gap > 1e-3 * sv_r[0]
    V, V_r = Vt.T, pod_ref.align_columns(Vt_r[:r].T, Vt.T)
    np.testing.assert_allclose(V[:, sep], V_r[:, sep], atol=1e-9)
    # each vector's largest-magnitude entry is positive
    assert np.all(V[np.argmax(np.abs(V), axis=0), np.arange(r)] > 0)
    np.testing.assert_allclose(Vt @ Vt.T, np.eye(r), atol=1e-12)


def test_pod_fit_low_rank_images(engine):
    """Snapshots whose marginal images span 3 dimensions plus components the
    marginal operator annihilates: the three modes, their singular values and
    the mean match LAPACK, and the images are reproduced exactly."""
    rng = np.random.default_rng(1)
    n_dof, n_q = 300, 25
    G = rng.standard_normal((n_q, n_dof))
    null = np.linalg.svd(G)[2][n_q:]
    A = rng.standard_normal((3, n_dof))
    Q = 2.0 + rng.standard_normal((2000, 3)) @ A + 50.0 * rng.standard_normal((2000, n_dof - n_q)) @ null
    mu, Phi, P, w0, sv = S.fit_pod(Q, G, k=3)
    mu_r, Phi_r, P_r, w0_r, sv_r = pod_ref.fit_pod(Q, G, k=3)
    np.testing.assert_allclose(mu, mu_r, rtol=0, atol=1e-12 * np.abs(mu_r).max())
    np.testing.assert_allclose(w0, w0_r, rtol=0, atol=1e-12 * np.abs(w0_r).max())
    np.testing.assert_allclose(sv[:3], sv_r[:3], rtol=1e-12)
    assert np.all(sv[3:] < 1e-10 * sv[0])
    np.testing.assert_allclose(P, pod_ref.align_columns(P_r, P), atol=1e-10)
    np.testing.assert_allclose(Phi, pod_ref.align_columns(Phi_r, Phi), atol=1e-8 * np.abs(Phi_r).max())
    np.testing.assert_allclose(P.T @ P, np.eye(3), atol=1e-12)
    W = Q @ G.T
    c = (W - w0) @ P
    np.testing.assert_allclose(c @ P.T + w0, W, atol=1e-9 * np.abs(W).max())


def test_pod_fit_on_solver_snapshots(engine):
    """1024 q(T) snapshots of the surrogate grid: singular values and the
    marginals reconstructed from 64 modes match the LAPACK POD of the same
    snapshots, for 64 held-out solves, in price units."""
    Pr = S.sample_params(1024 + 64, 4321)
    Q, st = S.terminal_states(Pr)
    assert np.all(st == 0)
    s_q, G = S.pricing_operator()
    Qf, Qh = Q[:1024], Q[1024:]
    t = time.perf_counter()
    mu, Phi, P, w0, sv = S.fit_pod(Qf, G, k=64)
    t_gpu = time.perf_counter() - t
    t = time.perf_counter()
    mu_r, Phi_r, P_r, w0_r, sv_r = pod_ref.fit_pod(Qf, G, k=64)
    t_np = time.perf_counter() - t
    rel = np.abs(sv[:64] - sv_r[:64]) / sv_r[:64]
    print(f"POD of 1024 x {Q.shape[1]} snapshots: GPU {t_gpu * 1e3:.0f} ms, numpy {t_np * 1e3:.0f} ms; "
          f"sv rel err modes < 40: {rel[:40].max():.2e}, 40..63: {rel[40:].max():.2e}")
    assert rel[:40].max() < 1e-9
    np.testing.assert_allclose(sv, sv_r[:len(sv)], rtol=0, atol=1e-12 * sv_r[0])
    np.testing.assert_allclose(mu, mu_r, rtol=0, atol=1e-13 * np.abs(mu_r).max())
    # held-out marginals projected on either basis; payoffs at 61 strikes
    W = Qh @ G.T
    proj = (W - w0) @ P @ P.T + w0
    proj_r = (W - w0_r) @ P_r @ P_r.T + w0_r
    A = S.payoff_matrix(s_q)
    d = np.abs((proj - proj_r) @ A.T).max()
    err = np.abs((proj - W) @ A.T).max()
    err_r = np.abs((proj_r - W) @ A.T).max()
    print(f"held-out payoff integrals: GPU vs LAPACK basis {d:.2e}; 64-mode projection error GPU {err:.2e}, "
          f"LAPACK {err_r:.2e}")
    # the two bases differ by far less than what 64 modes leave out
    assert d <= max(1e-9, 0.01 * err_r)
    assert err <= 1.01 * err_r + 1e-12
    # the q-space modes map onto the marginal modes (leading 40: sv above 1e-6 of the first)
    np.testing.assert_allclose(G @ Phi[:, :40], P[:, :40], atol=1e-9)
