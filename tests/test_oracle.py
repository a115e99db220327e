"""Pins the FP64 CPU oracle before the GPU path is compared against it.

Each check is chosen so that a plausible restatement mistake fails it:
reference-held printed prices (README.md:385-403,661-683 and
python/examples/Barrier_Call_Option_Pricing.ipynb outputs, printed to 4-6
digits), closed forms (Gauss-Legendre, B-spline identities, matrix
exponential, Heston 1993), library routines (numpy leggauss, scipy BSpline,
scipy nnls) and exact invariants (in-out parity, mass normalisation).
"""
import numpy as np
import pytest
import scipy.interpolate as si
import scipy.linalg as sl
import scipy.optimize as so

HESTON = dict(kappa=1.2, theta=0.05, sigma=0.35, rho=-0.4, r=0.1, T=0.6, S0=60.0, V0=0.1)


# ---------------------------------------------------------------- building blocks
def test_gauss_legendre_matches_numpy(oracle):
    # knots.mojo:26-29 gives the 2-point rule weights 0.5 (sum 1, not 2); the
    # path only uses the 4-point rule (domain.mojo:218).
    x, w = np.zeros(2), np.zeros(2)
    oracle.lib().orc_gauss_legendre(2, oracle._dp(x), oracle._dp(w))
    np.testing.assert_allclose(w, [0.5, 0.5])
    # knots.mojo:34-35 builds the 4-point nodes from sqrt(5/6); Gauss-Legendre needs sqrt(6/5)
    x, w = np.zeros(4), np.zeros(4)
    oracle.lib().orc_gauss_legendre(4, oracle._dp(x), oracle._dp(w))
    r = np.sqrt(5.0 / 6.0)
    np.testing.assert_allclose(x, [-np.sqrt(3 / 7 + 2 / 7 * r), -np.sqrt(3 / 7 - 2 / 7 * r),
                                   np.sqrt(3 / 7 - 2 / 7 * r), np.sqrt(3 / 7 + 2 / 7 * r)], atol=1e-15)
    np.testing.assert_allclose(w, np.polynomial.legendre.leggauss(4)[1], atol=1e-15)
    assert np.max(np.abs(x - np.polynomial.legendre.leggauss(4)[0])) > 0.06
    # orders 3 and 5 are exact; the 6-point rule of knots.mojo:55-64 has nodes
    # outside [-1, 1] and is unused on the path
    for order in (3, 5):
        x = np.zeros(order)
        w = np.zeros(order)
        oracle.lib().orc_gauss_legendre(order, oracle._dp(x), oracle._dp(w))
        xr, wr = np.polynomial.legendre.leggauss(order)
        np.testing.assert_allclose(x, xr, atol=1e-15)
        np.testing.assert_allclose(w, wr, atol=1e-15)


def test_de_boor_degree1_known_values(oracle):
    # tests/test_bspline.mojo:16-40
    t = np.array([0.0, 0.0, 0.5, 1.0, 1.0])
    f = lambda i, x: oracle.lib().orc_de_boor(oracle._dp(t), len(t), 1, i, x)
    assert abs(f(0, 0.25) - 0.5) < 1e-12 and abs(f(1, 0.25) - 0.5) < 1e-12 and abs(f(2, 0.25)) < 1e-12
    assert abs(f(0, 0.75)) < 1e-12 and abs(f(1, 0.75) - 0.5) < 1e-12 and abs(f(2, 0.75) - 0.5) < 1e-12
    for x in (0.1, 0.25, 0.5, 0.75, 0.9):
        assert abs(sum(f(i, x) for i in range(3)) - 1.0) < 1e-12


def test_cubic_basis_matches_scipy_and_partition_of_unity(oracle):
    P = oracle.Pipe(**HESTON, n_s=38, n_v=38, num_insert=50, s_max=180.0)
    sk, vk = P.knots()
    x = np.linspace(0.0, 0.999, 301)
    nb = len(sk) - 4
    vals = np.array([[oracle.lib().orc_de_boor(oracle._dp(sk), len(sk), 3, i, xx) for i in range(nb)] for xx in x])
    ref = si.BSpline.design_matrix(x, sk, 3).toarray()
    np.testing.assert_allclose(vals, ref, atol=1e-13)
    np.testing.assert_allclose(vals.sum(1), 1.0, atol=1e-13)


def test_reference_knot_counts(oracle):
    # notebook: 1638 dofs (2125100 x 1638 basis_2d) at n = 38, s in [0,180]
    for ot, B in ((8, 0.0), (2, 50.0)):
        P = oracle.Pipe(**HESTON, n_s=38, n_v=38, num_insert=251, s_max=180.0, option_type=ot, barrier=B)
        assert P.N == 1638, (ot, P.n_sb, P.n_vb)
        assert (P.n_qs, P.n_qv) == (1345, 1580)
        sk, vk = P.knots()
        assert np.all(np.diff(sk) >= 0) and np.all(np.diff(vk) >= 0)
        assert np.all(sk[:4] == 0.0) and np.all(sk[-4:] == 1.0)


def test_recombination_and_gram_factors(oracle):
    P = oracle.Pipe(**HESTON, n_s=16, n_v=16, num_insert=50, s_max=180.0)
    Bs, dBs = P.basis(0, 0), P.basis(0, 1)
    g = P.grid()
    # Dirichlet left drops B0, Neumann right merges: rows sum to 1 except where B0 is active
    rs = np.asarray(Bs.sum(1)).ravel()
    sk, _ = P.knots()
    B0 = np.array([oracle.lib().orc_de_boor(oracle._dp(sk), len(sk), 3, 0, xx) for xx in g["s_norm"]])
    np.testing.assert_allclose(rs + B0, 1.0, atol=1e-12)
    # Ns = Bs^T W Bs exactly the dense Gram matrix
    Ns = P.factor("Ns")
    ref = (Bs.T @ (Bs.multiply(g["sw"][:, None]))).toarray()
    np.testing.assert_allclose(Ns, ref, rtol=1e-13, atol=1e-15)
    # M = Ns (x) Nv (Kronecker identity) on the stencil
    M49, K49 = P.stencils()
    Mf = oracle.dense_to_stencil_matrix(M49, P.n_sb, P.n_vb)
    np.testing.assert_allclose(Mf, np.kron(Ns, P.factor("Nv")), rtol=1e-13, atol=1e-16)
    assert np.allclose(Mf, Mf.T, atol=1e-15)


def test_payoff_known_values(oracle):
    # tests/test_barrier_payoff.mojo:18-111
    cases = [(0, 80, 100, 70, 0), (0, 80, 100, 90, 0), (2, 80, 100, 120, 20), (2, 80, 100, 70, 0),
             (6, 120, 100, 110, 10), (6, 120, 100, 130, 0), (5, 120, 100, 130, 0), (5, 120, 100, 90, 0),
             (8, 0, 100, 120, 20), (8, 0, 100, 80, 0), (9, 0, 100, 80, 20), (9, 0, 100, 120, 0),
             (8, 0, 95, 110, 15), (8, 0, 105, 110, 5), (6, 120, 95, 110, 15), (6, 120, 105, 130, 0)]
    for ot, B, K, S, want in cases:
        assert oracle.lib().orc_payoff(ot, B, K, S) == want


# ---------------------------------------------------------------- Radau IIA
def _radau(oracle, M, K, y0, t1, rtol, atol, first=0.0, cmode=0, exact=0):
    n = len(y0)
    M = np.ascontiguousarray(M, dtype=np.float64)
    K = np.ascontiguousarray(K, dtype=np.float64)
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    out = np.zeros(n)
    st = oracle.OrcStats()
    import ctypes as C
    na = oracle.lib().orc_radau_dense(n, oracle._dp(M), oracle._dp(K), oracle._dp(y0), t1, rtol, atol, 0.0, first,
                                      cmode, exact, oracle._dp(out), C.byref(st))
    return out, na, st.as_dict()


def test_radau_diagonal_matches_exp(oracle):
    # tests/test_radau.mojo:72-162 (tolerance 1e-3 there)
    y, na, _ = _radau(oracle, np.eye(3), np.diag([1.0, 2.0, 3.0]), np.ones(3), 1.0, 1e-3, 1e-6, first=0.1)
    np.testing.assert_allclose(y, np.exp(-np.array([1.0, 2.0, 3.0])), atol=1e-3)
    M, K = np.diag([2.0, 3.0, 4.0]), np.diag([1.0, 2.0, 3.0])
    ref = np.exp(-np.array([0.5, 2 / 3, 0.75]))
    # exact RADAU5 constants: converges like Radau IIA
    y, na, _ = _radau(oracle, M, K, np.ones(3), 1.0, 1e-8, 1e-10, exact=1)
    np.testing.assert_allclose(y, ref, rtol=2e-7)


def test_radau_reference_constants_are_a_perturbed_method(oracle):
    """radau.mojo:59-61,73: perturbed U1/ALPH/BETA and a digit-transposed TI13 make
    the Newton fixed point a different one-step method: its error does not shrink
    with the step size, while the exact constants converge to round-off."""
    M, K = np.diag([2.0, 3.0, 4.0]), np.diag([1.0, 2.0, 3.0])
    ref = np.exp(-np.array([0.5, 2 / 3, 0.75]))
    errs = {}
    for exact in (0, 1):
        for ms in (0.01, 0.001):
            out = np.zeros(3)
            st = oracle.OrcStats()
            import ctypes as C
            oracle.lib().orc_radau_dense(3, oracle._dp(np.ascontiguousarray(M)), oracle._dp(np.ascontiguousarray(K)),
                                         oracle._dp(np.ones(3)), 1.0, 1e-8, 1e-10, ms, 0.0, 1, exact,
                                         oracle._dp(out), C.byref(st))
            errs[exact, ms] = np.max(np.abs(out / ref - 1))
    assert errs[1, 0.01] < 1e-12 and errs[1, 0.001] < 1e-12
    assert 2e-4 < errs[0, 0.01] < 5e-4 and abs(errs[0, 0.001] / errs[0, 0.01] - 1) < 1e-3


@pytest.mark.parametrize("cmode", [0, 1])
def test_radau_dense_matches_expm(oracle, cmode):
    """Exact constants on a non-symmetric 12 x 12 system vs scipy expm. (The
    reference constants' perturbed method is O(0.1) off on this oscillatory
    system, so only the exact mode is compared with the matrix exponential.)"""
    exact = 1
    rng = np.random.default_rng(7)
    n = 12
    A = rng.standard_normal((n, n))
    M = A @ A.T / n + np.eye(n)
    B = rng.standard_normal((n, n))
    K = B @ B.T / n + 0.3 * (B - B.T) + 0.5 * np.eye(n)  # non-symmetric, stable
    y0 = rng.standard_normal(n)
    y, na, st = _radau(oracle, M, K, y0, 2.0, 1e-9, 1e-12, cmode=cmode, exact=exact)
    ref = sl.expm(-2.0 * np.linalg.solve(M, K)) @ y0
    np.testing.assert_allclose(y, ref, rtol=1e-7, atol=1e-9)


def test_radau_gmres_and_direct_agree_when_tight(oracle):
    rng = np.random.default_rng(3)
    n = 20
    A = rng.standard_normal((n, n))
    M = A @ A.T / n + np.eye(n)
    K = np.diag(np.linspace(0.1, 50.0, n)) + 0.1 * rng.standard_normal((n, n))
    y0 = rng.standard_normal(n)
    a, _, sa = _radau(oracle, M, K, y0, 1.0, 1e-8, 1e-11, cmode=0)
    b, _, sb = _radau(oracle, M, K, y0, 1.0, 1e-8, 1e-11, cmode=1)
    np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------- ADMM NNLS
def test_admm_nnls_matches_scipy_nnls(oracle):
    rng = np.random.default_rng(11)
    n = 15
    A = np.eye(n) + 0.2 * rng.standard_normal((n, n))
    A = (A + A.T) / 2 + n * 0.05 * np.eye(n)
    b = rng.standard_normal(n)
    c = np.zeros(n)
    it = oracle.lib().orc_admm_nnls_dense(n, oracle._dp(np.ascontiguousarray(A)), oracle._dp(b), oracle._dp(c))
    lam = 1e-6  # osqp.mojo: objective 0.5|Ac-b|^2 + 0.5 lambda |c|^2
    ref, _ = so.nnls(np.vstack([A, np.sqrt(lam) * np.eye(n)]), np.concatenate([b, np.zeros(n)]))
    assert 0 < it < 50000
    np.testing.assert_allclose(c, ref, atol=5e-6)
    assert np.all(c >= 0)


# ---------------------------------------------------------------- reference-held prices
def test_readme_european_prices_and_greeks(oracle):
    # README.md:381-391: price(n=38, s_max=150, K=[80,90,100]) -> num_insert 50
    (p, d, g, v), rc = oracle.price_greeks([80.0, 90.0, 100.0], **HESTON, n_s=38, n_v=38, num_insert=50,
                                           s_max=150.0)
    assert rc == 0
    np.testing.assert_allclose(p, [0.966276, 0.280919, 0.078883], atol=6e-7)
    np.testing.assert_allclose(d, [0.175855, 0.060298, 0.018875], atol=6e-7)
    np.testing.assert_allclose(g, [0.017766, 0.009270, 0.002742], atol=6e-7)
    np.testing.assert_allclose(v, [14.057144, 6.197252, 2.283113], atol=6e-6)


def test_readme_down_and_out_prices(oracle):
    # README.md:395-404
    p, st, rc = oracle.price([80.0, 90.0, 100.0], **HESTON, n_s=38, n_v=38, num_insert=50, s_max=150.0,
                             option_type=2, barrier=50.0)
    np.testing.assert_allclose(p, [0.905038, 0.263874, 0.074223], atol=6e-7)


def test_notebook_quickstart_n16(oracle):
    # ipynb cell 8: price(S0=60, V0=0.1, T=0.6, r=0.1, n_s=n_v=16, K=100) with s_max = 3 S0
    (p, d, g, v), rc = oracle.price_greeks([100.0], **HESTON, n_s=16, n_v=16, num_insert=50, s_max=180.0)
    np.testing.assert_allclose([p[0], d[0], g[0], v[0]], [0.199286, 0.017885, 0.003026, -0.443417], atol=6e-7)


def test_notebook_barrier_prices_and_step_counts(oracle):
    # ipynb cell 38 (FPE_T = e^{rT} price), cells 28/32 (71 / 63 stored states)
    K = [65.0, 70.0, 75.0, 80.0, 85.0, 90.0, 95.0, 100.0]
    P = oracle.Pipe(**HESTON, n_s=38, n_v=38, num_insert=251, s_max=180.0, option_type=2, barrier=50.0)
    st = P.solve()
    assert P.n_states == 63
    p = P.price_at(K) * np.exp(0.1 * 0.6)
    np.testing.assert_allclose(p, [4.713256, 2.94113, 1.733301, 0.976814, 0.533807, 0.28632, 0.152251, 0.080859],
                               atol=6e-7)
    E = oracle.Pipe(**HESTON, n_s=38, n_v=38, num_insert=251, s_max=180.0, option_type=8)
    E.solve()
    assert E.n_states == 71


@pytest.mark.parametrize("ni,want", [(50, 0.077259), (75, 0.079342), (100, 0.080683), (150, 0.080849),
                                     (200, 0.080834)])
def test_notebook_num_insert_sweep(oracle, ni, want):
    # ipynb cell 41: DOC B=50, K=100, n=38, price_T = e^{rT} price
    P = oracle.Pipe(**HESTON, n_s=38, n_v=38, num_insert=ni, s_max=180.0, option_type=2, barrier=50.0)
    assert abs(P.price_at([100.0])[0] * np.exp(0.06) - want) < 6e-7


@pytest.mark.parametrize("n,want", [(22, 0.084058), (28, 0.080947), (34, 0.080822), (42, 0.080861)])
def test_notebook_knot_sweep(oracle, n, want):
    # ipynb cell 44 (n = 48 omitted for run time; 38 is covered above)
    P = oracle.Pipe(**HESTON, n_s=n, n_v=n, num_insert=251, s_max=180.0, option_type=2, barrier=50.0)
    assert abs(P.price_at([100.0])[0] * np.exp(0.06) - want) < 6e-7


def test_notebook_coarse_quadrature_prices_vanish(oracle):
    # ipynb cell 41: num_insert 10 -> price_T 1.4093e-27 (ADMM needs ~28000 iterations here)
    P = oracle.Pipe(**HESTON, n_s=38, n_v=38, num_insert=10, s_max=180.0, option_type=2, barrier=50.0)
    assert abs(P.price_at([100.0])[0] * np.exp(0.06)) < 1e-20


def test_readme_in_out_parity_table(oracle):
    # README.md:672-685 (s_max = 150, num_insert 251): vanilla / DOC / DIC to 4 decimals, parity exact
    K = [65.0, 70.0, 75.0, 80.0, 85.0, 90.0, 95.0, 100.0]
    base = dict(HESTON, n_s=38, n_v=38, num_insert=251, s_max=150.0)
    van, _, _ = oracle.price(K, **base, option_type=8)
    doc, _, _ = oracle.price(K, **base, option_type=2, barrier=50.0)
    dic, _, _ = oracle.price(K, **base, option_type=0, barrier=50.0)
    np.testing.assert_allclose(van, [4.6762, 2.8855, 1.6881, 0.9468, 0.5155, 0.2758, 0.1463, 0.0774], atol=5.1e-5)
    np.testing.assert_allclose(doc, [4.4383, 2.7694, 1.6319, 0.9195, 0.5023, 0.2692, 0.1430, 0.0758], atol=5.1e-5)
    np.testing.assert_allclose(dic, [0.2379, 0.1161, 0.0562, 0.0272, 0.0133, 0.0065, 0.0033, 0.0017], atol=5.1e-5)
    np.testing.assert_allclose(dic + doc, van, rtol=0, atol=1e-13)


# ---------------------------------------------------------------- invariants / closed forms
def test_initial_condition_invariants(oracle):
    P = oracle.Pipe(**HESTON, n_s=38, n_v=38, num_insert=251, s_max=180.0)
    q0, it = P.initial_condition()
    assert 0 < it < 50000
    assert np.all(q0 >= 0)
    Bs, Bv = P.basis(0, 0), P.basis(1, 0)
    g = P.grid()
    m = np.kron(Bs.T @ g["sw"], Bv.T @ g["vw"])
    assert abs(m @ q0 - 1.0) < 1e-12


def test_span_mode_config1_near_heston_closed_form(oracle):
    # Cubic 40 x 20 dofs, 4-pt GL per span, v in [0,1.5], IC sigma (0.5, 0.005), rtol 1e-6
    p, st, rc = oracle.price([60.0], **HESTON, disc_mode=1, n_s=40, n_v=20, v_max=1.5, s_max=180.0,
                             ic_sigma_s=0.5, ic_sigma_v=0.005, rtol=1e-6, atol=1e-9, complex_mode=1)
    ref = oracle.lib().orc_heston_call(1.2, 0.05, 0.35, -0.4, 0.1, 0.6, 60.0, 0.1, 60.0)
    assert abs(ref - 7.139000096) < 1e-6  # closed form sanity (Heston 1993)
    assert st["success"] == 1
    assert abs(p[0] / ref - 1.0) < 5e-3
