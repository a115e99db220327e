"""Batched Levenberg-Marquardt (the calibration caller's optimiser) on problems
with known minimisers, checked on the CPU with closed-form residuals."""
import numpy as np

from fpe_option_amd import calibration as cal


def test_lm_batch_recovers_exponential_fits():
    # r_i(x) = x0 exp(x1 t_i) + x2 - y_i for three independent problems
    t = np.linspace(0.0, 1.0, 12)
    truth = np.array([[2.0, -1.5, 0.3], [0.5, 0.8, -1.0], [1.2, 0.5, 0.1]])
    Y = truth[:, 0:1] * np.exp(truth[:, 1:2] * t[None]) + truth[:, 2:3]

    def residuals(idx, pts):
        return pts[..., 0:1] * np.exp(pts[..., 1:2] * t) + pts[..., 2:3] - Y[idx][:, None, :]

    X, iters, conv = cal.levenberg_marquardt_batch(residuals, truth + 0.2, max_iter=60, tol=1e-10)
    np.testing.assert_allclose(X, truth, atol=1e-6)
    assert conv.all() and np.all(iters >= 1)


def test_lm_batch_rosenbrock_as_least_squares():
    # r = (10 (x1 - x0^2), 1 - x0): the reference LM's fixed lambda schedule reaches (1, 1)
    def residuals(idx, pts):
        return np.stack([10.0 * (pts[..., 1] - pts[..., 0] ** 2), 1.0 - pts[..., 0]], axis=-1)

    X, iters, conv = cal.levenberg_marquardt_batch(residuals, np.array([[-1.2, 1.0], [0.5, 0.5]]), max_iter=200,
                                                   tol=1e-9)
    np.testing.assert_allclose(X, 1.0, atol=1e-6)


def test_clamps_follow_the_reference():
    X = cal.clamp([[-1.0, 0.0, 0.0, -2.0, -5.0], [3.0, 0.1, 0.5, 1.5, 0.2]])
    np.testing.assert_array_equal(X[0], [1e-4, 1e-5, 1e-4, -0.999, 1e-6])
    np.testing.assert_array_equal(X[1], [3.0, 0.1, 0.5, 0.999, 0.2])
