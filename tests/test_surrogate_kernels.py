"""Surrogate kernels (BF16 MFMA forward/backward, Adam, loss, POD pricing)
against the FP64 numpy reference in mlp_ref.py, which applies the same BF16
rounding points: each layer kernel is checked on identical inputs (FP32
accumulation vs FP64: 1e-5 relative), the whole network's gradients end to end."""
import numpy as np
import pytest

import mlp_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev(engine):
    from fpe_option_amd import surrogate as S
    return S


def _t(S, a, dtype=None):
    return S.DeviceArray.from_numpy(np.asarray(a, dtype or np.float32))


def _close(g, r, tol):
    g, r = np.asarray(g, np.float64), np.asarray(r, np.float64)
    scale = np.max(np.abs(r)) + 1e-30
    err = np.max(np.abs(g - r)) / scale
    assert err <= tol, f"max err / max |ref| = {err:.3e}"


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("M,K,N,act", [(37, 7, 50, 1), (128, 256, 256, 1), (70, 256, 64, 0), (1, 3, 5, 1)])
def test_linear_forward_and_backward(dev, M, K, N, act, fp32):
    S = dev
    rng = np.random.default_rng(M * 1000 + K)
    X = rng.standard_normal((M, K)).astype(np.float32)
    W = (rng.standard_normal((N, K)) / np.sqrt(K)).astype(np.float32)
    b = rng.standard_normal(N).astype(np.float32) * 0.1
    net = S.DeviceMLP([W], [b], fp32)
    lib, st = net.lib, net.stream
    fwd, bwd = (lib.fpe_mlp_forward_f32, lib.fpe_mlp_backward_f32) if fp32 else (lib.fpe_mlp_forward,
                                                                                  lib.fpe_mlp_backward)
    Wl = net.W[0] if fp32 else net.Wb[0]
    Xd, Z, A = _t(S, X), S.DeviceArray((M, N), np.float32), S.DeviceArray((M, N), np.float32)
    assert fwd(Xd.c, M, K, Wl.c, net.b[0].c, N, act, Z.c, A.c, st) == 0
    Zr, Ar = R.linear_forward(X, W, b.astype(np.float64), act, fp32)
    _close(Z.numpy(), Zr, 1e-5)
    _close(A.numpy(), Ar, 1e-5)
    dZ = rng.standard_normal((M, N)).astype(np.float32)
    Zp = rng.standard_normal((M, K)).astype(np.float32)
    dW, db, dX = S.DeviceArray((N, K), np.float32), S.DeviceArray((N,), np.float32), S.DeviceArray((M, K), np.float32)
    dZd, Zpd = _t(S, dZ), _t(S, Zp)
    assert bwd(dZd.c, Xd.c, Wl.c, M, N, K, dW.c, db.c, dX.c, Zpd.c, st) == 0
    dWr, dbr, dXr = R.linear_backward(dZ, X, W, Zp.astype(np.float64), fp32)
    _close(dW.numpy(), dWr, 1e-5)
    _close(db.numpy(), dbr, 1e-5)
    _close(dX.numpy(), dXr, 1e-5)


@pytest.mark.parametrize("fp32", [False, True])
def test_network_gradients_match_reference(dev, fp32):
    S = dev
    rng = np.random.default_rng(7)
    widths = (7, 32, 32, 16)
    Ws = [(rng.standard_normal((b, a)) / np.sqrt(a)).astype(np.float32) for a, b in zip(widths[:-1], widths[1:])]
    bs = [(0.1 * rng.standard_normal(b)).astype(np.float32) for b in widths[1:]]
    net = S.DeviceMLP(Ws, bs, fp32)
    tol = 1e-5 if fp32 else 1e-3
    X = rng.uniform(-1, 1, (64, 7)).astype(np.float32)
    T = rng.standard_normal((64, 16)).astype(np.float32)
    Td = _t(S, T)
    part = S.DeviceArray((1024,), np.float32)
    grads, loss_gpu = S.mse_step(net, _t(S, X), Td, None, part)
    out = net.forward(_t(S, X))
    # reference: forward with the same rounding, MSE gradient, backward
    h, cache = X.astype(np.float64), []
    for i, (W, b) in enumerate(zip(Ws, bs)):
        Z, A = R.linear_forward(h.astype(np.float32), W, b.astype(np.float64), i < len(Ws) - 1, fp32)
        cache.append((h, Z))
        h = A
    _close(out.numpy(), h, tol)
    loss = np.mean((h - T) ** 2)
    assert abs(loss_gpu - loss) <= tol * loss
    g = 2.0 * (h - T) / h.size
    for i in reversed(range(len(Ws))):
        Zp = cache[i - 1][1] if i > 0 else None
        dWr, dbr, dX = R.linear_backward(g.astype(np.float32), cache[i][0].astype(np.float32), Ws[i], Zp, fp32)
        _close(grads[i][0].numpy(), dWr, 2 * tol)
        _close(grads[i][1].numpy(), dbr, 2 * tol)
        g = dX


def test_adam_step_matches_formula(dev):
    S = dev
    rng = np.random.default_rng(3)
    W = rng.standard_normal(1000).astype(np.float32)
    net = S.DeviceMLP([W.reshape(10, 100)], [np.zeros(10, np.float32)])
    G = rng.standard_normal(1000).astype(np.float32)
    m, v = _t(S, np.zeros(1000)), _t(S, np.zeros(1000))
    Gd = _t(S, G)
    w_ref, m_ref, v_ref = W.astype(np.float64), np.zeros(1000), np.zeros(1000)
    for t in (1, 2, 3):
        assert net.lib.fpe_mlp_adam(net.W[0].c, net.Wb[0].c, Gd.c, m.c, v.c, 1000, 1e-2, 0.9, 0.999, 1e-8, t,
                                    net.stream) == 0
        m_ref = 0.9 * m_ref + 0.1 * G
        v_ref = 0.999 * v_ref + 0.001 * G.astype(np.float64) ** 2
        w_ref = w_ref - 1e-2 * (m_ref / (1 - 0.9 ** t)) / (np.sqrt(v_ref / (1 - 0.999 ** t)) + 1e-8)
    _close(net.W[0].numpy().ravel(), w_ref, 1e-5)
    wb = net.Wb[0].numpy().astype(np.uint32) << 16
    np.testing.assert_array_equal(wb.view(np.float32).astype(np.float64).ravel(), R.bf16(net.W[0].numpy()).ravel())


def test_surrogate_price_kernel(dev):
    S = dev
    rng = np.random.default_rng(11)
    n, k, nq = 203, 64, 316
    Cn = rng.standard_normal((n, k)).astype(np.float32)
    cs, cb = rng.uniform(0.5, 2, k).astype(np.float32), rng.standard_normal(k).astype(np.float32)
    P = (rng.standard_normal((nq, k)) * 1e-3).astype(np.float32)
    w0 = rng.uniform(0, 1e-2, nq).astype(np.float32)
    sq = np.linspace(0.1, 179.9, nq).astype(np.float32)
    K = rng.uniform(45, 75, n)
    ty = np.where(rng.uniform(size=n) < 0.5, 8, 9).astype(np.int32)
    r, T = rng.uniform(0, 0.1, n), rng.uniform(0.1, 2, n)
    out = S.DeviceArray((n,), np.float64)
    lib = S.L.lib()
    args = [_t(S, Cn), _t(S, cs), _t(S, cb), _t(S, np.ascontiguousarray(P.T)), _t(S, w0), _t(S, sq),
            _t(S, K, np.float64), _t(S, ty, np.int32), _t(S, r, np.float64), _t(S, T, np.float64)]
    c = [a.c for a in args]
    assert lib.fpe_surrogate_price(c[0], n, k, c[1], c[2], c[3], c[4], c[5], nq, c[6], c[7], c[8], c[9], out.c,
                                   None) == 0
    ref = R.surrogate_price(Cn, cs.astype(np.float64), cb.astype(np.float64), P.astype(np.float64),
                            w0.astype(np.float64), sq.astype(np.float64), K, ty, r, T)
    _close(out.numpy(), ref, 1e-5)


def test_gather_rows(dev):
    S = dev
    src = np.arange(40, dtype=np.float32).reshape(10, 4)
    idx = np.array([3, 0, 9, 3], np.int32)
    dst = S.DeviceArray((4, 4), np.float32)
    assert S.L.lib().fpe_mlp_gather_rows(_t(S, src).c, _t(S, idx, np.int32).c, 4, 4, dst.c, None) == 0
    np.testing.assert_array_equal(dst.numpy(), src[idx])
