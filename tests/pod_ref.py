"""FP64 numpy references of the surrogate's POD steps (test infrastructure
only): the library's GPU POD (fpe_pod_fit) and SVD (fpe_svd_jacobi) are
checked against these, and these against closed forms."""
import numpy as np


def pod(Q, k=64):
    """Mean and the k leading Euclidean POD modes of snapshot rows Q."""
    mu = Q.mean(axis=0)
    X = Q - mu
    ev, V = np.linalg.eigh(X.T @ X)
    order = np.argsort(ev)[::-1][:k]
    return mu, V[:, order], ev[order]


def fit_pod(Q, marg, k=64):
    """POD of snapshot rows Q in the pricing inner product <q1, q2> =
    (marg q1) . (marg q2): mu, Phi (n_dof x k), P = marg Phi (orthonormal
    columns), w0 = marg mu and the singular values of (Q - mu) marg^T."""
    mu = Q.mean(axis=0)
    X = Q - mu
    Wc = X @ marg.T
    U, sv, Vt = np.linalg.svd(Wc, full_matrices=False)
    k = min(k, len(sv))
    Phi = X.T @ (U[:, :k] / sv[:k])
    return mu, Phi, Vt[:k].T.copy(), marg @ mu, sv


def svd(A):
    """(singular values, V^T) with numpy's LAPACK SVD."""
    _, sv, Vt = np.linalg.svd(A, full_matrices=False)
    return sv, Vt


def align_columns(X, ref):
    """X with each column's sign flipped to best match ref's column."""
    s = np.sign(np.sum(X * ref, axis=0))
    s[s == 0] = 1.0
    return X * s
