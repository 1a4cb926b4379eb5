"""Closed forms of Gaussian-state Fock amplitudes, used to pin the oracle's site tensors and
sampler (and the GPU site tensors) against the mathematics rather than against themselves.

A pure zero-mean Gaussian state is |psi> = exp(a^dag B a^dag / 2)|0> / det(Q)^(1/4), with
B the upper-left block of X (I - Q^-1) (X = [[0, I], [I, 0]], Q = the Q-matrix of
make_cov.py:15-31).  Displacing it by alpha gives
D(alpha)|psi> ~ exp(a^dag B a^dag / 2 + gamma . a^dag)|0>,  gamma = alpha - B conj(alpha),
because D(alpha) f(a^dag) D(alpha)^dag = f(a^dag - conj(alpha)) and
D(alpha)|0> = exp(-|alpha|^2/2) exp(alpha . a^dag)|0>.  The Fock amplitude <n|.> is then
F[n] sqrt(n!), where F[n] are the Taylor coefficients of exp(z^T B z / 2 + gamma^T z).  They
follow from d/dz_i F = (gamma_i + sum_j B_ij z_j) F:

    (m_i + 1) F[m + e_i] = gamma_i F[m] + sum_j B_ij F[m - e_j].

This is a different algorithm from the hafnian recursion (MPS_cpu.py:143-187) the site tensors
are built with, and a different route from the Schmidt/Bloch-Messiah decomposition
(MPS_cpu.py:357-372) that defines their basis.
"""
from math import factorial

import numpy as np

from oracle import gbs_oracle as O


def pure_state_B(sq_cov):
    """B of the pure state with covariance ``sq_cov`` (xxpp, hbar = 2), and det(Q)."""
    M = len(sq_cov) // 2
    Q = O.Qmat(sq_cov)
    X = np.block([[np.zeros((M, M)), np.eye(M)], [np.eye(M), np.zeros((M, M))]])
    A = X @ (np.eye(2 * M) - np.linalg.inv(Q))
    return A[:M, :M], np.linalg.det(Q).real


def taylor_coefficients(B, gamma, c):
    """F[n], n in [0, c)^M, of exp(z^T B z / 2 + gamma^T z): mode 0 by the recursion above,
    the slice n_0 = 0 by the same construction on modes 1..M-1."""
    M = len(gamma)
    if M == 0:
        return np.ones((), dtype=np.complex128)
    rest = taylor_coefficients(B[1:, 1:], gamma[1:], c)
    F = np.zeros((c,) + rest.shape, dtype=np.complex128)
    F[0] = rest

    def lowered(X, j):  # X[m - e_j] (zero where m_j = 0)
        Y = np.zeros_like(X)
        src = [slice(None)] * X.ndim
        dst = [slice(None)] * X.ndim
        src[j], dst[j] = slice(0, c - 1), slice(1, c)
        Y[tuple(dst)] = X[tuple(src)]
        return Y

    for k in range(1, c):
        acc = gamma[0] * F[k - 1]
        if k >= 2:
            acc = acc + B[0, 0] * F[k - 2]
        for j in range(1, M):
            acc = acc + B[0, j] * lowered(F[k - 1], j - 1)
        F[k] = acc / k
    return F


def fock_amplitudes(B, gamma, c):
    """F[n] sqrt(n!) on [0, c)^M (unnormalised amplitudes of the state above)."""
    F = taylor_coefficients(B, np.asarray(gamma, dtype=np.complex128), c)
    sf = np.sqrt(np.array([factorial(k) for k in range(c)], dtype=np.float64))
    for ax in range(F.ndim):
        shape = [1] * F.ndim
        shape[ax] = c
        F = F * sf.reshape(shape)
    return F


def mps_amplitudes(Gammas, Lambdas):
    """psi[n_0, ..., n_{M-1}] = Gamma_0[0, a_0, n_0] L_0[a_0] Gamma_1[a_0, a_1, n_1] ...
    Gamma_{M-1}[a_{M-2}, 0, n_{M-1}]: the state the sampler walks (sampling_cpu.py:57, 75-81)."""
    v = np.transpose(np.asarray(Gammas[0])[0], (1, 0))           # [n_0, a_0]
    for k in range(1, len(Gammas)):
        v = v * np.asarray(Lambdas[k - 1])
        v = np.tensordot(v, np.asarray(Gammas[k]), axes=([v.ndim - 1], [0]))  # [..., a_k, n_k]
        v = np.moveaxis(v, -1, -2)
    return v[..., 0]


def right_canonical_sums(G, lam_next):
    """s_b = sum_{a, j} |Gamma[b, a, j]|^2 Lambda_k[a]^2 for every row b of one site tensor.
    Gamma_k[b, a, j] Lambda_k[a] = <j, R_a(k)| R_b(k-1)> is a matrix element of a unitary, so
    s_b <= 1, with equality up to the weight truncation drops."""
    G = np.asarray(G)
    w = np.asarray(lam_next, dtype=np.float64) ** 2
    return (np.abs(G) ** 2 * w[None, :, None]).sum(axis=(1, 2))


def left_canonical_sums(G, lam_prev):
    """t_a = sum_{b, j} Lambda_{k-1}[b]^2 |Gamma[b, a, j]|^2 for every column a (the norm of the
    left Schmidt vector a of cut k; <= 1 with equality up to truncation)."""
    G = np.asarray(G)
    w = np.asarray(lam_prev, dtype=np.float64) ** 2
    return (np.abs(G) ** 2 * w[:, None, None]).sum(axis=(0, 2))


def displaced_probabilities(B, alpha, c):
    """Unnormalised Fock probabilities of D(alpha)|psi> on [0, c)^M."""
    alpha = np.asarray(alpha, dtype=np.complex128)
    return np.abs(fock_amplitudes(B, alpha - B @ np.conj(alpha), c)) ** 2


def conditional(P, prefix, dd):
    """p(n_k = j | n_0..n_{k-1} = prefix), j < dd, normalised over j < dd, from a probability
    grid P (k = len(prefix); the modes after k are summed over the grid)."""
    c = P.shape[0]
    pk = P[tuple(int(x) for x in prefix)].reshape(c, -1).sum(axis=1)[:dd]
    return pk / pk.sum()
