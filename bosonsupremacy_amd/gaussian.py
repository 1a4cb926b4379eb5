"""Host FP64 Gaussian-state linear algebra for the MPS path.

These are O(M^3) per cut/site and run once per cut/site on the host; the device work
(Schmidt top-chi, hafnians, sampling) is in libbsmps.

The Williamson and Bloch-Messiah decompositions are not unique: the order and signs of the
normal modes (the "gauge") are whatever the particular sequence of LAPACK calls returns,
and the site tensors depend on that gauge.  To be a drop-in for the reference's files
(Gamma_k is only meaningful together with the S_k it was built from), these routines issue
the same LAPACK calls, on the same operands, in the same order as the reference:
williamson = kron_cpu.py:46-89, blochmessiah = MPS_cpu.py:189-229,
get_U2_sq_U1 = MPS_cpu.py:357-372, get_Sigma = MPS_cpu.py:305-312.
"""
import functools

import numpy as np
from scipy.linalg import schur, sqrtm, svd

try:
    from threadpoolctl import ThreadpoolController
    _BLAS = ThreadpoolController()
except Exception:  # pragma: no cover - threadpoolctl ships with this environment
    _BLAS = None


def single_thread_blas(fn):
    """Run ``fn`` with single-threaded BLAS/LAPACK.

    OpenBLAS results for these matrix sizes (up to a few hundred) depend on how many threads
    it splits a call over, so the same call gives different last bits on hosts with
    different core counts.  Pinning the decompositions to one thread makes them
    reproducible on any host, in any process (the host worker pool included)."""
    @functools.wraps(fn)
    def wrapper(*args, **kwargs):
        if _BLAS is None:
            return fn(*args, **kwargs)
        with _BLAS.limit(limits=1, user_api="blas"):
            return fn(*args, **kwargs)
    return wrapper


def sympmat(n):
    """Symplectic form [[0, I], [-I, 0]] of n modes (kron_cpu.py:23-27)."""
    eye = np.identity(n)
    zero = np.zeros((n, n))
    return np.block([[zero, eye], [-eye, zero]])


def xxpp_order(n):
    """Index permutation taking xpxp ordering to xxpp (kron_cpu.py:29-44)."""
    return np.arange(2 * n).reshape(-1, 2).T.flatten()


def xpxp_to_xxpp(S):
    shape = S.shape
    if shape[0] % 2:
        raise ValueError("The input array is not even-dimensional")
    ind = xxpp_order(shape[0] // 2)
    if S.ndim == 2:
        if shape[0] != shape[1]:
            raise ValueError("The input matrix is not square")
        return S[:, ind][ind]
    return S[ind]


@single_thread_blas
def williamson(V, tol=1e-11):
    """Williamson normal form of a real symmetric positive-definite V (xxpp).

    Returns (Db, S) with Db = diag(nu, nu) and S = inv(S_w).T, where V = S_w Db S_w^T; the
    errors raised match kron_cpu.py:49-66.
    """
    V = np.asarray(V)
    rows, cols = V.shape
    if rows != cols:
        raise ValueError("The input matrix is not square")
    if np.linalg.norm(V - np.transpose(V)) >= tol:
        raise ValueError("The input matrix is not symmetric")
    if rows % 2:
        raise ValueError("The input matrix must have an even number of rows/columns")
    n = rows // 2
    if np.any(np.linalg.eigvalsh(V) <= 0):
        raise ValueError("Input matrix is not positive definite")
    Mm12 = sqrtm(np.linalg.inv(V)).real
    s1, K = schur(Mm12 @ sympmat(n) @ Mm12)
    # orient every 2x2 Schur block so that its (0, 1) entry is positive
    blocks = [np.identity(2) if s1[2 * i, 2 * i + 1] > 0 else np.array([[0, 1], [1, 0]]) for i in range(n)]
    p = np.zeros((2 * n, 2 * n))
    for i, b in enumerate(blocks):
        p[2 * i:2 * i + 2, 2 * i:2 * i + 2] = b
    Kt = K @ p
    dd = xpxp_to_xxpp(p @ s1 @ p)
    Ktt = Kt[:, xpxp_to_xxpp(np.arange(2 * n))]
    inv_nu = [1 / dd[i, i + n] for i in range(n)]
    Db = np.diag(inv_nu + inv_nu)
    S = Mm12 @ Ktt @ sqrtm(Db)
    return Db, np.linalg.inv(S).T


def thermal_photons(nth, cutoff=20):
    """Geometric photon-number distribution of a thermal mode (kron_cpu.py:92-93)."""
    return 1 / (nth + 1) * (nth / (nth + 1)) ** np.arange(cutoff)


@single_thread_blas
def cut_inputs(sq_cov, L, cutoff):
    """Williamson of modes L..M-1 of sq_cov and their thermal photon distributions
    [M-L, cutoff] (kron_cpu.py:96-106)."""
    M = len(sq_cov) // 2
    mode = np.arange(L, M)
    modes = np.append(mode, mode + M)
    D, S = williamson(sq_cov[np.ix_(modes, modes)])
    nth = (np.diag(D) - 1) / 2
    nth[nth < 0] = 0
    n = M - L
    tp = np.stack([np.array(thermal_photons(nth[i], cutoff)) for i in range(n)])
    return tp, S


def photon_distribution(V, nmax):
    """Photon-number distribution p(0..nmax-1) of a zero-mean single-mode Gaussian state with
    covariance V (2 x 2, (x, p) order, hbar = 2).

    p(n) = haf(A_(n,n)) / (n! sqrt(det sigma_Q)) with sigma_Q = (V + I) / 2 in the (a, a^+)
    basis and A = X (I - sigma_Q^-1) (the Gaussian-state Fock-amplitude formula of Hamilton et
    al., PRL 119, 170501 (2017)).  For one mode, A = [[a, b], [b, a*]] and the hafnian of the
    matrix with each row repeated n times is a sum over the number k of within-row pairs:
    haf = sum_k (C(n, 2k) (2k-1)!!)^2 (n-2k)! |a|^(2k) b^(n-2k)."""
    from math import comb, factorial
    R = np.array([[1, 1j], [1, -1j]]) / np.sqrt(2)
    sQ = R @ ((np.asarray(V, dtype=float) + np.eye(2)) / 2) @ R.conj().T
    A = np.array([[0, 1], [1, 0]]) @ (np.eye(2) - np.linalg.inv(sQ))
    a2, b = abs(A[0, 0]) ** 2, A[0, 1]
    norm = 1.0 / np.sqrt(np.linalg.det(sQ).real)
    p = np.zeros(nmax)
    for n in range(nmax):
        h = 0j
        for k in range(n // 2 + 1):
            dfact = 1
            for i in range(1, 2 * k, 2):
                dfact *= i
            # exact integer ratio, rounded once: (C(n,2k) (2k-1)!!)^2 (n-2k)! / n!
            coef = (comb(n, 2 * k) * dfact) ** 2 * factorial(n - 2 * k) / factorial(n)
            h += coef * a2 ** k * b ** (n - 2 * k)
        p[n] = h.real * norm
    return p


def photon_tail(V, d):
    """P(n >= d) for the single-mode state of covariance V (see photon_distribution)."""
    if d <= 0:
        return 1.0
    return float(max(0.0, 1.0 - np.sum(photon_distribution(V, d))))


@single_thread_blas
def blochmessiah(S):
    """Bloch-Messiah decomposition S = uff @ dff @ vff of a real symplectic S (xxpp)."""
    N = S.shape[0]
    h = N // 2
    eye = np.eye(h)
    R = (1 / np.sqrt(2)) * np.block([[eye, 1j * eye], [eye, -1j * eye]])
    Rh = np.conjugate(R).T
    Sc = R @ S @ Rh
    u1, d1, v1 = svd(Sc, lapack_driver='gesvd')           # polar decomposition via SVD
    Sig = u1 @ np.diag(d1) @ np.conjugate(u1).T
    Unitary = u1 @ v1
    alpha = Unitary[0:h, 0:h]
    beta = Sig[0:h, h:N]
    u2, d2, v2 = np.linalg.svd(beta)                       # Takagi factorisation of beta
    sval = np.arcsinh(d2)
    tb = u2 @ sqrtm(np.conjugate(u2).T @ (v2.T))
    z = 0 * tb
    uf = np.block([[tb, z], [z, np.conjugate(tb)]])
    va = np.conjugate(tb).T @ alpha
    vf = np.block([[va, z], [z, np.conjugate(va)]])
    ch, sh = np.diag(np.cosh(sval)), np.diag(np.sinh(sval))
    df = np.block([[ch, sh], [sh, ch]])
    uff = np.real_if_close(Rh @ uf @ R)
    vff = np.real_if_close(Rh @ vf @ R)
    dff = np.real_if_close(Rh @ df @ R)
    return uff, dff, vff


@single_thread_blas
def get_U2_sq_U1(S_l, S_r):
    """Site unitary decomposition: S = (1 (+) S_l^-1) S_r = U2 . sq . U1 (complex form)."""
    M = len(S_r) // 2
    mode = np.arange(M - 1) + 1
    modes = np.append(mode, mode + M)
    S_l2_inv = np.eye(2 * M, dtype=float)
    S_l2_inv[np.ix_(modes, modes)] = np.linalg.inv(S_l)
    S2, SQ, S1 = blochmessiah(S_l2_inv @ S_r)
    U2 = S2[:M, :M] - 1j * S2[:M, M:]
    U1 = S1[:M, :M] - 1j * S1[:M, M:]
    sq = np.log(np.diag(SQ)[:M])
    return U2, sq, U1


@single_thread_blas
def get_Sigma(U2, sq, U1):
    """Fock-amplitude kernel of the site's Gaussian unitary, complex128 [2M', 2M']."""
    M = len(sq)
    th, sech = np.diag(np.tanh(sq)), np.diag(1 / np.cosh(sq))
    Sigma = np.zeros((2 * M, 2 * M), dtype=np.complex128)
    Sigma[:M, :M] = U2 @ th @ U2.T
    Sigma[:M, M:] = U2 @ sech @ U1
    Sigma[M:, :M] = U1.T @ sech @ U2.T
    Sigma[M:, M:] = -U1.T @ th @ U1
    return Sigma


@single_thread_blas
def site_kernel(k, M, sq_cov, S_cuts, S_full=None):
    """(Sigma, Z) of site k (MPS_cpu.py:395, 409-414, 439-442, 481-483).

    S_cuts[c] is the Williamson symplectic of cut c (modes c+1..M-1); S_full that of the
    whole state (computed here when not given)."""
    if k == 0:
        S_r = S_full if S_full is not None else williamson(sq_cov)[1]
        S_l = S_cuts[0]
    elif k == M - 1:
        S_r = S_cuts[k - 1]
        S_l = np.zeros((0, 0))
    else:
        S_r = S_cuts[k - 1]
        S_l = S_cuts[k]
    U2, sq, U1 = get_U2_sq_U1(S_l, S_r)
    return get_Sigma(U2, sq, U1), float(np.sqrt(np.prod(np.cosh(sq))))
