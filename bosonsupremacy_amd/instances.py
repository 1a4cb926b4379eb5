"""Synthetic lossy-GBS instances (host, float64).

An instance is the pair of files the reference's pipeline starts from: ``cov.npy`` (the
lossy output covariance, xxpp ordering, hbar = 2) and ``sq_cov.npy`` (its pure part).  The
reference builds ``cov`` from experiment data (make_cov.py:47-60, ``cov_from_T`` with the
Q-matrix helpers of make_cov.py:15-45) and obtains ``sq_cov`` from a semidefinite program
(get_decomposition.py:16-30: minimise tr X subject to X being a valid covariance and
cov - X >= 0).  Here the interferometer is a seeded Haar-random unitary and the transmission
is uniform, which makes both steps closed-form:

* T = sqrt(eta) * U with U unitary.  make_cov.py's Q-matrix convention (aidaj = <a_i^dag a_j>
  transformed as T N T^dag) acts on the quadratures through V = conj(U), so with the
  orthogonal symplectic O = [[Re V, -Im V], [Im V, Re V]] the output covariance is
  O diag(a, b) O^T -- equal to cov_from_T(r, sqrt(eta) U) -- with, per input mode i,
  a_i = eta e^{2 r_i} + 1 - eta + 2 nth and b_i = eta e^{-2 r_i} + 1 - eta + 2 nth;
* the minimum-trace pure covariance below it is O diag(1/b', b') O^T with
  b' = min(1, b) -- per mode, t + 1/t is decreasing on (0, 1], so the squeezed quadrature is
  kept and the anti-squeezed one is lowered to the uncertainty bound; the difference
  W = cov - sq_cov = O diag(a - 1/b', b - b') O^T is positive semidefinite.
"""
from dataclasses import dataclass, field

import numpy as np


def haar_unitary(M, seed):
    """Haar-random M x M unitary: QR of a seeded complex Gaussian with the phase fix
    U = Q diag(R_ii / |R_ii|) (Mezzadri 2007)."""
    rng = np.random.default_rng(seed)
    Z = (rng.normal(size=(M, M)) + 1j * rng.normal(size=(M, M))) / np.sqrt(2.0)
    Q, R = np.linalg.qr(Z)
    ph = np.diag(R) / np.abs(np.diag(R))
    return Q * ph[np.newaxis, :]


def _orthogonal_of(U):
    """Quadrature (xxpp) orthogonal matrix of the interferometer U in make_cov.py's
    convention, which applies conj(U) to the mode operators."""
    V = np.conj(U)
    return np.block([[V.real, -V.imag], [V.imag, V.real]])


def lossy_covariance(r, U, eta, nth=0.0):
    """Output covariance of squeezed vacua r through sqrt(eta) U plus thermal noise nth
    (xxpp, hbar = 2): O [eta diag(e^{2r}, e^{-2r}) + (1 - eta + 2 nth) I] O^T."""
    r = np.asarray(r, dtype=np.float64)
    O = _orthogonal_of(U)
    diag = np.concatenate([eta * np.exp(2 * r), eta * np.exp(-2 * r)]) + (1.0 - eta + 2.0 * nth)
    cov = (O * diag[np.newaxis, :]) @ O.T
    return 0.5 * (cov + cov.T)


def pure_part(r, U, eta, nth=0.0):
    """Closed-form minimum-trace pure covariance below ``lossy_covariance``."""
    r = np.asarray(r, dtype=np.float64)
    O = _orthogonal_of(U)
    b = np.minimum(1.0, eta * np.exp(-2 * r) + 1.0 - eta + 2.0 * nth)
    diag = np.concatenate([1.0 / b, b])
    sq = (O * diag[np.newaxis, :]) @ O.T
    return 0.5 * (sq + sq.T)


@dataclass
class Instance:
    """One synthetic GBS instance and the MPS/sampling parameters it is run with."""
    name: str
    M: int
    d: int
    chi: int
    dd: int
    n_samples: int
    eta: float
    nth: float
    r: np.ndarray
    U: np.ndarray
    cov: np.ndarray = field(repr=False)
    sq_cov: np.ndarray = field(repr=False)
    seed: int = 0


def make_instance(name, M, r_lo, r_hi, eta, d, chi, dd, n_samples, seed=0, nth=0.0):
    U = haar_unitary(M, seed)
    r = np.random.default_rng(seed + 1).uniform(r_lo, r_hi, size=M)
    cov = lossy_covariance(r, U, eta, nth)
    sq_cov = pure_part(r, U, eta, nth)
    return Instance(name=name, M=M, d=d, chi=chi, dd=dd, n_samples=n_samples, eta=eta, nth=nth,
                    r=r, U=U, cov=cov, sq_cov=sq_cov, seed=seed)


# The workloads, with their sizes.  dd (the photon cutoff after displacement) is not part of
# the instance definitions; 10 is the reference README's default (Readme.md:176).
WORKLOADS = {
    "m8_d4_chi64": dict(M=8, r_lo=0.5, r_hi=1.0, eta=0.3, d=4, chi=64, dd=10, n_samples=10 ** 4),
    "m32_d6_chi512": dict(M=32, r_lo=0.6, r_hi=1.2, eta=0.5, nth=0.05, d=6, chi=512, dd=10,
                          n_samples=10 ** 5),
    "m64_d8_chi2000": dict(M=64, r_lo=0.8, r_hi=1.4, eta=0.45, d=8, chi=2000, dd=16,
                           n_samples=10 ** 5),
    "m144_d10_chi10000": dict(M=144, r_lo=1.0, r_hi=1.6, eta=0.45, d=10, chi=10000, dd=20,
                              n_samples=10 ** 6),
}


def workload(name, seed=0, **overrides):
    params = dict(WORKLOADS[name])
    params.update(overrides)
    return make_instance(name, seed=seed, **params)


def random_batch(n_instances=256, seed=1000, n_samples=10 ** 4):
    """The batched workload: independent instances with M ~ U{16..96}, d ~ U{4..8},
    chi ~ U{128..1024}, r_i ~ U(0.5, 1.5), eta ~ U(0.3, 0.7), instance i with its own Haar
    unitary (seed + i); dd = 2 d (the photon cutoff after displacement)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_instances):
        M = int(rng.integers(16, 97))
        d = int(rng.integers(4, 9))
        chi = int(rng.integers(128, 1025))
        eta = float(rng.uniform(0.3, 0.7))
        out.append(make_instance(f"batch_{i}", M, 0.5, 1.5, eta, d, chi, 2 * d, n_samples, seed=seed + i))
    return out
