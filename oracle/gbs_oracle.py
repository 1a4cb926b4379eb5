"""FP64 NumPy restatement of the reference's CPU path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this module.

It follows, line for line, the reference scripts (paths relative to the reference repo):

* kron_cpu.py     -- Schmidt spectrum per cut (sympmat, xpxp_to_xxpp, williamson,
                     thermal_photons, get_cumsum_kron, __main__ loop);
* MPS_cpu.py      -- Gamma/Lambda construction (Sigma_select, push_to_end, hafnian,
                     recursive_hafnian, solve, blochmessiah, get_Sigma, get_target, A_elem,
                     get_U2_sq_U1, and the three site branches of __main__);
* sampling_cpu.py -- photon-number sampling (sampling, destroy, batch_displaces,
                     batch_mu_to_alpha, __main__ set-up of W);
* make_cov.py     -- Qmat / Covmat / cov_from_T, used only to build synthetic instances.

Deliberate deviations from the reference (each needed for an FP64 oracle):

* complex64 -> complex128 (MPS_cpu.py:17, 123, 126, 337, 351, 397, 488; sampling_cpu.py:55,
  77, 121) and float32 Lambda -> float64 (MPS_cpu.py:398, sampling_cpu.py:172);
* kron ``res`` is returned as float64, not cast to float16 (kron_cpu.py:127);
* ``np.product`` -> ``np.prod`` (MPS_cpu.py:413, 437, 468, 479; removed in NumPy 2);
* the global unseeded RNG (sampling_cpu.py:37, 52, 92) is replaced by explicit arrays of
  normals and uniforms passed in by the caller;
* a sampled photon number is clamped to dd-1: sampling_cpu.py:103-104 can return dd when
  rounding leaves cdf[dd-1] below the uniform, which makes the next site NaN;
* when W = cov - sq_cov is singular to working precision the regularised Cholesky of
  sampling_cpu.py:168-170 fails; an eigendecomposition factor is used instead.

Pinning (what checks each function against something other than itself):

* hafnian / recursive_hafnian / solve: brute-force perfect-matching sums (n <= 10), the
  closed forms haf(J_2n) = (2n-1)!!, haf of block-diagonal = product, haf([[0,B],[B^T,0]])
  = perm(B), and bit-exact golden vectors written by tests/golden/make_golden.py from the
  reference's own ``hafnian`` (imported in this container, never shipped);
* get_target / push_to_end / target assembly: multiset identity against the labels;
* get_cumsum_kron: brute-force enumeration of all d^n label products on tiny cuts, and
  golden vectors from the reference's ``get_cumsum_kron``;
* williamson / blochmessiah / get_U2_sq_U1 / get_Sigma: symplectic invariants
  (S Omega S^T = Omega, V = S Db S^T, unitarity) and golden vectors from the reference's
  functions (the gauge -- mode order and signs -- is only pinned by the latter);
* site tensors (build_site, build_mps): tests/test_oracle_pins.py -- the contracted MPS of a
  small weakly squeezed instance equals, amplitude by amplitude, the Taylor-coefficient form
  of the pure Gaussian state exp(a^dag B a^dag / 2)|0> / det(Q)^(1/4) (within the truncated
  weight, 2.6e-6); every site tensor satisfies the right- and left-canonical identities
  (sums <= 1, weighted deficits ~1e-6); a dropped sqrt(j!) or a transposed Gamma fails both;
* sampling: the same file -- every conditional distribution drawn from equals that of the
  displaced pure state D(alpha)|psi> given the sample's alpha and earlier draws (to 1e-6; a
  transposed displacement matrix fails); plus closed-form first-order marginals
  (tr V_ii - 2)/4 and the spectral form of the displacement exponential.
"""
import numpy as np
from scipy.special import factorial
from scipy.linalg import sqrtm, svd, block_diag, schur, expm

# ---------------------------------------------------------------------------------------
# Gaussian linear algebra (kron_cpu.py:23-93, MPS_cpu.py:49-115, 189-229, 305-372)
# ---------------------------------------------------------------------------------------


try:
    from threadpoolctl import ThreadpoolController as _TC
    _BLAS = _TC()
except Exception:  # pragma: no cover
    _BLAS = None


def _single_thread_blas(fn):
    """Single-threaded BLAS for the host decompositions: OpenBLAS results at these sizes
    depend on its thread count; one thread is the reproducible configuration (the product
    code pins the same)."""
    import functools

    @functools.wraps(fn)
    def wrapper(*args, **kwargs):
        if _BLAS is None:
            return fn(*args, **kwargs)
        with _BLAS.limit(limits=1, user_api="blas"):
            return fn(*args, **kwargs)
    return wrapper


def sympmat(N, dtype=np.float64):
    """kron_cpu.py:23-27."""
    I = np.identity(N, dtype=dtype)
    O = np.zeros_like(I, dtype=dtype)
    return np.block([[O, I], [-I, O]])


def xpxp_to_xxpp(S):
    """kron_cpu.py:29-44."""
    shape = S.shape
    n = shape[0]
    if n % 2 != 0:
        raise ValueError("The input array is not even-dimensional")
    n = n // 2
    ind = np.arange(2 * n).reshape(-1, 2).T.flatten()
    if len(shape) == 2:
        if shape[0] != shape[1]:
            raise ValueError("The input matrix is not square")
        return S[:, ind][ind]
    return S[ind]


@_single_thread_blas
def williamson(V, tol=1e-11):
    """kron_cpu.py:46-89 (identical copy at MPS_cpu.py:72-115)."""
    (n, m) = V.shape
    if n != m:
        raise ValueError("The input matrix is not square")
    diffn = np.linalg.norm(V - np.transpose(V))
    if diffn >= tol:
        raise ValueError("The input matrix is not symmetric")
    if n % 2 != 0:
        raise ValueError("The input matrix must have an even number of rows/columns")
    n = n // 2
    omega = sympmat(n)
    vals = np.linalg.eigvalsh(V)
    for val in vals:
        if val <= 0:
            raise ValueError("Input matrix is not positive definite")
    Mm12 = sqrtm(np.linalg.inv(V)).real
    r1 = Mm12 @ omega @ Mm12
    s1, K = schur(r1)
    X = np.array([[0, 1], [1, 0]])
    I = np.identity(2)
    seq = []
    for i in range(n):
        if s1[2 * i, 2 * i + 1] > 0:
            seq.append(I)
        else:
            seq.append(X)
    p = block_diag(*seq)
    Kt = K @ p
    s1t = p @ s1 @ p
    dd = xpxp_to_xxpp(s1t)
    perm_indices = xpxp_to_xxpp(np.arange(2 * n))
    Ktt = Kt[:, perm_indices]
    Db = np.diag([1 / dd[i, i + n] for i in range(n)] + [1 / dd[i, i + n] for i in range(n)])
    S = Mm12 @ Ktt @ sqrtm(Db)
    return Db, np.linalg.inv(S).T


def thermal_photons(nth, cutoff=20):
    """kron_cpu.py:92-93."""
    return 1 / (nth + 1) * (nth / (nth + 1)) ** np.arange(cutoff)


@_single_thread_blas
def blochmessiah(S):
    """MPS_cpu.py:189-229."""
    N, _ = S.shape
    R = (1 / np.sqrt(2)) * np.block(
        [[np.eye(N // 2), 1j * np.eye(N // 2)], [np.eye(N // 2), -1j * np.eye(N // 2)]]
    )
    Sc = R @ S @ np.conjugate(R).T
    u1, d1, v1 = svd(Sc, lapack_driver='gesvd')
    Sig = u1 @ np.diag(d1) @ np.conjugate(u1).T
    Unitary = u1 @ v1
    alpha = Unitary[0:N // 2, 0:N // 2]
    beta = Sig[0:N // 2, N // 2:N]
    u2, d2, v2 = np.linalg.svd(beta)
    sval = np.arcsinh(d2)
    takagibeta = u2 @ sqrtm(np.conjugate(u2).T @ (v2.T))
    uf = np.block([[takagibeta, 0 * takagibeta], [0 * takagibeta, np.conjugate(takagibeta)]])
    vf = np.block(
        [
            [np.conjugate(takagibeta).T @ alpha, 0 * takagibeta],
            [0 * takagibeta, np.conjugate(np.conjugate(takagibeta).T @ alpha)],
        ]
    )
    df = np.block(
        [
            [np.diag(np.cosh(sval)), np.diag(np.sinh(sval))],
            [np.diag(np.sinh(sval)), np.diag(np.cosh(sval))],
        ]
    )
    uff = np.conjugate(R).T @ uf @ R
    vff = np.conjugate(R).T @ vf @ R
    dff = np.conjugate(R).T @ df @ R
    dff = np.real_if_close(dff)
    vff = np.real_if_close(vff)
    uff = np.real_if_close(uff)
    return uff, dff, vff


@_single_thread_blas
def get_U2_sq_U1(S_l, S_r):
    """MPS_cpu.py:357-372."""
    M = len(S_r) // 2
    mode = np.arange(M - 1) + 1
    modes = np.append(mode, mode + M)
    S_l2_inv = np.eye(2 * M, dtype=float)
    S_l2_inv[np.ix_(modes, modes)] = np.linalg.inv(S_l)
    S = S_l2_inv @ S_r
    S2, SQ, S1 = blochmessiah(S)
    U2 = S2[:M, :M] - 1j * S2[:M, M:]
    U1 = S1[:M, :M] - 1j * S1[:M, M:]
    sq = np.log(np.diag(SQ)[:M])
    return U2, sq, U1


@_single_thread_blas
def get_Sigma(U2, sq, U1):
    """MPS_cpu.py:305-312, in complex128."""
    M = len(sq)
    Sigma = np.zeros((2 * M, 2 * M), dtype=np.complex128)
    Sigma[:M, :M] = U2 @ np.diag(np.tanh(sq)) @ U2.T
    Sigma[:M, M:] = U2 @ np.diag(1 / np.cosh(sq)) @ U1
    Sigma[M:, :M] = U1.T @ np.diag(1 / np.cosh(sq)) @ U2.T
    Sigma[M:, M:] = -U1.T @ np.diag(np.tanh(sq)) @ U1
    return Sigma


# ---------------------------------------------------------------------------------------
# Schmidt spectrum per cut (kron_cpu.py:95-153)
# ---------------------------------------------------------------------------------------


def kron_from_thermal(tps, chi, max_dim=10 ** 5, err_tol=10 ** (-12)):
    """The pruned Kronecker loop and final top-chi of kron_cpu.py:106-125, given the thermal
    distributions tps[i] = thermal_photons(nth_i, cutoff) of the cut's normal modes."""
    cutoff = len(tps[0])
    res = tps[0]
    num = np.arange(cutoff, dtype='int8')
    for i in range(1, len(tps)):
        res = np.outer(res, np.array(tps[i])).reshape(-1)
        keep_idx = np.where(res > err_tol)[0]
        res = res[keep_idx]
        idx = np.argsort(res)[-min(len(res), max_dim):]
        res = res[idx][::-1]
        if len(num.shape) == 1:
            num = num.reshape(-1, 1)
        keep_idx = keep_idx[idx][::-1]
        num = np.concatenate([num[keep_idx // cutoff],
                              np.arange(cutoff).reshape(-1, 1)[keep_idx % cutoff]], axis=1)
    len_ = min(chi, len(res))
    idx = np.argsort(res)[-len_:]
    idx_sorted = idx[np.argsort(res[idx])]
    res = res[idx_sorted][::-1]
    num = num[idx_sorted][::-1]
    return np.ascontiguousarray(res), np.ascontiguousarray(num.astype('int8'))


@_single_thread_blas
def cut_thermal(sq_cov, L, cutoff):
    """Williamson of modes L..M-1 and the thermal distribution of each normal mode
    (kron_cpu.py:96-106)."""
    M = len(sq_cov) // 2
    mode = np.arange(L, M)
    modes = np.append(mode, mode + M)
    sq_cov_A = sq_cov[np.ix_(modes, modes)]
    D, S = williamson(sq_cov_A)
    d = (np.diag(D) - 1) / 2
    d[d < 0] = 0
    return [thermal_photons(d[i], cutoff) for i in range(M - L)], S


def get_cumsum_kron(sq_cov, L, chi=100, max_dim=10 ** 5, cutoff=6, err_tol=10 ** (-12)):
    """kron_cpu.py:95-127, returning float64 ``res`` (no float16 cast)."""
    tps, S = cut_thermal(sq_cov, L, cutoff)
    res, num = kron_from_thermal(tps, chi, max_dim, err_tol)
    return res, num, S


def run_kron(sq_cov, d, chi, max_dim=10 ** 5):
    """kron_cpu.py:132-153 without file I/O: returns lists res_k, num_k, S_k for k=0..M-2."""
    M = len(sq_cov) // 2
    out_res, out_num, out_S = [], [], []
    for compute_site in range(M - 1):
        res, num, S_l = get_cumsum_kron(sq_cov, compute_site + 1, max_dim=max_dim, chi=chi,
                                        cutoff=d)
        out_res.append(res)
        out_num.append(num)
        out_S.append(S_l)
    return out_res, out_num, out_S


# ---------------------------------------------------------------------------------------
# Hafnians (MPS_cpu.py:19-47, 117-187, 314-355)
# ---------------------------------------------------------------------------------------


def Sigma_select(Sigma, target):
    """MPS_cpu.py:19-37 (gather A[l,r,c] = Sigma[target[l,r], target[l,c]])."""
    batch_size = 65535
    n_batch, n_select = target.shape
    Sigma2 = np.zeros([n_batch, n_select, n_select], dtype=Sigma.dtype)
    for batch_id in range(n_batch // batch_size + 1):
        begin = batch_id * batch_size
        end = min((batch_id + 1) * batch_size, n_batch)
        batch_target = target[begin:end]
        rows = np.repeat(batch_target, n_select, axis=1).reshape(-1)
        cols = np.tile(batch_target, (1, n_select)).reshape(-1)
        Sigma2[begin:end] = Sigma[rows, cols].reshape(end - begin, n_select, n_select)
    return Sigma2


def push_to_end(array):
    """MPS_cpu.py:39-47."""
    n_batch, n_select = array.shape
    new_array = np.zeros_like(array)
    idx = np.zeros(n_batch, dtype='int32')
    for i in range(1, n_select + 1):
        occupied = array[:, -i] != 0
        idx += occupied
        new_array[np.arange(n_batch), -idx] = array[:, -i]
    return new_array


def hafnian(A):
    """MPS_cpu.py:117-140 (batched dispatch), complex128."""
    matshape = A.shape[1:]
    n_batch = A.shape[0]
    if matshape == (0, 0):
        return np.ones(n_batch, dtype=np.complex128)
    if matshape[0] % 2 != 0:
        return np.zeros(n_batch, dtype=np.complex128)
    if matshape[0] == 2:
        return A[:, 0, 1]
    if matshape[0] == 4:
        return A[:, 0, 1] * A[:, 2, 3] + A[:, 0, 2] * A[:, 1, 3] + A[:, 0, 3] * A[:, 1, 2]
    return recursive_hafnian(A)


def recursive_hafnian(A):
    """MPS_cpu.py:143-160."""
    n_batch, nb_lines, nb_columns = A.shape
    if nb_lines != nb_columns:
        raise ValueError("Matrix must be square")
    if nb_lines % 2 != 0:
        raise ValueError("Matrix size must be even")
    n = A.shape[1] // 2
    z = np.zeros((n_batch, n * (2 * n - 1), n + 1), dtype=A.dtype)
    for j in range(1, 2 * n):
        ind = j * (j - 1) // 2
        for k in range(j):
            z[:, ind + k, 0] = A[:, j, k]
    g = np.zeros([n_batch, n + 1], dtype=A.dtype)
    g[:, 0] = 1
    return solve(z, 2 * n, 1, g, n)


def solve(b, s, w, g, n):
    """MPS_cpu.py:163-187: truncated-polynomial recursion, same loop and add order."""
    n_batch = b.shape[0]
    if s == 0:
        return w * g[:, n]
    c = np.zeros((n_batch, (s - 2) * (s - 3) // 2, n + 1), dtype=g.dtype)
    i = 0
    for j in range(1, s - 2):
        for k in range(j):
            c[:, i] = b[:, (j + 1) * (j + 2) // 2 + k + 2]
            i += 1
    h = solve(c, s - 2, -w, g, n)
    e = g.copy()
    for u in range(n):
        for v in range(n - u):
            e[:, u + v + 1] += g[:, u] * b[:, 0, v]
    for j in range(1, s - 2):
        for k in range(j):
            for u in range(n):
                for v in range(n - u):
                    c[:, j * (j - 1) // 2 + k, u + v + 1] += (
                        b[:, (j + 1) * (j + 2) // 2, u] * b[:, (k + 1) * (k + 2) // 2 + 1, v]
                        + b[:, (k + 1) * (k + 2) // 2, u] * b[:, (j + 1) * (j + 2) // 2 + 1, v]
                    )
    return h + solve(c, s - 2, w, e, n)


def get_target(num):
    """MPS_cpu.py:314-332: occupation rows -> 1-based repeated mode indices, zero padded."""
    num = np.asarray(num)
    n_select = np.sum(num, axis=1).max() if num.shape[0] else 0
    n_batch, n_len = num.shape
    idx_x = np.tile(np.arange(n_select, dtype='int32').reshape(1, -1), (n_batch, 1))
    target = np.zeros([n_batch, n_select], dtype='int32')
    idx_end = np.zeros(n_batch)
    idx_begin = np.zeros(n_batch)
    for i in range(n_len):
        vals_n = num[:, i]
        idx_end += vals_n
        mask = idx_x >= idx_begin.reshape(-1, 1)
        mask *= idx_x < idx_end.reshape(-1, 1)
        target += mask * (i + 1)
        idx_begin = np.copy(idx_end)
    return np.array(target, dtype='int32')


def A_elem(Sigma, target, denominator, max_memory_in_gb=0.5):
    """MPS_cpu.py:334-355 (chunked gather + hafnian, then / denominator), complex128."""
    n_batch, n_select = target.shape
    all_haf = np.zeros([0], dtype=np.complex128)
    if n_select == 0:
        n_batch_max = 99999999999
    else:
        n_batch_max = int(max_memory_in_gb * (10 ** 9) // (n_select ** 2 * 8))
    for begin_batch in range(0, n_batch, n_batch_max):
        end_batch = min(n_batch, begin_batch + n_batch_max)
        Sigma2 = Sigma_select(Sigma, target[begin_batch:end_batch])
        haf = hafnian(Sigma2).astype(np.complex128)
        all_haf = np.append(all_haf, haf)
    return all_haf / denominator


# ---------------------------------------------------------------------------------------
# Site tensors (MPS_cpu.py:377-525)
# ---------------------------------------------------------------------------------------


@_single_thread_blas
def site_sigma(compute_site, M, sq_cov, kron_S):
    """Sigma, sq and Z of one site, as MPS_cpu.py:395, 409-414, 439-442, 481-483 compute them."""
    if compute_site == 0:
        _, S_r = williamson(sq_cov)
        S_l = kron_S[0]
    elif compute_site == M - 1:
        S_r = kron_S[compute_site - 1]
        S_l = np.zeros((0, 0))
    else:
        S_r = kron_S[compute_site - 1]
        S_l = kron_S[compute_site]
    U2, sq, U1 = get_U2_sq_U1(S_l, S_r)
    Sigma = get_Sigma(U2, sq, U1)
    Z = np.sqrt(np.prod(np.cosh(sq)))
    return Sigma, Z


def build_site(compute_site, M, sq_cov, kron_res, kron_num, kron_S, d, chi, err_tol=1e-10,
               max_memory_in_gb=0.5):
    """One iteration of MPS_cpu.py:384-525; returns (Gamma [chi,chi,d] c128, Lambda [chi] f64).

    Lambda is all zeros for the last site (the reference saves none there, :524-525).
    """
    Sigma, Z = site_sigma(compute_site, M, sq_cov, kron_S)
    Gamma = np.zeros([chi, chi, d], dtype=np.complex128)
    Lambda = np.zeros([chi], dtype=np.float64)

    if compute_site == 0:  # MPS_cpu.py:402-428
        res = kron_res[0]
        num = kron_num[0].reshape(kron_num[0].shape[0], -1)
        num = num[res > err_tol]
        res = res[res > err_tol]
        left_target = get_target(num)
        left_sum = np.sum(num, axis=1)
        left_denominator = np.sqrt(np.prod(np.array(factorial(num)), axis=1))
        Lambda[:len(res)] = np.array(np.sqrt(res))
        for j in np.arange(d):
            for size in np.arange(np.max(left_sum) + 1):
                left_idx = np.where(left_sum == size)[0]
                if (Lambda[left_idx] <= err_tol).all():
                    continue
                n_batch = left_idx.shape[0]
                target = np.append(np.zeros([n_batch, j], dtype='int32'),
                                   left_target[:, :size][left_idx], axis=1)
                denominator = np.sqrt(factorial(j)) * left_denominator[left_idx]
                haf = A_elem(Sigma, target, denominator, max_memory_in_gb)
                Gamma[0, left_idx, j] = haf / Z / Lambda[left_idx]

    elif compute_site == M - 1:  # MPS_cpu.py:430-459
        num_pre = kron_num[compute_site - 1]
        num_pre = num_pre.reshape(num_pre.shape[0], -1)
        right_target = get_target(num_pre)
        right_sum = np.array(np.sum(num_pre, axis=1))
        right_denominator = np.sqrt(np.prod(np.array(factorial(num_pre)), axis=1))
        for j in np.arange(d):
            for size in np.arange(int(np.nanmax(right_sum)) + 1):
                right_idx = np.where(right_sum == size)[0]
                n_batch = right_idx.shape[0]
                if size == 0 and j == 0:
                    Gamma[right_idx, 0, j] = np.ones(n_batch) / Z
                    continue
                target = np.copy(right_target[:, :size][right_idx])
                if size == 0:
                    target = np.zeros([n_batch, 0], dtype='int32')
                target = np.append(np.zeros([n_batch, j], dtype=int), target, axis=1)
                denominator = np.sqrt(factorial(j)) * right_denominator[right_idx]
                haf = A_elem(Sigma, target, denominator, max_memory_in_gb)
                Gamma[right_idx, 0, j] = haf / Z

    else:  # MPS_cpu.py:461-520
        num_pre = kron_num[compute_site - 1]
        num_pre = num_pre.reshape(num_pre.shape[0], -1)
        right_target = np.array(push_to_end(get_target(num_pre)))
        right_sum = np.array(np.sum(num_pre, axis=1))
        right_denominator = np.sqrt(np.prod(np.array(factorial(num_pre)), axis=1))
        num = kron_num[compute_site]
        res = kron_res[compute_site]
        num = num.reshape(num.shape[0], -1)
        num = num[res > err_tol]
        left_target = get_target(num)
        left_n_select = left_target.shape[1]
        left_sum = np.array(np.sum(num, axis=1))
        full_sum = (np.repeat(left_sum.reshape(-1, 1), right_sum.shape[0], axis=1)
                    + np.repeat(right_sum.reshape(1, -1), left_sum.shape[0], axis=0))
        left_denominator = np.sqrt(np.prod(np.array(factorial(num)), axis=1))
        res = res[res > err_tol]
        Lambda[:len(res)] = np.array(np.sqrt(res))
        for j in np.arange(d):
            gpu_Gamma = np.zeros([chi, chi], dtype=np.complex128)
            for size in np.arange(int(np.nanmax(full_sum)) + 1):
                left_idx, right_idx = np.where(full_sum == size)
                n_batch = left_idx.shape[0]
                if (Lambda[left_idx] <= err_tol).all():
                    continue
                if size == 0 and j == 0:
                    gpu_Gamma[right_idx, left_idx] = np.ones(n_batch) / Z / Lambda[left_idx]
                    continue
                if size == 0:
                    n_batch_max = 99999999999
                else:
                    n_batch_max = int(max_memory_in_gb * (10 ** 9) // (size * 8))
                for begin_batch in range(0, n_batch, n_batch_max):
                    end_batch = min(n_batch, begin_batch + n_batch_max)
                    target = np.zeros([end_batch - begin_batch, size], dtype='int32')
                    target[:, :left_n_select] = np.copy(
                        left_target[:, :size][left_idx[begin_batch:end_batch]])
                    right_target_chosen = np.copy(
                        right_target[:, -size:][right_idx[begin_batch:end_batch]])
                    if size == 0:
                        right_target_chosen = np.zeros([end_batch - begin_batch, 0], dtype='int32')
                    right_n_select = right_target_chosen.shape[1]
                    non_zero_locations = np.where(right_target_chosen != 0)
                    right_target_chosen[non_zero_locations] += num.shape[1]
                    if right_n_select:
                        target[:, -right_n_select:] += right_target_chosen
                    target = np.append(np.zeros([end_batch - begin_batch, j], dtype=int), target,
                                       axis=1)
                    denominator = (np.sqrt(factorial(j)) * left_denominator[left_idx[begin_batch:end_batch]]
                                   * right_denominator[right_idx[begin_batch:end_batch]])
                    haf = A_elem(Sigma, target, denominator, max_memory_in_gb)
                    gpu_Gamma[right_idx[begin_batch:end_batch], left_idx[begin_batch:end_batch]] = (
                        haf / Z / Lambda[left_idx[begin_batch:end_batch]])
            Gamma[:, :, j] = gpu_Gamma
    return Gamma, Lambda


def build_mps(sq_cov, d, chi, max_dim=10 ** 5, kron=None):
    """kron_cpu.py __main__ followed by MPS_cpu.py __main__: lists of Gamma_k, Lambda_k."""
    M = len(sq_cov) // 2
    if kron is None:
        kron = run_kron(sq_cov, d, chi, max_dim)
    kres, knum, kS = kron
    Gammas, Lambdas = [], []
    for k in range(M):
        G, L = build_site(k, M, sq_cov, kres, knum, kS, d, chi)
        Gammas.append(G)
        Lambdas.append(L)
    return Gammas, Lambdas


# ---------------------------------------------------------------------------------------
# Sampling (sampling_cpu.py:30-183)
# ---------------------------------------------------------------------------------------


def destroy(N):
    """sampling_cpu.py:120-122 (complex128)."""
    data = np.sqrt(np.arange(1, N, dtype=np.complex128))
    return np.diag(data, 1)


def batch_displaces(N, alphas):
    """sampling_cpu.py:133-148: expm(alpha a^dag - alpha^* a) per sample and mode."""
    samples_in_parallel = alphas.shape[0]
    M = alphas.shape[1]
    a = destroy(N)
    a_h = np.conj(a).T
    a = np.repeat(a[np.newaxis], samples_in_parallel, axis=0)
    a_h = np.repeat(a_h[np.newaxis], samples_in_parallel, axis=0)
    results = []
    for i in range(M):
        alpha = alphas[:, i].reshape(-1, 1, 1)
        results.append(expm(alpha * a_h - np.conj(alpha) * a))
    return np.transpose(np.array(results), (1, 0, 2, 3))


def batch_mu_to_alpha(mu, hbar=2):
    """sampling_cpu.py:156-159."""
    M = mu.shape[1] // 2
    return (mu[:, :M] + 1j * mu[:, M:]) / np.sqrt(2 * hbar)


def thermal_sqrtW(cov, sq_cov):
    """sampling_cpu.py:168-170: regularised Cholesky factor of W = cov - sq_cov.

    Deviation: when W is singular to working precision (synthetic instances have an exactly
    rank-M W) that Cholesky fails; the factor is then V diag(sqrt(max(lambda, 0))) from the
    eigendecomposition of W.  Any L with L L^T = W gives the same distribution of L z."""
    thermal_cov = cov - sq_cov
    reg = thermal_cov + 1.000001 * np.eye(len(thermal_cov)) * np.abs(
        np.min(np.linalg.eigvalsh(thermal_cov)))
    try:
        return np.linalg.cholesky(reg)
    except np.linalg.LinAlgError:
        lam, V = np.linalg.eigh(0.5 * (thermal_cov + thermal_cov.T))
        return V * np.sqrt(np.maximum(lam, 0.0))[np.newaxis, :]


def sampling(Gammas, Lambda, sqrtW, dd, normals, uniforms, displacements=None):
    """sampling_cpu.py:30-117 with explicit random arrays.

    Gammas   : list of M arrays [chi, chi, d] complex128 (site tensors)
    Lambda   : [chi, M-1] float64 (column i = Lambda_i)
    normals  : [2M, n] standard normals (sampling_cpu.py:37)
    uniforms : [M, n] uniforms in [0,1) (sampling_cpu.py:52 and :92, one row per site)
    Returns (samples [n, M] int64, probs [n, M, dd] float64 -- the normalised conditional
    distribution each site's draw was made from).
    """
    chi, _, d = Gammas[0].shape
    M = len(sqrtW) // 2
    samples_in_parallel = normals.shape[1]
    if displacements is None:
        pure_mu = (sqrtW @ normals).T
        pure_alpha = batch_mu_to_alpha(pure_mu, hbar=2)
        displacements = batch_displaces(dd, pure_alpha)
    res = []
    all_probs = []
    pre_tensor = None
    for i in range(M):
        if i == 0:
            temp_tensor = np.zeros([chi, chi, dd], dtype=np.complex128)
            temp_tensor[:, :, :d] = Gammas[0]
            temp_tensor = np.sum(temp_tensor, axis=0)
            temp_tensor = np.einsum('mj,Bkj->Bmk', temp_tensor, displacements[:, i])
            pre_tensor = np.copy(temp_tensor)
            temp_tensor = np.abs(temp_tensor) ** 2
            probs = [np.dot(temp_tensor[:, :, j], Lambda[:, 0] ** 2) for j in range(dd)]
        else:
            tensor = pre_tensor * Lambda[:, i - 1]
            Gamma_temp = np.zeros([chi, chi, dd], dtype=np.complex128)
            Gamma_temp[:, :, :d] = Gammas[i]
            temp_tensor = (tensor @ Gamma_temp.reshape(chi, chi * dd)).reshape(
                samples_in_parallel, chi, dd)
            temp_tensor = np.einsum('Bmj,Bkj->Bmk', temp_tensor, displacements[:, i])
            pre_tensor = np.copy(temp_tensor)
            temp_tensor = np.abs(temp_tensor) ** 2
            probs = []
            for j in range(dd):
                if i == M - 1:
                    probs.append(temp_tensor[:, 0, j])
                else:
                    probs.append(np.dot(temp_tensor[:, :, j], Lambda[:, i] ** 2))
        probs = np.array(probs).T
        probs = probs / np.sum(probs, axis=1)[:, np.newaxis]
        all_probs.append(probs)
        cumulative_probs = np.cumsum(probs, axis=1)
        thresholds = np.repeat(uniforms[i].reshape(-1, 1), dd, axis=1)
        n_photons = np.minimum(np.sum(thresholds > cumulative_probs, axis=1), dd - 1)
        res.append(n_photons)
        if i == M - 1:
            break
        pre_tensor = pre_tensor[np.arange(samples_in_parallel), :, n_photons]
    return np.array(res).T.astype(np.int64), np.stack(all_probs, axis=1)


# ---------------------------------------------------------------------------------------
# Covariance construction for synthetic instances (make_cov.py:15-60)
# ---------------------------------------------------------------------------------------


def Qmat(cov, hbar=2):
    """make_cov.py:15-31."""
    N = len(cov) // 2
    I = np.identity(N)
    x = cov[:N, :N] * 2 / hbar
    xp = cov[:N, N:] * 2 / hbar
    p = cov[N:, N:] * 2 / hbar
    aidaj = (x + p + 1j * (xp - xp.T) - 2 * I) / 4
    aiaj = (x - p + 1j * (xp + xp.T)) / 4
    return np.block([[aidaj, aiaj.conj()], [aiaj, aidaj.conj()]]) + np.identity(2 * N)


def Covmat(Q, hbar=2):
    """make_cov.py:34-45."""
    n = len(Q) // 2
    I = np.identity(n)
    N = Q[0:n, 0:n] - I
    M = Q[n:2 * n, 0:n]
    mm11a = 2 * (N.real + M.real) + np.identity(n)
    mm22a = 2 * (N.real - M.real) + np.identity(n)
    mm12a = 2 * (M.imag + N.imag)
    cov = np.block([[mm11a, mm12a], [mm12a.T, mm22a]])
    return (hbar / 2) * cov


def cov_from_T(r_array, T):
    """make_cov.py:47-60."""
    cov0 = np.diag([np.exp(2 * r) for r in r_array] + [np.exp(-2 * r) for r in r_array])
    Q_in = Qmat(cov0)
    Tc = T.conj()
    Tt = T.T
    Th = Tt.conj()
    a = block_diag(T, Tc)
    b = block_diag(Th, Tt)
    A = (np.eye(len(T) * 2) - a @ b)
    B = a @ Q_in @ b
    Q_out = A + B
    return Covmat(Q_out)


def photon_number_moments(cov, hbar=2):
    """Exact first moments E[n_i] = (tr V_ii - 2)/4 (hbar=2) and the covariance
    Cov(n_i, n_j) = |N_ij|^2 + |M_ij|^2 (i != j) of a zero-mean Gaussian state, from the
    Q-matrix blocks of make_cov.py:15-31 (formula used by the reference's
    analysis/Correlation_photon_number.ipynb)."""
    M = len(cov) // 2
    Q = Qmat(cov, hbar)
    N = Q[:M, :M] - np.eye(M)          # <a_i^dag a_j>
    Mm = Q[M:, :M]                     # <a_i a_j>
    mean = np.real(np.diag(N))
    covn = np.abs(N) ** 2 + np.abs(Mm) ** 2
    np.fill_diagonal(covn, mean * (mean + 1) + np.abs(np.diag(Mm)) ** 2)
    return mean, covn
