"""Photon-number sampling (L4) -- drop-in for sampling_cpu.py.

The MPS is walked site by site with all samples of a batch advanced together (site-major):
for site i the carried vectors go through the FP64 MFMA contraction with Gamma_i
(``bsmps_site_contract``), then the per-sample displacement, conditional probabilities and
draw (``bsmps_sample_site``).  Random numbers come from counter-based Philox streams on the
device, so a sample's randomness depends only on (seed, sample index):

* normals  z[k][s]  -- stream k (k = 0..2M-1), element s;
* uniforms u[i][s]  -- stream UNIFORM_STREAM + i (site i), element s.

Deviation from the reference, by design: a draw is clamped to dd-1 when rounding leaves the
last cumulative probability below the uniform (sampling_cpu.py:103-104 would return dd).

CLI (same flags and files as the reference)::

    python -m bosonsupremacy_amd.sampling --N 1000 --n 100 --iter 1 --d 4 --dd 10 --chi 64 --dir ROOT/
"""
import argparse
import ctypes

import numpy as np
import torch

from . import runtime

UNIFORM_STREAM = 0x10000
HBAR = 2


def destroy(N):
    """Truncated annihilation operator (sampling_cpu.py:120-122), complex128."""
    return np.diag(np.sqrt(np.arange(1, N, dtype=np.complex128)), 1)


def displacement_basis(dd):
    """Eigen-decomposition of the truncated a + a^dag: (W [dd, dd], x [dd])."""
    a = np.diag(np.sqrt(np.arange(1, dd, dtype=np.float64)), 1)
    x, W = np.linalg.eigh(a + a.T)
    return np.ascontiguousarray(W), np.ascontiguousarray(x)


def batch_mu_to_alpha(mu, hbar=HBAR):
    """sampling_cpu.py:156-159 (host helper, for API parity)."""
    M = mu.shape[1] // 2
    return (mu[:, :M] + 1j * mu[:, M:]) / np.sqrt(2 * hbar)


def thermal_sqrtW(cov, sq_cov):
    """Factor L (L L^T = W) of the classical noise W = cov - sq_cov: the regularised Cholesky
    of sampling_cpu.py:168-170, or, when W is singular to working precision and that
    Cholesky fails, V diag(sqrt(max(lambda, 0))) from the eigendecomposition of W."""
    W = cov - sq_cov
    reg = W + 1.000001 * np.eye(len(W)) * np.abs(np.min(np.linalg.eigvalsh(W)))
    try:
        return np.linalg.cholesky(reg)
    except np.linalg.LinAlgError:
        lam, V = np.linalg.eigh(0.5 * (W + W.T))
        return V * np.sqrt(np.maximum(lam, 0.0))[np.newaxis, :]


def batch_displaces(N, alphas):
    """expm(alpha a^dag - alpha^* a) for alphas [B, M] -> [B, M, N, N] (sampling_cpu.py:133-148),
    computed on the GPU in FP64."""
    ctx = runtime.context()
    al = np.ascontiguousarray(alphas, dtype=np.complex128)
    B, M = al.shape
    W, x = displacement_basis(N)
    a_d = runtime.to_device(al)
    W_d, x_d = runtime.to_device(W), runtime.to_device(x)
    out = torch.empty((B * M, N, N), dtype=torch.complex128, device=ctx.device)
    ctx.call("bsmps_batch_displaces", N, runtime.ptr(a_d), B * M, 1, runtime.ptr(W_d), runtime.ptr(x_d),
             runtime.ptr(out))
    return runtime.to_host(out).reshape(B, M, N, N)


def philox(n_streams, n, seed, stream0, first=0, normal=False):
    """Device array [n_streams, n]: elements first..first+n-1 of Philox streams stream0.."""
    ctx = runtime.context()
    out = torch.empty((n_streams, n), dtype=torch.float64, device=ctx.device)
    ctx.call("bsmps_philox", int(seed), int(stream0), int(n_streams), int(first), int(n), int(bool(normal)),
             runtime.ptr(out))
    return out


def random_inputs(M, n, seed, offset=0):
    """Device normals [2M, n] and uniforms [M, n] for samples offset..offset+n-1."""
    normals = philox(2 * M, n, seed, 0, offset, normal=True)
    uniforms = philox(M, n, seed, UNIFORM_STREAM, offset)
    return normals, uniforms


class Sampler:
    """Samples photon numbers from an MPS held as device tensors.

    gammas : list of M device tensors [chi_{i-1}, chi_i, d_i] complex128 (the plain method:
             all [chi, chi, d]; site 0 uses row 0 only, so its first extent may be 1)
    lambdas: list of M float64 arrays [chi_i] (the last one unused)
    sqrtW  : [2M, 2M] Cholesky factor of the classical noise covariance
    """

    PRECISIONS = ("fp64", "fp32", "bf16x3", "bf16x6", "fp16x3")

    def __init__(self, gammas, lambdas, sqrtW, dd, precision="fp64", ctx=None):
        self.ctx = ctx or runtime.context()
        dev = self.ctx.device
        self.M = len(gammas)
        self.chis = [int(g.shape[1]) for g in gammas]
        self.ds = [int(g.shape[2]) for g in gammas]
        self.chi, self.d = max(self.chis), max(self.ds)
        self.dd = dd
        if dd < self.d:
            raise ValueError("dd must be >= d")
        for i in range(1, self.M):
            if gammas[i].shape[0] < self.chis[i - 1]:
                raise ValueError(f"Gamma_{i} has {gammas[i].shape[0]} rows, fewer than the {self.chis[i - 1]} "
                                 f"Schmidt vectors of the previous cut")
            if len(lambdas[i - 1]) < self.chis[i - 1]:
                raise ValueError(f"Lambda_{i - 1} has fewer than {self.chis[i - 1]} entries")
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {self.PRECISIONS}")
        self.precision = precision
        # device preparation runs on the context's stream (after the caller's pending work), so a
        # context bound to a side stream sees its inputs ready
        self.ctx.stream.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(self.ctx.stream):
            self._prepare(gammas, lambdas, sqrtW, dd, dev)
        self.dev = dev

    def _prepare(self, gammas, lambdas, sqrtW, dd, dev):
        """Device copies of the site tensors in the form's precision (and their bf16 splits)."""
        precision = self.precision
        self.lambdas = [runtime.to_device(np.asarray(l, dtype=np.float64)) for l in lambdas]
        ctype = torch.complex128 if precision == "fp64" else torch.complex64
        self.gammas = [g.to(ctype).contiguous() for g in gammas]
        self.splits = None
        if precision == "bf16x3":
            # hi / lo bf16 planes of every Gamma, split once and reused by every batch
            self.splits = []
            for g in self.gammas:
                cnt = 2 * g.numel()
                hi = torch.empty((cnt,), dtype=torch.bfloat16, device=dev)
                lo = torch.empty((cnt,), dtype=torch.bfloat16, device=dev)
                self.ctx.call("bsmps_split_bf16", runtime.ptr(g), cnt, runtime.ptr(hi), runtime.ptr(lo))
                self.splits.append((hi, lo))
        if precision == "bf16x6":
            # six bf16 planes of every contracted Lambda_{i-1} Gamma_i (FP32-exact split), reused
            # by every batch
            self.splits = [None]
            for i in range(1, self.M):
                chi_p, chi, d = self.chis[i - 1], self.chis[i], self.ds[i]
                nbytes = self.ctx.lib.bsmps_bf16x6_plane_bytes(chi_p, chi, d)
                planes = torch.empty((nbytes // 2,), dtype=torch.bfloat16, device=dev)
                self.ctx.call("bsmps_split_bf16x6", chi_p, chi, d, runtime.ptr(self.gammas[i]),
                              runtime.ptr(self.lambdas[i - 1]), runtime.ptr(planes))
                self.splits.append(planes)
        if precision == "fp16x3":
            # two fp16 planes per part of every contracted Lambda_{i-1} Gamma_i, columns scaled to a
            # largest component of 1 (the scales are multiplied back by the contraction)
            self.splits = [None]
            for i in range(1, self.M):
                chi_p, chi, d = self.chis[i - 1], self.chis[i], self.ds[i]
                nbytes = self.ctx.lib.bsmps_fp16x3_plane_bytes(chi_p, chi, d)
                planes = torch.empty((nbytes // 2,), dtype=torch.float16, device=dev)
                colscale = torch.empty((-(-(chi * d) // 64) * 64,), dtype=torch.float32, device=dev)
                self.ctx.call("bsmps_split_fp16x3", chi_p, chi, d, runtime.ptr(self.gammas[i]),
                              runtime.ptr(self.lambdas[i - 1]), runtime.ptr(planes), runtime.ptr(colscale))
                self.splits.append((planes, colscale))
        self.sqrtW = runtime.to_device(np.ascontiguousarray(sqrtW, dtype=np.float64))
        W, x = displacement_basis(dd)
        self.W = runtime.to_device(W)
        self.x = runtime.to_device(x)

    def run(self, normals, uniforms, return_probs=False, timer=None):
        """normals [2M, n], uniforms [M, n] (device float64) -> samples [n, M] int32 (device),
        and optionally the conditional distributions [M, n, dd].  ``timer`` (optional) is a
        context-manager factory ``timer(name)`` wrapped around each kernel launch."""
        import contextlib
        timer = timer or (lambda name: contextlib.nullcontext())
        ctx, M, dd = self.ctx, self.M, self.dd
        n = normals.shape[1]
        normals = normals.contiguous()
        uniforms = uniforms.contiguous()
        alpha = torch.empty((n, M), dtype=torch.complex128, device=self.dev)
        ctx.call("bsmps_mu_to_alpha", runtime.ptr(self.sqrtW), M, runtime.ptr(normals), n, runtime.ptr(alpha))
        samples = torch.empty((n, M), dtype=torch.int32, device=self.dev)
        ctype = torch.complex128 if self.precision == "fp64" else torch.complex64
        draw = "bsmps_sample_site" if self.precision == "fp64" else "bsmps_sample_site_c64"
        pre = torch.empty((n * self.chi,), dtype=ctype, device=self.dev)
        T = torch.empty((n * max(c * dk for c, dk in zip(self.chis, self.ds)),), dtype=ctype, device=self.dev)
        probs = torch.empty((M, n, dd), dtype=torch.float64, device=self.dev) if return_probs else None
        dwork = torch.empty((2 * n * dd * self.d,), dtype=torch.float64, device=self.dev)  # D_s per sample
        pre_planes = rowscale = None
        if self.precision == "bf16x6":
            nbytes = max(self.ctx.lib.bsmps_bf16x6_row_plane_bytes(n, c) for c in self.chis)
            pre_planes = torch.empty((nbytes // 2,), dtype=torch.bfloat16, device=self.dev)
        if self.precision == "fp16x3":
            nbytes = max(self.ctx.lib.bsmps_fp16x3_row_plane_bytes(n, c) for c in self.chis)
            pre_planes = torch.empty((nbytes // 2,), dtype=torch.float16, device=self.dev)
            rowscale = torch.empty((n,), dtype=torch.float32, device=self.dev)
        for i in range(M):
            last = i == M - 1
            chi, d = self.chis[i], self.ds[i]
            if i == 0:
                src, stride = self.gammas[0], 0  # Gamma_0[0, a, j] is the site-0 tensor for every sample
            else:
                chi_p = self.chis[i - 1]
                with timer("site_contract"):
                    if self.precision == "fp64":
                        ctx.call("bsmps_site_contract", n, chi_p, chi, d, runtime.ptr(pre),
                                 runtime.ptr(self.lambdas[i - 1]), runtime.ptr(self.gammas[i]), runtime.ptr(T))
                    elif self.precision == "fp32":
                        ctx.call("bsmps_site_contract_c64", n, chi_p, chi, d, runtime.ptr(pre),
                                 runtime.ptr(self.lambdas[i - 1]), runtime.ptr(self.gammas[i]), runtime.ptr(T))
                    elif self.precision == "fp16x3":
                        planes, colscale = self.splits[i]
                        ctx.call("bsmps_split_rows_fp16x3", n, chi_p, runtime.ptr(pre), runtime.ptr(pre_planes),
                                 runtime.ptr(rowscale))
                        ctx.call("bsmps_site_contract_fp16x3", n, chi_p, chi, d, runtime.ptr(pre_planes),
                                 runtime.ptr(rowscale), runtime.ptr(planes), runtime.ptr(colscale), runtime.ptr(T))
                    elif self.precision == "bf16x6":
                        ctx.call("bsmps_split_rows_bf16x6", n, chi_p, runtime.ptr(pre), runtime.ptr(pre_planes))
                        ctx.call("bsmps_site_contract_bf16x6", n, chi_p, chi, d, runtime.ptr(pre_planes),
                                 runtime.ptr(self.splits[i]), runtime.ptr(T))
                    else:
                        hi, lo = self.splits[i]
                        ctx.call("bsmps_site_contract_bf16x3", n, chi_p, chi, d, runtime.ptr(pre),
                                 runtime.ptr(self.lambdas[i - 1]), runtime.ptr(hi), runtime.ptr(lo), runtime.ptr(T))
                src, stride = T, chi * d
            a_ptr = ctypes.c_void_p(alpha.data_ptr() + 16 * i)
            s_ptr = ctypes.c_void_p(samples.data_ptr() + 4 * i)
            p_ptr = runtime.ptr(probs[i]) if probs is not None else None
            with timer("sample_site"):
                ctx.call(draw, n, chi, d, dd, runtime.ptr(src), stride, a_ptr, M,
                         runtime.ptr(self.W), runtime.ptr(self.x), runtime.ptr(dwork),
                         None if last else runtime.ptr(self.lambdas[i]),
                         runtime.ptr(uniforms[i]), s_ptr, M, None if last else runtime.ptr(pre), p_ptr)
        return (samples, probs) if return_probs else samples


def sampling(gammas, lambdas, sqrtW, dd, n, seed=0, offset=0):
    """Draw n samples (host int64 array [n, M]) with the device Philox streams."""
    s = Sampler(gammas, lambdas, sqrtW, dd)
    normals, uniforms = random_inputs(s.M, n, seed, offset)
    return runtime.to_host(s.run(normals, uniforms)).astype(np.int64)


def main(argv=None):
    parser = argparse.ArgumentParser(description="Photon-number sampling (drop-in for sampling_cpu.py)")
    parser.add_argument('--N', type=int, required=True, help='total number of samples')
    parser.add_argument('--n', type=int, required=True, help='samples per batch')
    parser.add_argument('--iter', type=int, default=1)
    parser.add_argument('--d', type=int, required=True)
    parser.add_argument('--dd', type=int, required=True)
    parser.add_argument('--chi', type=int, required=True)
    parser.add_argument('--dir', type=str, required=True)
    parser.add_argument('--seed', type=int, default=0)
    args = parser.parse_args(argv)
    path = args.dir + f'd_{args.d}_chi_{args.chi}/'
    sq_cov = np.load(args.dir + "sq_cov.npy")
    cov = np.load(args.dir + "cov.npy")
    sqrtW = thermal_sqrtW(cov, sq_cov)
    M = sqrtW.shape[0] // 2
    dev = runtime.context().device
    gammas = [torch.from_numpy(np.load(path + f'Gamma_{i}.npy').astype(np.complex128)).to(dev) for i in range(M)]
    lambdas = [np.load(path + f"Lambda_{i}.npy").astype(np.float64) for i in range(M - 1)] + [np.zeros(args.chi)]
    sampler = Sampler(gammas, lambdas, sqrtW, args.dd)
    for it in range(args.iter):
        chunks = []
        for b0 in range(0, args.N, args.n):
            nb = min(args.N, b0 + args.n) - b0
            normals, uniforms = random_inputs(M, nb, args.seed + it, b0)
            chunks.append(runtime.to_host(sampler.run(normals, uniforms)).astype(np.int64))
            np.save(args.dir + f"samples_{it}.npy", np.concatenate(chunks, axis=0))


if __name__ == "__main__":
    main()
