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


def live_ranks(gammas, lambdas):
    """Schmidt ranks the sampler contracts: r_i = 1 + the index of the last non-zero entry of
    Lambda_i (i < M-1), capped by Gamma_i's column count, and r_{M-1} = 1 (the last site's one
    column, sampling_cpu.py:88-89).  Rows and columns past these ranks are multiplied by a zero
    Lambda (sampling_cpu.py:75, 90), so leaving them out changes no probability."""
    M = len(gammas)
    ranks = []
    for i in range(M - 1):
        lam = lambdas[i].detach().cpu().numpy() if isinstance(lambdas[i], torch.Tensor) else np.asarray(lambdas[i])
        nz = np.flatnonzero(lam)
        ranks.append(max(1, min(int(nz[-1]) + 1 if len(nz) else 1, int(gammas[i].shape[1]))))
    ranks.append(1)
    return ranks


def _as_tensor(g):
    return g if isinstance(g, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(g))


# Photon-number parity of the site tensors.  Every entry of Gamma_i is a hafnian of the zero-mean
# pure state (MPS_cpu.py:117-187), whose matrix size is the photon total of (left label, j, right
# label); an odd total gives a zero entry.  So Gamma_i[a, b, j] = 0 unless p(a) = p(b) xor (j mod 2)
# for the label parities p, and the contraction splits into two independent products of half the
# size: the left labels with p = 0 against the columns (b, j) with p(b) xor j = 0, and the labels
# with p = 1 against the other columns.  The parities are read off Gamma itself (a two-colouring of
# its non-zero pattern, checked entry by entry), so a tensor with this structure -- built here or
# read from the reference's files -- takes the split, and one without it keeps the full product.
# The split costs a search of Gamma's non-zero pattern and two gathers once per site, and a gather
# and a scatter per batch of samples: below about 2^24 live entries (rows x cols x d) a site's
# product saves less than that, so smaller sites keep the full product.
PARITY_MIN_ENTRIES = 1 << 24


def parity_blocks(g):
    """For a device tensor g [rows][cols][d]: ((rows0, cols0), (rows1, cols1)), index tensors of
    the two parity blocks into g's rows and into its flattened cols * d columns, or None when g's
    non-zero entries do not split by parity, a block would be empty, or g is below
    PARITY_MIN_ENTRIES."""
    rows, cols, d = (int(x) for x in g.shape)
    if d < 2 or rows * cols * d < PARITY_MIN_ENTRIES:
        return None
    nz = g != 0
    E = nz[:, :, 0::2].any(dim=2)   # (a, b) joined through an even j
    O = nz[:, :, 1::2].any(dim=2)   # ... through an odd j
    del nz
    dev = g.device
    pa = torch.full((rows,), -1, dtype=torch.int8, device=dev)
    pb = torch.full((cols,), -1, dtype=torch.int8, device=dev)
    row_any = (E | O).any(dim=1)

    def propagate(pa, pb):
        a0, a1 = (pa == 0)[:, None], (pa == 1)[:, None]
        b_even = ((E & a0) | (O & a1)).any(dim=0)
        b_odd = ((E & a1) | (O & a0)).any(dim=0)
        pb = torch.where(b_even & (pb < 0), 0, torch.where(b_odd & (pb < 0), 1, pb.long())).to(torch.int8)
        b0, b1 = (pb == 0)[None, :], (pb == 1)[None, :]
        a_even = ((E & b0) | (O & b1)).any(dim=1)
        a_odd = ((E & b1) | (O & b0)).any(dim=1)
        pa = torch.where(a_even & (pa < 0), 0, torch.where(a_odd & (pa < 0), 1, pa.long())).to(torch.int8)
        return pa, pb

    # colour one connected part at a time from a seed label (even); a dense pattern is coloured in
    # one or two rounds, so rounds run in fours between the checks for uncoloured labels
    for _ in range(rows):
        unseen = torch.nonzero((pa < 0) & row_any)
        if len(unseen) == 0:
            break
        pa[int(unseen[0])] = 0
        for _ in range(rows + cols):
            before = int((pa >= 0).sum()) + int((pb >= 0).sum())
            for _ in range(4):
                pa, pb = propagate(pa, pb)
            if int((pa >= 0).sum()) + int((pb >= 0).sum()) == before:
                break
    pa = torch.where(pa < 0, 0, pa.long()).bool()   # labels without a non-zero entry: either block
    pb = torch.where(pb < 0, 0, pb.long()).bool()
    same = pa[:, None] == pb[None, :]
    if bool(((E & ~same) | (O & same)).any()):    # every non-zero entry checked (an (a, b) pair
        return None                                 # joined through both parities fails here too)
    pc = (pb[:, None] ^ (torch.arange(d, device=dev) % 2 == 1)[None, :]).reshape(-1)
    blocks = []
    for cls in (False, True):
        ia = torch.nonzero(pa == cls).squeeze(1)
        ic = torch.nonzero(pc == cls).squeeze(1)
        if len(ia) == 0 or len(ic) == 0:
            return None
        blocks.append((ia, ic))
    return tuple(blocks)


class _SiteForm:
    """Device data of one site in one precision: the live block Gamma_i[:rows, :cols, :d] and, for
    the split forms, its pre-split operand planes (built once per site, reused by every batch).
    When the site splits by photon-number parity, ``blocks`` holds the two blocks instead (each a
    _SiteBlock) and ``gamma`` is dropped past site 0."""
    __slots__ = ("rows", "cols", "d", "gamma", "split", "blocks")


class _SiteBlock:
    """One parity block of a site: its left labels ``ia`` and columns ``ic`` (index tensors), the
    columns again as the device int32 map the contraction stores through (``cmap``), the block of
    Gamma as a [len(ia), len(ic), 1] tensor, Lambda_{i-1} of its labels and, for the split forms,
    its operand planes."""
    __slots__ = ("ia", "ic", "cmap", "rows", "cols", "gamma", "lam", "split")


def _split_operand(ctx, precision, gamma, rows, cols, d, lam_prev):
    """The reduced forms' once-per-site operand of gamma [rows][cols][d] (device, complex64 for the
    split forms): hi / lo bf16 planes (bf16x3), six bf16 planes of Lambda_{i-1} Gamma (bf16x6), or
    the scaled fp16 planes and their column scales (fp16x3); None for fp64 / fp32."""
    dev = ctx.device
    if precision == "bf16x3":
        # hi / lo bf16 planes of Gamma_i
        cnt = 2 * gamma.numel()
        hi = torch.empty((cnt,), dtype=torch.bfloat16, device=dev)
        lo = torch.empty((cnt,), dtype=torch.bfloat16, device=dev)
        ctx.call("bsmps_split_bf16", runtime.ptr(gamma), cnt, runtime.ptr(hi), runtime.ptr(lo))
        return hi, lo
    if precision == "bf16x6":
        # six bf16 planes of Lambda_{i-1} Gamma_i (FP32-exact split)
        nbytes = ctx.lib.bsmps_bf16x6_plane_bytes(rows, cols, d)
        planes = torch.empty((nbytes // 2,), dtype=torch.bfloat16, device=dev)
        ctx.call("bsmps_split_bf16x6", rows, cols, d, runtime.ptr(gamma), runtime.ptr(lam_prev), runtime.ptr(planes))
        return planes
    if precision == "fp16x3":
        # two fp16 planes per part of Lambda_{i-1} Gamma_i, columns scaled to a largest component of
        # 1 (the scales are multiplied back by the contraction)
        nbytes = ctx.lib.bsmps_fp16x3_plane_bytes(rows, cols, d)
        planes = torch.empty((nbytes // 2,), dtype=torch.float16, device=dev)
        colscale = torch.empty((-(-(cols * d) // 64) * 64,), dtype=torch.float32, device=dev)
        ctx.call("bsmps_split_fp16x3", rows, cols, d, runtime.ptr(gamma), runtime.ptr(lam_prev), runtime.ptr(planes),
                 runtime.ptr(colscale))
        return planes, colscale
    return None


def _prepare_site(ctx, precision, i, g, rows, cols, lam_prev):
    """_SiteForm of Gamma_i (device or host tensor, or NumPy array) restricted to its live block;
    ``lam_prev`` is Lambda_{i-1} on the device (folded into the split forms' planes)."""
    dev = ctx.device
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    f = _SiteForm()
    if not isinstance(g, torch.Tensor):   # NumPy (possibly memory-mapped): read the live block only
        g = np.ascontiguousarray(g[:rows, :cols])
    g = _as_tensor(g)
    f.rows, f.cols, f.d = rows, cols, int(g.shape[2])
    f.gamma = g[:rows, :cols].to(device=dev, dtype=ctype).contiguous()
    f.split = f.blocks = None
    if i == 0:
        return f
    pb = parity_blocks(f.gamma)
    if pb is None:
        f.split = _split_operand(ctx, precision, f.gamma, rows, cols, f.d, lam_prev)
        return f
    flat = f.gamma.view(rows, cols * f.d)
    f.blocks = []
    for ia, ic in pb:
        b = _SiteBlock()
        b.ia, b.ic, b.rows, b.cols = ia, ic, len(ia), len(ic)
        b.cmap = ic.to(torch.int32).contiguous()      # ic < cols * d, the row length of T
        b.gamma = flat.index_select(0, ia).index_select(1, ic).contiguous().view(b.rows, b.cols, 1)
        b.lam = lam_prev.index_select(0, ia).contiguous()
        b.split = _split_operand(ctx, precision, b.gamma, b.rows, b.cols, 1, b.lam)
        f.blocks.append(b)
    f.gamma = None
    return f


def contract_site(ctx, precision, f, nc, pre, lam_prev, T, work, scatter=False):
    """T [nc][cols * d] = (pre [nc][rows] Lambda_{i-1}) Gamma_i for the site form f (every sample's
    carried vector against the site tensor), in one product or, when f splits by parity, in two
    half-size ones: each block's labels gathered from the carried vectors and its product stored
    straight into its columns of T through the block's column map (the ``_cols`` entry points).
    BF16x3, and any form with ``scatter``, instead writes the block's product to ``work.T`` and
    scatters its columns into T (the same values).  ``work`` holds the reduced forms' row planes
    (``pre_planes``, ``rowscale``) and the block output buffer ``T`` (at least nc * cols * d)."""
    if f.blocks is None:
        _contract(ctx, precision, nc, f.rows, f.cols, f.d, runtime.ptr(pre), lam_prev, f.gamma, f.split, T, work)
        return
    ld = f.cols * f.d
    Tv = T[:nc * ld].view(nc, ld)
    pre_v = pre[:nc * f.rows].view(nc, f.rows)
    mapped = precision in _MAPPED and not scatter
    for b in f.blocks:
        pre_b = pre_v.index_select(1, b.ia).contiguous()
        if mapped:
            _contract(ctx, precision, nc, b.rows, b.cols, 1, runtime.ptr(pre_b), b.lam, b.gamma, b.split, T, work,
                      out_cols=(ld, b.cmap))
        else:
            Tb = work.T[:nc * b.cols]
            _contract(ctx, precision, nc, b.rows, b.cols, 1, runtime.ptr(pre_b), b.lam, b.gamma, b.split, Tb, work)
            Tv.index_copy_(1, b.ic, Tb.view(nc, b.cols))


_MAPPED = ("fp64", "fp32", "fp16x3", "bf16x6")   # forms with a column-mapped entry point


def _contract(ctx, precision, nc, rows, cols, d, pre_in, lam_prev, gamma, split, T, work, out_cols=None):
    """One contraction call; ``out_cols=(ldT, cmap)`` stores its columns through cmap into rows
    of ldT (the ``_cols`` entry points)."""
    T = runtime.ptr(T)
    tail, sfx = ((out_cols[0], runtime.ptr(out_cols[1])), "_cols") if out_cols is not None else ((), "")
    if precision == "fp64":
        ctx.call("bsmps_site_contract" + sfx, nc, rows, cols, d, pre_in, runtime.ptr(lam_prev), runtime.ptr(gamma), T,
                 *tail)
    elif precision == "fp32":
        ctx.call("bsmps_site_contract_c64" + sfx, nc, rows, cols, d, pre_in, runtime.ptr(lam_prev),
                 runtime.ptr(gamma), T, *tail)
    elif precision == "fp16x3":
        planes, colscale = split
        ctx.call("bsmps_split_rows_fp16x3", nc, rows, pre_in, runtime.ptr(work.pre_planes),
                 runtime.ptr(work.rowscale))
        ctx.call("bsmps_site_contract_fp16x3" + sfx, nc, rows, cols, d, runtime.ptr(work.pre_planes),
                 runtime.ptr(work.rowscale), runtime.ptr(planes), runtime.ptr(colscale), T, *tail)
    elif precision == "bf16x6":
        ctx.call("bsmps_split_rows_bf16x6", nc, rows, pre_in, runtime.ptr(work.pre_planes))
        ctx.call("bsmps_site_contract_bf16x6" + sfx, nc, rows, cols, d, runtime.ptr(work.pre_planes),
                 runtime.ptr(split), T, *tail)
    else:
        hi, lo = split
        ctx.call("bsmps_site_contract_bf16x3", nc, rows, cols, d, pre_in, runtime.ptr(lam_prev), runtime.ptr(hi),
                 runtime.ptr(lo), T)


class _Work:
    """Scratch of the contraction: row planes of the reduced forms and a block output buffer."""

    def __init__(self, ctx, precision, chunk, chi_max, cd_max, ctype):
        dev = ctx.device
        self.pre_planes = self.rowscale = None
        if precision == "bf16x6":
            nbytes = ctx.lib.bsmps_bf16x6_row_plane_bytes(chunk, chi_max)
            self.pre_planes = torch.empty((nbytes // 2,), dtype=torch.bfloat16, device=dev)
        if precision == "fp16x3":
            nbytes = ctx.lib.bsmps_fp16x3_row_plane_bytes(chunk, chi_max)
            self.pre_planes = torch.empty((nbytes // 2,), dtype=torch.float16, device=dev)
            self.rowscale = torch.empty((chunk,), dtype=torch.float32, device=dev)
        self.T = torch.empty((chunk * cd_max,), dtype=ctype, device=dev)


def _offset(t, elems):
    """Device pointer ``elems`` elements into device tensor t."""
    runtime.ptr(t)  # device and contiguity checks
    return ctypes.c_void_p(t.data_ptr() + elems * t.element_size())


class _Walk:
    """Device state of one sampling pass over n samples, advanced one site at a time (all
    samples through site i before site i+1, sampling_cpu.py:54-113): the displacements, the
    carried vectors, the draws and (optionally) the conditional distributions.

    Samples go through a site in chunks of ``chunk``; with more than one chunk the carried
    vectors are double-buffered (a chunk's new vectors must not overwrite vectors of chunks
    that have not been contracted yet)."""

    def __init__(self, ctx, precision, M, dd, sqrtW, W, x, normals, chi_max, cd_max, d_max, chunk=None,
                 return_probs=False):
        self.ctx, self.precision, self.M, self.dd = ctx, precision, M, dd
        self.W, self.x = W, x
        dev = ctx.device
        n = normals.shape[1]
        self.n = n
        self.chunk = n if chunk is None else max(1, min(int(chunk), n))
        ctype = torch.complex128 if precision == "fp64" else torch.complex64
        normals = normals.contiguous()
        self.alpha = torch.empty((n, M), dtype=torch.complex128, device=dev)
        ctx.call("bsmps_mu_to_alpha", runtime.ptr(sqrtW), M, runtime.ptr(normals), n, runtime.ptr(self.alpha))
        self.samples = torch.empty((n, M), dtype=torch.int32, device=dev)
        n_buf = 1 if self.chunk >= n else 2
        self.pre = [torch.empty((n * chi_max,), dtype=ctype, device=dev) for _ in range(n_buf)]
        self.cur = 0                         # the buffer holding the carried vectors
        self.T = torch.empty((self.chunk * cd_max,), dtype=ctype, device=dev)
        self.probs = torch.empty((M, n, dd), dtype=torch.float64, device=dev) if return_probs else None
        self.dwork = torch.empty((2 * self.chunk * dd * d_max,), dtype=torch.float64, device=dev)  # D_s
        self.work = _Work(ctx, precision, self.chunk, chi_max, cd_max, ctype)

    def site(self, i, f, lam_prev, lam, uniforms_row, timer):
        """Advance every sample through site i (form ``f``); ``lam_prev`` / ``lam`` are
        Lambda_{i-1} / Lambda_i on the device, ``uniforms_row`` the [n] uniforms of the site."""
        ctx, M, dd, prec = self.ctx, self.M, self.dd, self.precision
        last = i == M - 1
        draw = "bsmps_sample_site" if prec == "fp64" else "bsmps_sample_site_c64"
        src_buf = self.pre[self.cur]
        dst_buf = self.pre[1 - self.cur] if len(self.pre) == 2 else src_buf
        for c0 in range(0, self.n, self.chunk):
            nc = min(self.n, c0 + self.chunk) - c0
            if i == 0:
                src, stride = f.gamma, 0   # Gamma_0[0, a, j] is the site-0 tensor of every sample
            else:
                with timer("site_contract"):
                    contract_site(ctx, prec, f, nc, src_buf[c0 * f.rows:], lam_prev, self.T, self.work)
                src, stride = self.T, f.cols * f.d
            with timer("sample_site"):
                ctx.call(draw, nc, f.cols, f.d, dd, runtime.ptr(src), stride, _offset(self.alpha, c0 * M + i), M,
                         runtime.ptr(self.W), runtime.ptr(self.x), runtime.ptr(self.dwork),
                         None if last else runtime.ptr(lam), _offset(uniforms_row, c0),
                         _offset(self.samples, c0 * M + i), M, None if last else _offset(dst_buf, c0 * f.cols),
                         _offset(self.probs[i], c0 * dd) if self.probs is not None else None)
        if len(self.pre) == 2:
            self.cur = 1 - self.cur


class Sampler:
    """Samples photon numbers from an MPS whose site tensors are all held on the device.

    gammas : list of M tensors [chi_{i-1}, chi_i, d_i] complex128 (device or host tensors, or
             NumPy arrays; the plain method: all [chi, chi, d], site 0 uses row 0 only)
    lambdas: list of M float64 arrays [chi_i] (the last one unused)
    sqrtW  : [2M, 2M] Cholesky factor of the classical noise covariance

    Only the live block of each site tensor is kept and contracted: rows and columns up to
    the Schmidt ranks of ``live_ranks`` (zero-padded Lambda entries multiply the rest away) and
    the last site's single column."""

    PRECISIONS = ("fp64", "fp32", "bf16x3", "bf16x6", "fp16x3")

    def __init__(self, gammas, lambdas, sqrtW, dd, precision="fp64", ctx=None):
        self.ctx = ctx or runtime.context()
        dev = self.ctx.device
        self.M = len(gammas)
        if len(lambdas) < self.M - 1:
            raise ValueError("one Lambda per cut is needed")
        self.chis = live_ranks(gammas, lambdas)
        self.rows = [1] + self.chis[:-1]
        self.ds = [int(g.shape[2]) for g in gammas]
        self.chi, self.d = max(self.chis), max(self.ds)
        self.dd = dd
        if dd < self.d:
            raise ValueError("dd must be >= d")
        for i in range(self.M):
            if gammas[i].shape[0] < self.rows[i] or gammas[i].shape[1] < self.chis[i]:
                raise ValueError(f"Gamma_{i} of shape {tuple(gammas[i].shape)} is smaller than its live block "
                                 f"[{self.rows[i]}, {self.chis[i]}]")
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {self.PRECISIONS}")
        self.precision = precision
        # device preparation runs on the context's stream (after the caller's pending work), so a
        # context bound to a side stream sees its inputs ready
        self.ctx.stream.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(self.ctx.stream):
            self.lambdas = [runtime.to_device(np.asarray(runtime.host(l) if l is not None else np.zeros(1),
                                                         dtype=np.float64)) for l in lambdas]
            self.forms = [_prepare_site(self.ctx, precision, i, gammas[i], self.rows[i], self.chis[i],
                                        self.lambdas[i - 1] if i else None) for i in range(self.M)]
            self.sqrtW = runtime.to_device(np.ascontiguousarray(sqrtW, dtype=np.float64))
            W, x = displacement_basis(dd)
            self.W = runtime.to_device(W)
            self.x = runtime.to_device(x)
        self.dev = dev

    def run(self, normals, uniforms, return_probs=False, timer=None):
        """normals [2M, n], uniforms [M, n] (device float64) -> samples [n, M] int32 (device),
        and optionally the conditional distributions [M, n, dd].  ``timer`` (optional) is a
        context-manager factory ``timer(name)`` wrapped around each kernel launch."""
        import contextlib
        timer = timer or (lambda name: contextlib.nullcontext())
        uniforms = uniforms.contiguous()
        walk = _Walk(self.ctx, self.precision, self.M, self.dd, self.sqrtW, self.W, self.x, normals, self.chi,
                     max(c * dk for c, dk in zip(self.chis, self.ds)), self.d, return_probs=return_probs)
        for i in range(self.M):
            walk.site(i, self.forms[i], self.lambdas[i - 1] if i else None,
                      self.lambdas[i] if i < self.M - 1 else None, uniforms[i], timer)
        return (walk.samples, walk.probs) if return_probs else walk.samples


class SiteSampler:
    """Site-major sampling with one site tensor on the device at a time, as the reference
    loads Gamma_i per site inside its sampling loop (sampling_cpu.py:76-80): ``push(i, Gamma_i,
    Lambda_{i-1}, Lambda_i)`` advances every sample through site i -- in chunks of ``chunk``
    samples, so the contraction output stays small -- and then lets the site's device data go.
    The carried vectors of all n samples stay on the device between sites.

    ``chi_max`` / ``d_max`` bound the live ranks and cutoffs of the sites that will be pushed.
    ``push`` takes the site's uniforms from ``uniforms`` unless a row is passed (the replay of
    another form's sample paths, ``mid_uniforms_row``)."""

    def __init__(self, M, dd, sqrtW, normals, uniforms, chi_max, d_max, precision="fp64", chunk=4096,
                 return_probs=True, ctx=None):
        if precision not in Sampler.PRECISIONS:
            raise ValueError(f"precision must be one of {Sampler.PRECISIONS}")
        if dd < d_max:
            raise ValueError("dd must be >= d")
        self.ctx = ctx or runtime.context()
        self.M, self.dd, self.precision = M, dd, precision
        self.chi_max, self.d_max = chi_max, d_max
        self.uniforms = uniforms.contiguous() if uniforms is not None else None
        self.sqrtW = runtime.to_device(np.ascontiguousarray(sqrtW, dtype=np.float64))
        W, x = displacement_basis(dd)
        self.W, self.x = runtime.to_device(W), runtime.to_device(x)
        self.walk = _Walk(self.ctx, precision, M, dd, self.sqrtW, self.W, self.x, normals, chi_max,
                          chi_max * d_max, d_max, chunk=chunk, return_probs=return_probs)
        self.next_site = 0
        self.rows = 1
        self.lam_prev = None

    def push(self, i, gamma, lam, uniforms_row=None, timer=None):
        """Advance all samples through site i.  ``gamma`` is Gamma_i (device/host tensor or
        NumPy array, at least [rank_{i-1}, rank_i, d]); ``lam`` is Lambda_i (ignored at the last
        site).  Sites must come in order 0, 1, ..., M-1."""
        import contextlib
        if i != self.next_site:
            raise ValueError(f"site {i} pushed, site {self.next_site} expected")
        last = i == self.M - 1
        if last:
            cols = 1
            lam_dev = None
        else:
            lam_h = np.asarray(runtime.host(lam), dtype=np.float64)
            nz = np.flatnonzero(lam_h)
            cols = max(1, min(int(nz[-1]) + 1 if len(nz) else 1, int(gamma.shape[1])))
            lam_dev = runtime.to_device(lam_h)
        if cols > self.chi_max or int(gamma.shape[2]) > self.d_max:
            raise ValueError(f"site {i}: rank {cols} / cutoff {gamma.shape[2]} beyond chi_max {self.chi_max} / "
                             f"d_max {self.d_max}")
        if gamma.shape[0] < self.rows or gamma.shape[1] < cols:
            raise ValueError(f"Gamma_{i} of shape {tuple(gamma.shape)} is smaller than its live block "
                             f"[{self.rows}, {cols}]")
        f = _prepare_site(self.ctx, self.precision, i, gamma, self.rows, cols, self.lam_prev)
        u = uniforms_row if uniforms_row is not None else self.uniforms[i]
        self.walk.site(i, f, self.lam_prev, lam_dev, u.contiguous(),
                       timer or (lambda name: contextlib.nullcontext()))
        self.rows, self.lam_prev = cols, lam_dev
        self.next_site += 1

    def state(self):
        """Host copies of what continuing at the next site needs: the carried vectors
        ([n][rank of the last cut walked]), the draws so far, and Lambda of that cut.  The
        displacements are recomputed from the normals."""
        w = self.walk
        pre = w.pre[w.cur][:w.n * self.rows] if self.next_site > 0 else w.pre[w.cur][:0]
        return {"next_site": np.array(self.next_site), "rows": np.array(self.rows),
                "pre": runtime.to_host(pre), "samples": runtime.to_host(w.samples),
                "lam_prev": runtime.to_host(self.lam_prev) if self.lam_prev is not None else np.zeros(0)}

    def restore(self, st):
        """Continue from ``state()`` of a sampler over the same normals and uniforms."""
        w = self.walk
        self.next_site = int(st["next_site"])
        self.rows = int(st["rows"])
        if self.rows > self.chi_max:
            raise ValueError("checkpoint rank beyond chi_max")
        w.samples.copy_(torch.from_numpy(np.ascontiguousarray(st["samples"])))
        w.cur = 0
        pre = np.ascontiguousarray(st["pre"])
        if pre.size:
            w.pre[0][:pre.size].copy_(torch.from_numpy(pre))
        lam = np.asarray(st["lam_prev"], dtype=np.float64)
        self.lam_prev = runtime.to_device(lam) if lam.size else None

    @property
    def samples(self):
        return self.walk.samples

    @property
    def probs(self):
        return self.walk.probs

    def mid_uniforms_row(self, i):
        """Uniforms [n] in the middle of each sample's cdf interval at site i (the draw is the
        number of cdf entries below the uniform), so another form replaying site i with them
        repeats these draws unless its probabilities differ by half an interval."""
        if self.walk.probs is None:
            raise ValueError("mid_uniforms_row needs return_probs=True")
        p = self.walk.probs[i]
        idx = self.walk.samples[:, i].long().unsqueeze(1)
        cdf = torch.cumsum(p, dim=1)
        return (cdf.gather(1, idx) - 0.5 * p.gather(1, idx)).squeeze(1).contiguous()


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
    # Lambda_i as float64 whatever dtype the files hold (the reference saves float32)
    lambdas = [np.load(path + f"Lambda_{i}.npy").astype(np.float64) for i in range(M - 1)] + [None]
    for it in range(args.iter):
        chunks = []
        for b0 in range(0, args.N, args.n):
            nb = min(args.N, b0 + args.n) - b0
            normals, uniforms = random_inputs(M, nb, args.seed + it, b0)
            walker = SiteSampler(M, args.dd, sqrtW, normals, uniforms, args.chi, args.d, chunk=nb,
                                 return_probs=False)
            for i in range(M):
                # one site tensor at a time, as sampling_cpu.py:76 loads Gamma_i per site; the
                # memory map reads only the live block (complex64 files are widened on the device)
                walker.push(i, np.load(path + f'Gamma_{i}.npy', mmap_mode='r'), lambdas[i])
            chunks.append(runtime.to_host(walker.samples).astype(np.int64))
            np.save(args.dir + f"samples_{it}.npy", np.concatenate(chunks, axis=0))


if __name__ == "__main__":
    main()
