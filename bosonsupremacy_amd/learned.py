"""Learned-cutoff form: per-site photon cutoffs and bond ranks predicted by a small MLP.

The plain method (MPS_cpu.py) uses one local photon cutoff d for every site and keeps every
Schmidt vector of weight > 1e-10 up to chi at every cut.  Many sites need less: a site
whose mode carries few photons needs a smaller cutoff, and cuts near the chain ends have
fast-decaying Schmidt spectra.  Skipping the entries j >= d_k removes the largest hafnians of
site k, and a rank r_k < chi shrinks both the hafnian grid and the sampling contraction.

A 3-layer MLP (16 -> 128 -> 128 -> 2, ReLU, BF16 weights, FP32 accumulation) maps local
features of site k -- the 4 x 4 pure-state covariance block of modes k, k+1, the classical
displacement variance and mean photon numbers of mode k, the site position and the leading
Schmidt weight of cut k -- to (needed cutoff, log2 of the needed rank fraction).  It is
trained on pairs generated by the plain method's own runs (Schmidt spectra on the GPU,
photon tails on the host) with this library's forward/backward/Adam kernels
(csrc/mlp.hip).

Every prediction is judged exactly before it is used (``plan_sites``):
  * cutoff: the weight the MPS loses by cutting site k at d_k instead of d is at most
    P(d_k <= n_k < d), the photon-number probability of mode k in the pure state, computed
    in closed form from its 2 x 2 reduced covariance (gaussian.photon_distribution); the
    bound holds because the Schmidt truncation only projects the state;
  * rank: the Schmidt weight dropped by keeping r_k instead of all kept vectors is the
    exact tail sum of the cut's spectrum.
Whenever a bound exceeds its tolerance, that site falls back to the plain cutoff (or rank).
"""
import os

import numpy as np
import torch

from . import gaussian, runtime

N_FEATURES = 16
WIDTH = 128
N_OUT = 2
# Tolerances of the exact checks.  Both kinds of truncation drop an amplitude, not a
# probability: the conditional distributions the sampler draws from involve modes on both sides
# of every earlier cut (the draws on the left, site k on the right), so a dropped Schmidt vector
# of weight delta moves them by O(sqrt(delta)) through its cross terms with the kept ones, and a
# dropped local component j >= d_k, mixed with the kept ones by the displacement D(alpha), does
# the same.  Measured along common sample paths: 1e-7 of dropped photon weight moved M=64
# first-order marginals by 4e-5, and ranks cut to 1e-7 of Schmidt weight per cut (23% less
# hafnian work at M=32) moved them by 1.3e-5.  A 1e-6 marginal bar therefore allows ~1e-12 of
# photon weight and ~1e-9 of Schmidt weight in all -- less than one kept label (every kept
# label weighs more than 1e-10) -- so TAU_CHI keeps every kept vector.  The network is trained
# to predict the ranks of the looser tolerance TAU_CHI_TRAIN; plan_sites(tau_chi=TAU_CHI_TRAIN)
# uses them, at a marginal error of ~1e-5.
TAU_D = 1e-12          # extra photon weight a site may drop by its cutoff
TAU_CHI = 1e-10        # extra Schmidt weight a cut may drop by its rank at the 1e-6 bar
TAU_CHI_TRAIN = 1e-7   # the tolerance the rank targets are computed at
D_SCALE = 10.0   # target scaling: cutoff / 10, log2(needed rank / kept rank) / 2
R_SCALE = 2.0
MARGIN_D = 0.5   # safety margins added to the predictions before rounding up (photons; log2 of rank)
MARGIN_R = 0.05


# ---------------------------------------------------------------------------------------
# features and labels
# ---------------------------------------------------------------------------------------

def site_features(sq_cov, cov, kres):
    """Features [M, 16] float32 of every site."""
    M = len(sq_cov) // 2
    W = cov - sq_cov
    iu = np.triu_indices(4)
    out = np.zeros((M, N_FEATURES), dtype=np.float32)
    for k in range(M):
        k1 = min(k + 1, M - 1)
        idx = [k, k1, k + M, k1 + M]
        f = list(sq_cov[np.ix_(idx, idx)][iu])                          # 10
        f += [W[k, k], W[k + M, k + M]]                                 # 2: displacement variance
        f += [(sq_cov[k, k] + sq_cov[k + M, k + M] - 2) / 4,            # 2: mean photons,
              (cov[k, k] + cov[k + M, k + M] - 2) / 4]                  #    pure and full state
        lead = float(np.max(kres[k])) if k < M - 1 and len(kres[k]) else 1.0
        f += [np.log10(max(1.0 - lead, 1e-16))]                         # 1: entanglement of cut k
        f += [_tail_feature(kres[k]) if k < M - 1 else 0.0]             # 1: weight near the kept end
        out[k] = f
    return out


def _tail_feature(res, err_tol=1e-10):
    """log10 of the Schmidt weight in the last tenth of the kept spectrum (how fast it ends)."""
    r = np.asarray(res, dtype=np.float64)
    r = r[r > err_tol]
    if len(r) == 0:
        return -16.0
    return float(np.log10(max(np.sum(r[int(0.9 * len(r)):]), 1e-16)))


def mode_covariance(sq_cov, k):
    M = len(sq_cov) // 2
    return sq_cov[np.ix_([k, k + M], [k, k + M])]


def needed_cutoff(sq_cov, k, d_cap, tau=TAU_D):
    """Smallest cutoff d' (<= d_cap) with P(n_k >= d') <= tau in the pure state (d_cap when
    even d_cap does not reach tau)."""
    p = gaussian.photon_distribution(mode_covariance(sq_cov, k), d_cap)
    tail = 1.0 - np.cumsum(p)
    ok = np.nonzero(tail <= tau)[0]
    return int(ok[0]) + 1 if len(ok) else d_cap


def kept_rank(res, err_tol=1e-10):
    return int((np.asarray(res) > err_tol).sum())


def needed_rank(res, tau=TAU_CHI, err_tol=1e-10):
    """Smallest r with sum(res[r:kept]) <= tau (res sorted descending, as the spectra are)."""
    r = np.asarray(res, dtype=np.float64)[:kept_rank(res, err_tol)]
    if len(r) == 0:
        return 0
    tail = np.concatenate([np.cumsum(r[::-1])[::-1], [0.0]])   # tail[i] = sum r[i:]
    return int(np.nonzero(tail <= tau)[0][0])


def site_labels(sq_cov, kres, chi, d_cap=12):
    """Targets [M, 2] float32: needed cutoff / D_SCALE, log2(needed rank / kept rank) / R_SCALE
    (the kept rank -- Schmidt weights above 1e-10, at most chi -- is what the pruned plain form
    keeps; the needed rank drops at most TAU_CHI_TRAIN more)."""
    M = len(sq_cov) // 2
    out = np.zeros((M, N_OUT), dtype=np.float32)
    for k in range(M):
        out[k, 0] = needed_cutoff(sq_cov, k, d_cap) / D_SCALE
        if k < M - 1:
            out[k, 1] = np.log2(max(needed_rank(kres[k], TAU_CHI_TRAIN), 1) / max(kept_rank(kres[k]), 1)) / R_SCALE
    return out


def training_pairs(instances, max_dim=10 ** 5):
    """(X, T) from the plain method's runs on a list of instances: the Schmidt spectra of
    every cut come from the GPU kron kernel, the photon tails from the closed form."""
    from . import kron
    Xs, Ts = [], []
    for inst in instances:
        kres, _, _ = kron.kron_all(inst.sq_cov, inst.d, inst.chi, max_dim)
        Xs.append(site_features(inst.sq_cov, inst.cov, kres))
        Ts.append(site_labels(inst.sq_cov, kres, inst.chi))
    return np.concatenate(Xs), np.concatenate(Ts)


# ---------------------------------------------------------------------------------------
# the network (parameters and activations on the GPU)
# ---------------------------------------------------------------------------------------

class MLP:
    """16 -> 128 -> 128 -> 2 ReLU network trained with Adam by the library's kernels."""

    def __init__(self, n_in=N_FEATURES, width=WIDTH, n_out=N_OUT, seed=0):
        self.ctx = runtime.context()
        self.dev = self.ctx.device
        self.n_in, self.width, self.n_out = n_in, width, n_out
        rng = np.random.default_rng(seed)
        shapes = [(width, n_in), (width,), (width, width), (width,), (n_out, width), (n_out,)]
        host = []
        for s in shapes:
            if len(s) == 2:  # He initialisation
                host.append((rng.normal(size=s) * np.sqrt(2.0 / s[1])).astype(np.float32))
            else:
                host.append(np.zeros(s, np.float32))
        self.params = [torch.from_numpy(h).to(self.dev) for h in host]
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.bf = [torch.empty(p.shape, dtype=torch.bfloat16, device=self.dev) for p in self.params[0::2]]
        self.mu = np.zeros(n_in, np.float32)
        self.sd = np.ones(n_in, np.float32)
        self.step = 0
        self._cast()

    def _cast(self):
        for w, b in zip(self.params[0::2], self.bf):
            self.ctx.call("bsmps_f32_to_bf16", w.numel(), runtime.ptr(w), runtime.ptr(b))

    def _norm(self, X):
        return torch.from_numpy(((np.asarray(X, np.float32) - self.mu) / self.sd).astype(np.float32)).to(self.dev)

    def _forward(self, Xd):
        n = Xd.shape[0]
        H1 = torch.empty((n, self.width), dtype=torch.float32, device=self.dev)
        H2 = torch.empty((n, self.width), dtype=torch.float32, device=self.dev)
        Y = torch.empty((n, self.n_out), dtype=torch.float32, device=self.dev)
        W1, W2, W3 = self.bf
        b1, b2, b3 = self.params[1::2]
        self.ctx.call("bsmps_mlp_forward", n, self.n_in, self.width, self.n_out, runtime.ptr(Xd), runtime.ptr(W1),
                      runtime.ptr(b1), runtime.ptr(W2), runtime.ptr(b2), runtime.ptr(W3), runtime.ptr(b3),
                      runtime.ptr(H1), runtime.ptr(H2), runtime.ptr(Y))
        return H1, H2, Y

    def _backward(self, Xd, H1, H2, dY):
        n = Xd.shape[0]
        grads = [torch.empty_like(p) for p in self.params]
        scratch = torch.empty((2 * n * self.width,), dtype=torch.float32, device=self.dev)
        W2, W3 = self.bf[1], self.bf[2]
        self.ctx.call("bsmps_mlp_backward", n, self.n_in, self.width, self.n_out, runtime.ptr(Xd), runtime.ptr(W2),
                      runtime.ptr(W3), runtime.ptr(H1), runtime.ptr(H2), runtime.ptr(dY), runtime.ptr(scratch),
                      *[runtime.ptr(g) for g in grads])
        return grads

    def fit(self, X, T, steps=3000, lr=2e-3, beta1=0.9, beta2=0.999, eps=1e-8, log_every=0):
        """Full-batch Adam on the mean squared error; returns the loss history (host)."""
        X = np.asarray(X, np.float32)
        self.mu = X.mean(axis=0)
        self.sd = X.std(axis=0) + 1e-6
        Xd = self._norm(X)
        Td = torch.from_numpy(np.asarray(T, np.float32)).to(self.dev)
        n = X.shape[0]
        dY = torch.empty((n, self.n_out), dtype=torch.float32, device=self.dev)
        err = torch.empty((n, self.n_out), dtype=torch.float32, device=self.dev)
        history = []
        for it in range(steps):
            H1, H2, Y = self._forward(Xd)
            self.ctx.call("bsmps_mse_grad", n, self.n_out, runtime.ptr(Y), runtime.ptr(Td), runtime.ptr(dY),
                          runtime.ptr(err))
            grads = self._backward(Xd, H1, H2, dY)
            self.step += 1
            for p, g, m, v in zip(self.params, grads, self.m, self.v):
                self.ctx.call("bsmps_adam_step", p.numel(), runtime.ptr(p), runtime.ptr(g), runtime.ptr(m),
                              runtime.ptr(v), lr, beta1, beta2, eps, self.step)
            self._cast()
            if log_every and (it % log_every == 0 or it == steps - 1):
                history.append((it, float(err.mean().item())))
        if not history:
            history.append((steps - 1, float(err.mean().item())))
        return history

    def predict(self, X):
        """Raw network outputs [n, 2] (host float32)."""
        _, _, Y = self._forward(self._norm(X))
        return runtime.to_host(Y)

    def state(self):
        d = {f"p{i}": runtime.to_host(p) for i, p in enumerate(self.params)}
        d.update(mu=self.mu, sd=self.sd)
        return d

    def load_state(self, state):
        for i, p in enumerate(self.params):
            p.copy_(torch.from_numpy(np.asarray(state[f"p{i}"], np.float32)).to(self.dev))
        self.mu = np.asarray(state["mu"], np.float32)
        self.sd = np.asarray(state["sd"], np.float32)
        self._cast()


MODEL_VERSION = 3  # bump when features, labels or tolerances change


def _cache_path():
    return os.environ.get("BSMPS_MLP_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "bsmps",
                                                          f"cutoff_mlp_v{MODEL_VERSION}.npz"))


TRAINING_WORKLOADS = ("m8_d4_chi64", "m32_d6_chi512", "m64_d8_chi2000")


def training_instances(seeds=range(1, 9)):
    """Instances of the three plain-method workloads with seeds other than the benchmark's
    seed 0, so the benchmarked instance is never seen in training."""
    from . import instances
    return [instances.workload(name, seed=s) for name in TRAINING_WORKLOADS for s in seeds]


def load_or_train(steps=3000, verbose=False):
    """The trained network: from the cache when present, else trained now and cached."""
    model = MLP()
    path = _cache_path()
    if os.path.exists(path):
        try:
            with np.load(path, allow_pickle=False) as z:
                model.load_state(z)
            return model
        except (OSError, ValueError, KeyError):
            pass
    X, T = training_pairs(training_instances())
    hist = model.fit(X, T, steps=steps, log_every=500 if verbose else 0)
    if verbose:
        print("mlp training loss", hist, flush=True)
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, **model.state())
    except OSError:
        pass
    return model


# ---------------------------------------------------------------------------------------
# prediction + exact judgement
# ---------------------------------------------------------------------------------------

def plan_sites(model, sq_cov, cov, kres, d, chi, tau_d=TAU_D, tau_chi=TAU_CHI):
    """Per-site cutoffs and per-cut ranks for the learned-cutoff form, each checked exactly.

    Returns (site_d [M], site_chi [M-1], report): a site whose predicted cutoff would drop more
    than tau_d photon weight beyond the plain cutoff d keeps d; a cut whose predicted rank would
    drop more than tau_chi Schmidt weight beyond the plain truncation keeps every kept vector."""
    M = len(sq_cov) // 2
    Y = model.predict(site_features(sq_cov, cov, kres))
    site_d, site_chi = [], []
    fall_d = fall_chi = 0
    extra_d, extra_chi = [], []
    for k in range(M):
        dk = int(np.clip(np.ceil(Y[k, 0] * D_SCALE + MARGIN_D), 1, d))
        p = gaussian.photon_distribution(mode_covariance(sq_cov, k), d)
        drop = float(np.sum(p[dk:d]))                 # P(dk <= n_k < d), exact
        if drop > tau_d:
            dk, drop, fall_d = d, 0.0, fall_d + 1
        site_d.append(dk)
        extra_d.append(drop)
    for k in range(M - 1):
        kept = kept_rank(kres[k])
        rk = int(np.clip(np.ceil(kept * 2.0 ** (Y[k, 1] * R_SCALE + MARGIN_R)), 1, kept))
        r = np.asarray(kres[k], dtype=np.float64)
        drop = float(np.sum(r[rk:kept]))               # exact Schmidt weight dropped
        if drop > tau_chi:
            rk, drop, fall_chi = kept, 0.0, fall_chi + 1
        site_chi.append(rk)
        extra_chi.append(drop)
    report = {"fallback_cutoffs": fall_d, "fallback_ranks": fall_chi,
              "max_extra_photon_weight": max(extra_d), "max_extra_schmidt_weight": max(extra_chi + [0.0]),
              "mean_cutoff": float(np.mean(site_d)), "mean_rank": float(np.mean(site_chi)) if site_chi else 0.0}
    return site_d, site_chi, report
