"""Site-major construction and sampling: each site tensor is built, every sample is advanced
through it, and it is dropped before the next site is built.

This is the reference's per-site loop turned around.  MPS_cpu.py writes every Gamma_k to disk
(:523-525), and sampling_cpu.py then reads each Gamma_k back inside its per-site loop (:76-80).
At M = 144 and chi = 10^4 a site tensor is 16 GB, so the chain never exists as a whole on one
GPU.  The device holds one site tensor, its reduced-precision copies, and the carried vectors
of all samples.

Several forms walk the chain together:

* The first form ("fp64" by default) draws with the Philox uniforms.
* Every other form replays the first form's draws. At each site it takes the uniform in the
  middle of the first form's cdf interval. Its own probabilities then decide the same photon
  number unless they differ by half an interval. Such flips are counted. Without replay, a
  difference of one ULP near a cdf boundary would send the two forms down different paths.
  Their marginals would then differ by sampling noise, not arithmetic.
* A form may keep only a sub-block of each site tensor: leading ranks and photon cutoffs from
  a learned plan. Gamma's entries do not depend on which other labels are kept, so the
  sub-block of the pruned tensor is exactly the tensor that plan would build.

Per site and form the driver keeps the Rao-Blackwellised statistics. They are the mean
conditional distribution and each sample's conditional mean. The chain can stop after any
site and resume from a checkpoint. The checkpoint holds the carried vectors, the draws and
those statistics, so a chain longer than one session is built in several.
"""
import time

import numpy as np
import torch

from . import kron as kron_mod, mps, runtime, sampling


class Form:
    """One way of walking the chain: a sampling precision, and optionally per-cut ranks and
    per-site cutoffs (a learned plan) that select a leading sub-block of every site tensor."""

    def __init__(self, name, precision="fp64", site_chi=None, site_d=None):
        self.name, self.precision = name, precision
        self.site_chi = None if site_chi is None else [int(r) for r in site_chi]
        self.site_d = None if site_d is None else [int(x) for x in site_d]


class Chain:
    """Construct-and-sample driver of one instance (sq_cov, cov) at cutoff d, bond dimension chi
    and displaced cutoff dd, for n samples with Philox seed ``seed``.  ``forms[0]`` draws; the
    others replay its draws.  Site tensors are built in the pruned plain form
    (mps.build_mps(pruned=True))."""

    def __init__(self, sq_cov, cov, d, chi, dd, n, seed=0, forms=(Form("fp64"),), chunk=4096, kron=None):
        self.sq_cov, self.cov = sq_cov, cov
        self.M = len(sq_cov) // 2
        self.d, self.chi, self.dd, self.n, self.seed = d, chi, dd, n, seed
        self.forms = list(forms)
        self.kron = kron if kron is not None else kron_mod.kron_all(sq_cov, d, chi)
        self.ranks = [int((np.asarray(r) > mps.ERR_TOL).sum()) for r in self.kron[0]]
        self.sqrtW = sampling.thermal_sqrtW(cov, sq_cov)
        normals, uniforms = sampling.random_inputs(self.M, n, seed)
        chi_max = max(self.ranks) if self.ranks else 1
        self.walkers = [sampling.SiteSampler(self.M, dd, self.sqrtW, normals, uniforms, chi_max, d,
                                             precision=f.precision, chunk=chunk, return_probs=True)
                        for f in self.forms]
        M = self.M
        self.mean_probs = np.zeros((len(self.forms), M, dd))   # E_s p_k(j | s)
        self.cond_mean = np.zeros((len(self.forms), n, M))     # E[n_k | s's draws before k]
        self.off_path = np.zeros((len(self.forms), M), dtype=np.int64)
        self.site_seconds = np.zeros(M)
        self.next_site = 0

    # -- walking ------------------------------------------------------------------------------
    def _sub(self, f, k, G, L):
        """Form f's block of site k: Gamma[:r_{k-1}, :r_k, :d_k] and Lambda[:r_k]."""
        if f.site_chi is None and f.site_d is None:
            return G, L
        rows = f.site_chi[k - 1] if (k > 0 and f.site_chi is not None) else G.shape[0]
        cols = f.site_chi[k] if (k < self.M - 1 and f.site_chi is not None) else G.shape[1]
        dk = f.site_d[k] if f.site_d is not None else G.shape[2]
        lam = None if L is None else np.asarray(L)[:cols]
        return G[:rows, :cols, :dk], lam

    def push(self, k, G, L):
        """Advance every form's samples through site k (Gamma_k on the device or host)."""
        ref = self.walkers[0]
        u = None
        for fi, (f, w) in enumerate(zip(self.forms, self.walkers)):
            g, lam = self._sub(f, k, G, L)
            w.push(k, g, lam if k < self.M - 1 else None, uniforms_row=u)
            if fi == 0:
                u = ref.mid_uniforms_row(k)
            else:
                self.off_path[fi, k] = int((w.samples[:, k] != ref.samples[:, k]).sum())
            p = w.probs[k]
            self.mean_probs[fi, k] = runtime.to_host(p.mean(dim=0))
            j = torch.arange(self.dd, dtype=p.dtype, device=p.device)
            self.cond_mean[fi, :, k] = runtime.to_host((p * j).sum(dim=1))
        self.next_site = k + 1

    def run(self, sites=None, on_site=None, right_runs=None, on_run=None):
        """Build and walk ``sites`` (default: from the next site to the end; they must start at
        the next site and be consecutive).  ``on_site(k, seconds)`` reports each finished site."""
        sites = list(range(self.next_site, self.M)) if sites is None else list(sites)
        if sites and (sites[0] != self.next_site or sites != list(range(sites[0], sites[-1] + 1))):
            raise ValueError(f"sites must run consecutively from site {self.next_site}")
        clock = [time.perf_counter()]

        def consume(k, G, L):
            self.push(k, G, L)
            torch.cuda.synchronize()
            t = time.perf_counter()
            self.site_seconds[k] = t - clock[0]
            clock[0] = t
            if on_site is not None:
                on_site(k, self.site_seconds[k])

        mps.build_mps(self.sq_cov, self.d, self.chi, kron=self.kron, sites=sites, pruned=True, consume=consume,
                      right_runs=right_runs, on_run=on_run)

    # -- results ------------------------------------------------------------------------------
    def samples(self, fi=0):
        return runtime.to_host(self.walkers[fi].samples[:, :self.next_site]).astype(np.int64)

    def first_order(self, fi):
        """Rao-Blackwellised E[n_k] of form fi over the sites walked so far."""
        return (self.mean_probs[fi, :self.next_site] * np.arange(self.dd)).sum(axis=1)

    def second_order(self, fi):
        """Rao-Blackwellised E[n_i n_j], i < j, of form fi: mean over samples of
        n_i(s) E[n_j | s's earlier draws], with the draws of the first form (every replaying
        form follows them)."""
        K = self.next_site
        s = self.samples(0).astype(np.float64)[:, :K]
        full = s.T @ self.cond_mean[fi, :, :K] / self.n
        return full[np.triu_indices(K, k=1)]

    def errors(self):
        """Largest relative difference of every replaying form's first- and second-order marginals
        from the first form's, the largest absolute difference of the single-mode distributions,
        and the draws that left the first form's paths."""
        out = {}
        f0, s0 = self.first_order(0), self.second_order(0)
        for fi in range(1, len(self.forms)):
            f1, s1 = self.first_order(fi), self.second_order(fi)
            big = np.abs(s0) > 1e-6 * np.abs(s0).max() if len(s0) else np.zeros(0, bool)
            out[self.forms[fi].name] = {
                "first_order_max_rel_err": float(np.max(np.abs(f1 - f0) / np.abs(f0))) if len(f0) else 0.0,
                "second_order_max_rel_err": float(np.max(np.abs(s1 - s0)[big] / np.abs(s0[big]))) if big.any() else 0.0,
                "single_mode_marginal_max_abs_err": float(np.abs(self.mean_probs[fi, :self.next_site]
                                                                 - self.mean_probs[0, :self.next_site]).max()),
                "off_path_draws": int(self.off_path[fi, :self.next_site].sum()),
            }
        return out

    # -- checkpoints --------------------------------------------------------------------------
    def state(self):
        """Host arrays of everything needed to continue at the next site."""
        st = {"next_site": np.array(self.next_site), "n": np.array(self.n), "seed": np.array(self.seed),
              "mean_probs": self.mean_probs, "cond_mean": self.cond_mean, "off_path": self.off_path,
              "site_seconds": self.site_seconds,
              "forms": np.array([f"{f.name}:{f.precision}" for f in self.forms])}
        for fi, f in enumerate(self.forms):   # a learned plan travels with the checkpoint
            if f.site_chi is not None:
                st[f"plan{fi}_site_chi"] = np.asarray(f.site_chi)
            if f.site_d is not None:
                st[f"plan{fi}_site_d"] = np.asarray(f.site_d)
        for fi, w in enumerate(self.walkers):
            for key, v in w.state().items():
                st[f"w{fi}_{key}"] = v
        return st

    def save(self, path):
        np.savez(path, **self.state())

    @staticmethod
    def forms_of(path):
        """The forms a checkpoint was walked with (learned plans included), so every part of a
        chain walks the same forms."""
        st = np.load(path, allow_pickle=False)
        out = []
        for fi, spec in enumerate(st["forms"]):
            name, precision = str(spec).split(":")
            out.append(Form(name, precision,
                            site_chi=st[f"plan{fi}_site_chi"] if f"plan{fi}_site_chi" in st.files else None,
                            site_d=st[f"plan{fi}_site_d"] if f"plan{fi}_site_d" in st.files else None))
        return out

    def restore(self, path):
        st = np.load(path, allow_pickle=False)
        if int(st["n"]) != self.n or int(st["seed"]) != self.seed:
            raise ValueError("checkpoint of a different sample set")
        if list(st["forms"]) != [f"{f.name}:{f.precision}" for f in self.forms]:
            raise ValueError("checkpoint of different forms")
        self.next_site = int(st["next_site"])
        self.mean_probs[...] = st["mean_probs"]
        self.cond_mean[...] = st["cond_mean"]
        self.off_path[...] = st["off_path"]
        self.site_seconds[...] = st["site_seconds"]
        for fi, w in enumerate(self.walkers):
            w.restore({key[len(f"w{fi}_"):]: st[key] for key in st.files if key.startswith(f"w{fi}_")})
