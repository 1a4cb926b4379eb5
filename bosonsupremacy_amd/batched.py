"""Batched form: many independent instances through construction and sampling.

Every instance goes through the pipelined construction (host FP64 work in the worker pool,
hafnians on the GPU).  Its sampling pass -- per site, the contraction and the displacement +
draw kernels -- is captured once as a HIP graph on a side stream and replayed for every batch
of samples, so a pass is one submission however many sites the instance has; the random
numbers for a batch are generated into the graph's static input buffers before each replay.
Launch shapes come from the project's autotuner (``autotune.load_or_tune``).

``run_batch`` times, per instance, the construction and one graph-replayed sampling pass in
each of the FP64 and FP32 contraction forms (instances/s for either form), then measures the
FP32 form's error against FP64 on the same sample paths: the FP32 pass is replayed with each
uniform moved to the middle of the FP64 cumulative-probability interval of the FP64 draw, so
both passes follow the FP64 samples (unless the two cdfs differ by half a probability, which
is counted), and the Rao-Blackwellised single-mode marginals P(n_k = j) -- the mean over
samples of the conditional distributions -- are compared.  It also reports how many draws
differ when both forms use the same uniforms.
"""
import sys
import time

import numpy as np
import torch

from . import mps, runtime, sampling


class GraphSampler:
    """A Sampler whose pass over all sites is captured as a HIP graph for a fixed batch size n.

    The graph is recorded on a side stream with its own library context (a context enqueues on
    the stream current at its creation); static buffers hold the random inputs and outputs."""

    def __init__(self, gammas, lambdas, sqrtW, dd, n, precision="fp64", return_probs=False):
        self.stream = torch.cuda.Stream()
        with torch.cuda.stream(self.stream):
            self.ctx = runtime.Context(runtime.context().device.index)
        self.ctx.copy_settings(runtime.context())
        self.sampler = sampling.Sampler(gammas, lambdas, sqrtW, dd, precision=precision, ctx=self.ctx)
        M = self.sampler.M
        dev = self.sampler.dev
        self.n = n
        self.normals = torch.zeros((2 * M, n), dtype=torch.float64, device=dev)
        self.uniforms = torch.zeros((M, n), dtype=torch.float64, device=dev)
        self.return_probs = return_probs
        torch.cuda.synchronize()
        with torch.cuda.stream(self.stream):  # load the kernels outside capture (small batch)
            w = min(n, 256)
            self.sampler.run(self.normals[:, :w], self.uniforms[:, :w], return_probs=return_probs)
        self.stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self.out = self.sampler.run(self.normals, self.uniforms, return_probs=return_probs)
        torch.cuda.synchronize()

    def run(self, normals, uniforms):
        """Replay on new inputs (device [2M, n] and [M, n]); returns the static outputs."""
        self.normals.copy_(normals)
        self.uniforms.copy_(uniforms)
        self.graph.replay()
        return self.out


def _mid_uniforms(samples, probs):
    """Uniforms [M, n] in the middle of each draw's cumulative-probability interval (the draw
    is the number of cdf entries below the uniform), so a pass with them repeats ``samples``."""
    idx = samples.T.long().unsqueeze(2)                                  # [M, n, 1]
    cdf = torch.cumsum(probs, dim=2)
    return (cdf.gather(2, idx) - 0.5 * probs.gather(2, idx)).squeeze(2).contiguous()


def run_batch(instances, n_samples=10 ** 4, verbose=False):
    """Construct and sample every instance; returns a summary dict (progress lines go to
    stderr when ``verbose``)."""
    from ._lib import BsmpsError

    def now():
        torch.cuda.synchronize()
        return time.perf_counter()
    t_start = time.perf_counter()
    tot = {"construction": 0.0, "setup_fp64": 0.0, "setup_fp32": 0.0, "sampling_fp64": 0.0, "sampling_fp32": 0.0}
    abs_errs, rows, skipped = [], [], []
    for i, inst in enumerate(instances):
        t0 = now()
        try:
            G, L, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi)
        except BsmpsError as e:  # e.g. hafnians beyond the supported size
            torch.cuda.synchronize()
            skipped.append({"index": i, "M": inst.M, "d": inst.d, "chi": inst.chi, "reason": str(e)})
            if verbose:
                print(f"[batched] instance {i} skipped: {e}", file=sys.stderr, flush=True)
            continue
        t1 = now()
        sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
        g64 = GraphSampler(G, L, sqrtW, inst.dd, n_samples, "fp64", return_probs=True)
        t2 = now()
        g32 = GraphSampler(G, L, sqrtW, inst.dd, n_samples, "fp32", return_probs=True)
        t3 = now()
        normals, uniforms = sampling.random_inputs(inst.M, n_samples, seed=inst.seed)
        t4 = now()
        s64, p64 = g64.run(normals, uniforms)
        t5 = now()
        s32, _ = g32.run(normals, uniforms)
        t6 = now()
        flips = int((s32 != s64).any(dim=1).sum())          # samples whose draws differ (same uniforms)
        s32, p32 = g32.run(normals, _mid_uniforms(s64, p64))  # FP32 along the FP64 sample paths
        off_path = int((s32 != s64).any(dim=1).sum())
        abs_err = float(torch.abs(p32.mean(dim=1) - p64.mean(dim=1)).max())
        times = {"construction": t1 - t0, "setup_fp64": t2 - t1, "setup_fp32": t3 - t2,
                 "sampling_fp64": t5 - t4, "sampling_fp32": t6 - t5}
        for k, v in times.items():
            tot[k] += v
        abs_errs.append(abs_err)
        rows.append({"M": inst.M, "d": inst.d, "chi": inst.chi, **{k + "_s": v for k, v in times.items()},
                     "fp32_marginal_abs_err": abs_err, "fp32_draws_differing_same_uniforms": flips,
                     "fp32_off_fp64_path": off_path})
        if verbose:
            print(f"[batched] instance {i}: M={inst.M} d={inst.d} chi={inst.chi} construction {t1 - t0:.2f} s "
                  f"sampling fp64 {t5 - t4:.3f} s fp32 {t6 - t5:.3f} s marginal err {abs_err:.2e} "
                  f"differing draws {flips}/{n_samples}", file=sys.stderr, flush=True)
        del G, L, g64, g32
    wall = time.perf_counter() - t_start
    done = len(rows)
    t64 = tot["construction"] + tot["setup_fp64"] + tot["sampling_fp64"]
    t32 = tot["construction"] + tot["setup_fp32"] + tot["sampling_fp32"]
    return {"instances": done, "skipped": skipped, "samples_per_instance": n_samples, "wall_s": wall,
            "instances_per_s_fp64": done / t64, "instances_per_s_fp32": done / t32,
            **{k + "_s": v for k, v in tot.items()},
            "fp32_marginal_abs_err_max": float(np.max(abs_errs)),
            "fp32_marginal_abs_err_median": float(np.median(abs_errs)),
            "fp32_off_fp64_path_total": int(sum(r["fp32_off_fp64_path"] for r in rows)),
            "fp32_draws_differing_same_uniforms_total": int(sum(r["fp32_draws_differing_same_uniforms"] for r in rows)),
            "per_instance": rows}
