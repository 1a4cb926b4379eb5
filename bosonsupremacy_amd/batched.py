"""Batched form: many independent instances through construction and sampling.

Every instance goes through the pipelined construction (host FP64 work in the worker pool,
hafnians on the GPU).  Its sampling pass -- per site, the contraction and the displacement +
draw kernels -- is captured once as a HIP graph on a side stream and replayed for every batch
of samples, so a pass is one submission however many sites the instance has; the random
numbers for a batch are generated into the graph's static input buffers before each replay.
Launch shapes come from the project's autotuner (``autotune.load_or_tune``).

``run_batch`` times, per instance, the construction and one graph-replayed sampling pass in
the FP64 form and in each reduced-precision form (FP32, BF16x6, FP16x3 by default; instances/s for
each), then measures every reduced form's error against FP64 on the same sample paths: its
pass is replayed with each uniform moved to the middle of the FP64 cumulative-probability
interval of the FP64 draw, so both passes follow the FP64 samples (unless the two cdfs differ
by half a probability, which is counted), and the Rao-Blackwellised single-mode marginals
P(n_k = j) -- the mean over samples of the conditional distributions -- are compared.  It also
reports how many draws differ when both forms use the same uniforms.
"""
import sys
import time

import numpy as np
import torch

from . import metrics, mps, runtime, sampling


class GraphSampler:
    """A Sampler whose pass over all sites is captured as a HIP graph for a fixed batch size n.

    The graph is recorded on a side stream with its own library context (a context enqueues on
    the stream current at its creation); static buffers hold the random inputs and outputs, and
    the graph keeps its own memory pool.  ``prepare_s`` times the device preparation of the
    site tensors (the form's conversions and splits), ``capture_s`` the warm-up pass (kernels
    launched outside capture) and the capture."""

    def __init__(self, gammas, lambdas, sqrtW, dd, n, precision="fp64", return_probs=False):
        import time
        self.stream = torch.cuda.Stream()
        with torch.cuda.stream(self.stream):
            self.ctx = runtime.Context(runtime.context().device.index)
        self.ctx.copy_settings(runtime.context())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.sampler = sampling.Sampler(gammas, lambdas, sqrtW, dd, precision=precision, ctx=self.ctx)
        M = self.sampler.M
        dev = self.sampler.dev
        self.n = n
        self.normals = torch.zeros((2 * M, n), dtype=torch.float64, device=dev)
        self.uniforms = torch.zeros((M, n), dtype=torch.float64, device=dev)
        self.return_probs = return_probs
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        with torch.cuda.stream(self.stream):  # load the kernels outside capture (small batch)
            w = min(n, 256)
            self.sampler.run(self.normals[:, :w], self.uniforms[:, :w], return_probs=return_probs)
        self.stream.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=self.stream):
            self.out = self.sampler.run(self.normals, self.uniforms, return_probs=return_probs)
        torch.cuda.synchronize()
        self.prepare_s, self.capture_s = t1 - t0, time.perf_counter() - t1

    def run(self, normals, uniforms):
        """Replay on new inputs (device [2M, n] and [M, n]); returns the static outputs."""
        self.normals.copy_(normals)
        self.uniforms.copy_(uniforms)
        self.graph.replay()
        return self.out


def mid_uniforms(samples, probs):
    """Uniforms [M, n] in the middle of each draw's cumulative-probability interval (the draw
    is the number of cdf entries below the uniform), so a pass with them repeats ``samples``."""
    idx = samples.T.long().unsqueeze(2)                                  # [M, n, 1]
    cdf = torch.cumsum(probs, dim=2)
    return (cdf.gather(2, idx) - 0.5 * probs.gather(2, idx)).squeeze(2).contiguous()


def run_batch(instances, n_samples=10 ** 4, forms=("fp32", "bf16x6", "fp16x3"), verbose=False,
              learned_model=None, learned_precision="fp16x3"):
    """Construct and sample every instance; returns a summary dict (progress lines go to
    stderr when ``verbose``).  Each form in ``forms`` is timed and checked against FP64.

    With ``learned_model`` (learned.MLP) every instance is also run in the learned-cutoff form:
    its spectra, the exactly judged plan (learned.plan_sites), the per-site-shaped MPS and one
    graph-replayed pass in ``learned_precision``, timed as a whole ("learned" instances/s) and
    checked along the FP64 sample paths of the plain MPS."""
    from ._lib import BsmpsError
    from . import kron as kron_mod, learned as learned_mod

    def now():
        torch.cuda.synchronize()
        return time.perf_counter()
    t_start = time.perf_counter()
    names = ("fp64",) + tuple(forms)
    tot = {"construction": 0.0}
    for f in names:
        tot["setup_" + f] = tot["sampling_" + f] = 0.0
    errs = {f: [] for f in forms}
    if learned_model is not None:
        tot.update({"construction_learned": 0.0, "setup_learned": 0.0, "sampling_learned": 0.0})
        errs["learned"] = []
    rows, skipped = [], []
    cap = {f: [0.0, 0.0] for f in names + (("learned",) if learned_model is not None else ())}
    for i, inst in enumerate(instances):
        t0 = now()
        try:
            G, L, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, pruned=True)
        except BsmpsError as e:  # e.g. hafnians beyond the supported size
            torch.cuda.synchronize()
            skipped.append({"index": i, "M": inst.M, "d": inst.d, "chi": inst.chi, "reason": str(e)})
            if verbose:
                print(f"[batched] instance {i} skipped: {e}", file=sys.stderr, flush=True)
            continue
        t1 = now()
        times = {"construction": t1 - t0}
        sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
        graphs = {}
        for f in names:
            ta = now()
            graphs[f] = GraphSampler(G, L, sqrtW, inst.dd, n_samples, f, return_probs=True)
            times["setup_" + f] = now() - ta
            cap[f][0] += graphs[f].prepare_s
            cap[f][1] += graphs[f].capture_s
        normals, uniforms = sampling.random_inputs(inst.M, n_samples, seed=inst.seed)
        row = {"M": inst.M, "d": inst.d, "chi": inst.chi}
        ta = now()
        s64, p64 = graphs["fp64"].run(normals, uniforms)
        times["sampling_fp64"] = now() - ta
        mid = mid_uniforms(s64, p64)
        m64 = p64.mean(dim=1)
        e1_64, e2_64 = metrics.rb_first_order(p64), metrics.rb_second_order(s64, p64)
        big = np.abs(e2_64) > 1e-3 * np.abs(e2_64).max()

        def rel_errs(sf, pf):
            """Largest relative differences of the Rao-Blackwellised first- and second-order
            marginals from FP64's (both along the FP64 sample paths)."""
            e1, e2 = metrics.rb_first_order(pf), metrics.rb_second_order(sf, pf)
            return (float(np.max(np.abs(e1 - e1_64) / np.abs(e1_64))),
                    float(np.max(np.abs(e2[big] - e2_64[big]) / np.abs(e2_64[big]))) if big.any() else 0.0)
        line = []
        for f in forms:
            ta = now()
            sf, _ = graphs[f].run(normals, uniforms)
            times["sampling_" + f] = now() - ta
            flips = int((sf != s64).any(dim=1).sum())          # samples whose draws differ (same uniforms)
            sf, pf = graphs[f].run(normals, mid)                # along the FP64 sample paths
            off_path = int((sf != s64).any(dim=1).sum())
            err = float(torch.abs(pf.mean(dim=1) - m64).max())
            r1, r2 = rel_errs(sf, pf)
            errs[f].append((err, r1, r2))
            row.update({f"{f}_marginal_abs_err": err, f"{f}_first_order_rel_err": r1, f"{f}_second_order_rel_err": r2,
                        f"{f}_draws_differing_same_uniforms": flips, f"{f}_off_fp64_path": off_path})
            line.append(f"{f} {times['sampling_' + f]:.3f} s err {err:.1e} flips {flips}")
        if learned_model is not None:
            ta = now()
            ko = kron_mod.kron_all(inst.sq_cov, inst.d, inst.chi)
            site_d, site_chi, plan = learned_mod.plan_sites(learned_model, inst.sq_cov, inst.cov, ko[0],
                                                            inst.d, inst.chi)
            Gl, Ll, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=ko, site_d=site_d, site_chi=site_chi)
            tb = now()
            gl = GraphSampler(Gl, Ll, sqrtW, inst.dd, n_samples, learned_precision, return_probs=True)
            cap["learned"][0] += gl.prepare_s
            cap["learned"][1] += gl.capture_s
            tc = now()
            gl.run(normals, uniforms)
            td = now()
            times.update({"construction_learned": tb - ta, "setup_learned": tc - tb, "sampling_learned": td - tc})
            sl, pl = gl.run(normals, mid)
            err = float(torch.abs(pl.mean(dim=1) - m64).max())
            r1, r2 = rel_errs(sl, pl)
            errs["learned"].append((err, r1, r2))
            row.update({"learned_marginal_abs_err": err, "learned_first_order_rel_err": r1,
                        "learned_second_order_rel_err": r2, "learned_off_fp64_path": int((sl != s64).any(dim=1).sum()),
                        "learned_mean_cutoff": plan["mean_cutoff"], "learned_mean_rank": plan["mean_rank"]})
            line.append(f"learned+{learned_precision} construction {tb - ta:.2f} s err {err:.1e}")
            del Gl, Ll, gl
        for k, v in times.items():
            tot[k] += v
        row.update({k + "_s": v for k, v in times.items()})
        rows.append(row)
        if verbose:
            print(f"[batched] instance {i}: M={inst.M} d={inst.d} chi={inst.chi} construction {t1 - t0:.2f} s "
                  f"fp64 {times['sampling_fp64']:.3f} s " + " ".join(line), file=sys.stderr, flush=True)
        del G, L, graphs
    wall = time.perf_counter() - t_start
    done = len(rows)
    out = {"instances": done, "skipped": skipped, "samples_per_instance": n_samples, "wall_s": wall}
    for f in names:
        out[f"instances_per_s_{f}"] = done / (tot["construction"] + tot["setup_" + f] + tot["sampling_" + f])
    out.update({k + "_s": v for k, v in tot.items()})
    if learned_model is not None:
        out[f"instances_per_s_learned_{learned_precision}"] = done / (
            tot["construction_learned"] + tot["setup_learned"] + tot["sampling_learned"])
        out["learned_marginal_abs_err_max"] = float(np.max([e[0] for e in errs["learned"]]))
        out["learned_marginal_abs_err_median"] = float(np.median([e[0] for e in errs["learned"]]))
        out["learned_first_order_rel_err_max"] = float(np.max([e[1] for e in errs["learned"]]))
        out["learned_second_order_rel_err_max"] = float(np.max([e[2] for e in errs["learned"]]))
        out["learned_off_fp64_path_total"] = int(sum(r["learned_off_fp64_path"] for r in rows))
    for f in forms:
        out[f"{f}_marginal_abs_err_max"] = float(np.max([e[0] for e in errs[f]]))
        out[f"{f}_marginal_abs_err_median"] = float(np.median([e[0] for e in errs[f]]))
        out[f"{f}_first_order_rel_err_max"] = float(np.max([e[1] for e in errs[f]]))
        out[f"{f}_second_order_rel_err_max"] = float(np.max([e[2] for e in errs[f]]))
        out[f"{f}_off_fp64_path_total"] = int(sum(r[f"{f}_off_fp64_path"] for r in rows))
        out[f"{f}_draws_differing_same_uniforms_total"] = int(
            sum(r[f"{f}_draws_differing_same_uniforms"] for r in rows))
    out["graph_prepare_s"] = {f: v[0] for f, v in cap.items()}
    out["graph_capture_s"] = {f: v[1] for f, v in cap.items()}
    out["per_instance"] = rows
    return out


class EagerPass:
    """A Sampler run launch by launch on the process context, with GraphSampler's interface:
    for a pass that is run once, capturing it costs more than its launches (the warm-up pass and
    the capture of an FP64 pass took ~100 ms per instance of the batched workload, its launches a
    few)."""

    def __init__(self, gammas, lambdas, sqrtW, dd, n, precision="fp64", return_probs=False):
        t0 = time.perf_counter()
        self.sampler = sampling.Sampler(gammas, lambdas, sqrtW, dd, precision=precision)
        torch.cuda.synchronize()
        self.return_probs = return_probs
        self.prepare_s, self.capture_s = time.perf_counter() - t0, 0.0

    def run(self, normals, uniforms):
        return self.sampler.run(normals, uniforms, return_probs=self.return_probs)


def _pass_and_errors(inst, G, L, sqrtW, names, n_samples, now, plan=None, learned_precision="fp16x3",
                     graphs=True):
    """One instance's sampling in every form of ``names`` (the first is FP64): a graph per form,
    one replay with the Philox uniforms (timed), and each reduced form replayed along the FP64
    sample paths (mid-interval uniforms).  ``plan`` = (site_d, site_chi) adds the learned-cutoff
    form: the leading blocks of the pruned tensors that the plan keeps, in ``learned_precision``.
    Returns (times, row, errors per form)."""
    from . import metrics
    times, row, errs = {}, {"M": inst.M, "d": inst.d, "chi": inst.chi}, {}
    forms = list(names)
    shapes = {f: (G, L) for f in forms}
    if plan is not None:
        site_d, site_chi = plan
        M = inst.M
        Gl, Ll = [], []
        for k in range(M):
            rows = site_chi[k - 1] if k > 0 else G[k].shape[0]
            cols = site_chi[k] if k < M - 1 else G[k].shape[1]
            Gl.append(G[k][:rows, :cols, :site_d[k]].contiguous())
            Ll.append(None if L[k] is None else np.asarray(L[k])[:cols])
        forms.append("learned")
        shapes["learned"] = (Gl, Ll)
    runner = GraphSampler if graphs else EagerPass
    graphs = {}
    for f in forms:
        ta = now()
        g, lam = shapes[f]
        graphs[f] = runner(g, lam, sqrtW, inst.dd, n_samples, learned_precision if f == "learned" else f,
                           return_probs=True)
        times["setup_" + f] = now() - ta
        times["prepare_" + f] = graphs[f].prepare_s
        times["capture_" + f] = graphs[f].capture_s
    normals, uniforms = sampling.random_inputs(inst.M, n_samples, seed=inst.seed)
    ta = now()
    s64, p64 = graphs[forms[0]].run(normals, uniforms)
    times["sampling_" + forms[0]] = now() - ta
    mid = mid_uniforms(s64, p64)
    e1_64, e2_64 = metrics.rb_first_order(p64), metrics.rb_second_order(s64, p64)
    big = np.abs(e2_64) > 1e-3 * np.abs(e2_64).max()
    m64 = p64.mean(dim=1)
    for f in forms[1:]:
        ta = now()
        sf, _ = graphs[f].run(normals, uniforms)
        times["sampling_" + f] = now() - ta
        flips = int((sf != s64).any(dim=1).sum())
        sf, pf = graphs[f].run(normals, mid)
        e1, e2 = metrics.rb_first_order(pf), metrics.rb_second_order(sf, pf)
        r1 = float(np.max(np.abs(e1 - e1_64) / np.abs(e1_64)))
        r2 = float(np.max(np.abs(e2[big] - e2_64[big]) / np.abs(e2_64[big]))) if big.any() else 0.0
        err = float(torch.abs(pf.mean(dim=1) - m64).max())
        errs[f] = (err, r1, r2)
        row.update({f"{f}_marginal_abs_err": err, f"{f}_first_order_rel_err": r1, f"{f}_second_order_rel_err": r2,
                    f"{f}_draws_differing_same_uniforms": flips,
                    f"{f}_off_fp64_path": int((sf != s64).any(dim=1).sum())})
    return times, row, errs


def run_batch_merged(instances, n_samples=10 ** 4, forms=("fp32", "bf16x6", "fp16x3"), group_bytes=32 << 30,
                     verbose=False, learned_model=None, learned_precision="fp16x3", graphs=True):
    """Batched form with cross-instance construction.  Instances are taken in groups whose site
    tensors together fit ``group_bytes``; the site tensors of a whole group -- every site of
    every instance in it -- are built by one ``bsmps_gamma_sites`` call, i.e. one hafnian launch
    per photon total over the (site, j, p, q) rectangles of all of them, whatever their cutoffs,
    ranks and hafnian sizes.  While the GPU builds group g the host prepares group g + 1: its
    Schmidt spectra on a side stream with a context of their own, its FP64 site kernels and
    label tables in the worker pool, its uploads and output tensors.  Each instance of group g
    is then sampled as in run_batch: one HIP graph per form (the pruned plain form's tensors,
    FP64 first), replayed once with the Philox uniforms and once along the FP64 sample paths,
    with relative errors of the Rao-Blackwellised marginals.  With ``learned_model`` each
    instance's exactly judged plan is walked as one more form on the leading blocks of its
    pruned tensors (the tensors that plan would build), so its construction is the pruned one.
    Instances whose hafnians exceed the library's sizes are skipped and listed.  ``graphs=False``
    runs each pass launch by launch (EagerPass) instead of capturing it.

    Timing: the sampling of each instance (graphs, passes, error replays) is timed as a block;
    construction is the rest of the wall time.  instances/s of a form = instances /
    (construction + that form's graph set-up and pass)."""
    from . import kron as kron_mod, hostpool, learned as learned_mod

    def now():
        torch.cuda.synchronize()
        return time.perf_counter()
    t_start = now()
    names = ("fp64",) + tuple(forms)
    all_forms = names + (("learned",) if learned_model is not None else ())
    tot = {f: 0.0 for f in ("check", "prepare", "merged_launch")}
    for f in all_forms:
        tot["setup_" + f] = tot["sampling_" + f] = tot["prepare_" + f] = tot["capture_" + f] = 0.0
    errs = {f: [] for f in all_forms[1:]}
    rows, skipped, group_sizes = [], [], []
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        kctx = runtime.Context(runtime.context().device.index)
    max_total = _lib_max_hafnian_size() + 1   # odd photon totals are never computed

    def prepare(i, inst):
        """Spectra (side stream), host work and uploads of one instance: (i, inst, calls, kron)
        or None when it is skipped."""
        cuts = hostpool.cut_inputs_all(inst.sq_cov, inst.d)
        with torch.cuda.stream(side):
            spec = kron_mod.kron_cuts([tp for tp, _ in cuts], inst.chi, ctx=kctx)
        kron = ([r for r, _ in spec], [n for _, n in spec], [S for _, S in cuts])
        lc = [int(np.asarray(n).reshape(len(np.asarray(r)), -1)[np.asarray(r) > mps.ERR_TOL].sum(axis=1).max(initial=0))
              for r, n in zip(kron[0], kron[1])]
        # the largest photon total of any site: d - 1 + (photons of a label of cut k) + (of cut k - 1)
        top = max([inst.d - 1 + (lc[k] if k < inst.M - 1 else 0) + (lc[k - 1] if k > 0 else 0)
                   for k in range(inst.M)])
        if top > max_total:
            skipped.append({"index": i, "M": inst.M, "d": inst.d, "chi": inst.chi,
                            "reason": f"hafnians up to size {top}"})
            return None
        calls, _ = mps.site_calls(inst.sq_cov, inst.d, inst.chi, kron=kron, pruned=True)
        return (i, inst, calls, kron)

    work = iter(enumerate(instances))
    carry = [None]

    def next_group():
        """The next group, prepared on the host (nothing waits for the main stream)."""
        t0 = time.perf_counter()
        group, nbytes = [], 0
        if carry[0] is not None:
            group.append(carry[0])
            nbytes = sum(c.gamma.numel() * 16 for c in carry[0][2])
            carry[0] = None
        for i, inst in work:
            item = prepare(i, inst)
            if item is None:
                continue
            b = sum(c.gamma.numel() * 16 for c in item[2])
            if group and nbytes + b > group_bytes:
                carry[0] = item
                break
            group.append(item)
            nbytes += b
        tot["prepare"] += time.perf_counter() - t0
        return group

    def launch(group):
        t0 = time.perf_counter()
        mps.gamma_sites([c for _, _, calls, _ in group for c in calls])
        tot["merged_launch"] += time.perf_counter() - t0
        group_sizes.append(len(group))

    def check(group):
        for i, inst, calls, kron in group:
            t0 = now()
            G = [c.gamma for c in calls]
            L = [c.Lambda for c in calls]
            plan = None
            if learned_model is not None:
                site_d, site_chi, _ = learned_mod.plan_sites(learned_model, inst.sq_cov, inst.cov, kron[0],
                                                             inst.d, inst.chi)
                plan = (site_d, site_chi)
            times, row, e = _pass_and_errors(inst, G, L, sampling.thermal_sqrtW(inst.cov, inst.sq_cov), names,
                                             n_samples, now, plan, learned_precision, graphs)
            for k, v in times.items():
                tot[k] += v
            for f, v in e.items():
                errs[f].append(v)
            row["index"] = i
            rows.append(row)
            tot["check"] += now() - t0
            if verbose:
                print(f"[batched] instance {i}: M={inst.M} d={inst.d} chi={inst.chi} "
                      + " ".join(f"{f} {times['sampling_' + f]:.3f} s" for f in all_forms), file=sys.stderr, flush=True)

    group = next_group()
    while group:
        launch(group)             # enqueued; the GPU builds this group ...
        nxt = next_group()        # ... while the host prepares the next one
        check(group)              # waits for the construction, then samples and checks
        del group
        group = nxt
    wall = now() - t_start
    done = len(rows)
    construction = wall - tot["check"]
    out = {"instances": done, "skipped": skipped, "samples_per_instance": n_samples, "wall_s": wall,
           "construction": "merged: one bsmps_gamma_sites call per group of instances, the next group "
                           "prepared on the host meanwhile",
           "passes": "one HIP graph per instance and form" if graphs else "launch by launch",
           "groups": len(group_sizes), "instances_per_group": group_sizes, "construction_s": construction,
           "host_prepare_s": tot["prepare"], "sampling_and_checks_s": tot["check"]}
    for f in all_forms:
        out[f"instances_per_s_{f}"] = done / (construction + tot["setup_" + f] + tot["sampling_" + f])
    out.update({k + "_s": v for k, v in tot.items() if k.startswith(("setup_", "sampling_"))})
    for f in all_forms[1:]:
        e = errs[f]
        out[f"{f}_marginal_abs_err_max"] = float(np.max([x[0] for x in e])) if e else 0.0
        out[f"{f}_first_order_rel_err_max"] = float(np.max([x[1] for x in e])) if e else 0.0
        out[f"{f}_second_order_rel_err_max"] = float(np.max([x[2] for x in e])) if e else 0.0
        out[f"{f}_off_fp64_path_total"] = int(sum(r.get(f"{f}_off_fp64_path", 0) for r in rows))
    out["graph_setup_s"] = {f: tot["setup_" + f] for f in all_forms}
    out["graph_setup_split_s"] = {f: {"form_preparation": tot["prepare_" + f], "warm_up_and_capture": tot["capture_" + f]}
                                  for f in all_forms}
    out["per_instance"] = rows
    return out


def _lib_max_hafnian_size():
    from . import _lib
    return _lib.MAX_HAFNIAN_SIZE
