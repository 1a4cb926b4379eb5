"""Launch-shape autotuner for the hafnian kernels.

For every even hafnian size it times the candidate launch shapes -- one lane per matrix
(sizes <= 10), 16 / 32 / 64 cooperating lanes per matrix, and for sizes 12..34 the
frontier-split form (upper recursion tree breadth-first in HBM, one lane per subtree below)
-- on synthetic targets drawn from a random Sigma, and keeps the fastest.  It also times the
two forms of the site contraction (complex product as one real GEMM of 2x2 blocks, or as
three real products) for FP64 and FP32 at a sampling-sized shape and keeps the faster.
Every candidate computes bit-identical results (the tests check all shapes against the same
oracle), so the choice only affects speed.  Results are cached per device name in
``~/.cache/bsmps/tuning.json`` (or $BSMPS_TUNING_CACHE) and applied to the process-wide
context.  A hafnian choice is "lane", "group:G" or "split"; a contraction choice
("contract_fp64", "contract_fp32") is "blocks" or "gauss".
"""
import json
import os

import numpy as np
import torch

from . import runtime

GROUP_CANDIDATES = (16, 32, 64)
LANE_MAX_SIZE = 10
SPLIT_SIZES = range(12, 35, 2)
CACHE_VERSION = "v5"
CONTRACT_KEYS = {"contract_fp64": 0, "contract_fp32": 1}


def _cache_path():
    return os.environ.get("BSMPS_TUNING_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "bsmps", "tuning.json"))


def _time_call(fn, reps=2):
    fn()  # warm (plan build, code load)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune(sizes=(6, 8, 10, 12, 14, 16, 18), ld=64, items=None, seed=0, verbose=False):
    """Time every candidate per size; returns {n: "lane" | "group:G" | "split"}."""
    ctx = runtime.context()
    dev = ctx.device
    rng = np.random.default_rng(seed)
    S = (rng.normal(size=(ld, ld)) + 1j * rng.normal(size=(ld, ld))) * 0.3
    S_d = runtime.to_device(S)
    best = {}
    for n in sizes:
        count = items or max(2000, int(2e9 / max(1, _flops(n))))
        tg = torch.from_numpy(rng.integers(0, ld, size=(count, n)).astype(np.int32)).to(dev)
        out = torch.empty((count,), dtype=torch.complex128, device=dev)

        def run():
            ctx.call("bsmps_hafnian_targets", runtime.ptr(S_d), ld, runtime.ptr(tg), count, n, None,
                     runtime.ptr(out))
        timings = {}
        if n <= LANE_MAX_SIZE:
            ctx.set_tuning(n // 2, 64)
            timings["lane"] = _time_call(run)
        ctx.set_tuning(min(2, n // 2 - 1) if n >= 6 else 2, 64)
        if n in SPLIT_SIZES:
            ctx.call("bsmps_set_split", n, 1)
            timings["split"] = _time_call(run)
            ctx.call("bsmps_set_split", n, 0)
        for g in GROUP_CANDIDATES:
            try:
                ctx.call("bsmps_set_group_lanes", n, g)
                timings[f"group:{g}"] = _time_call(run)
            except Exception:
                continue
        best[n] = min(timings, key=timings.get)
        if verbose:
            print(f"n={n} items={count} " + " ".join(f"{k}:{v:.2f}ms" for k, v in timings.items()), flush=True)
    return best


def tune_contract(n=16384, chi=512, d=10, verbose=False):
    """Time both site-contraction forms per precision on a [n, chi] x [chi, chi, d] product;
    returns {"contract_fp64": form, "contract_fp32": form}."""
    ctx = runtime.context()
    dev = ctx.device
    gen = torch.Generator(device=dev).manual_seed(0)
    best = {}
    for key, prec in CONTRACT_KEYS.items():
        dt = torch.complex128 if prec == 0 else torch.complex64
        entry = "bsmps_site_contract" if prec == 0 else "bsmps_site_contract_c64"
        pre = torch.randn((n, chi), dtype=dt, device=dev, generator=gen)
        lam = torch.rand((chi,), dtype=torch.float64, device=dev, generator=gen)
        G = torch.randn((chi, chi, d), dtype=dt, device=dev, generator=gen)
        T = torch.empty((n, chi, d), dtype=dt, device=dev)

        def run():
            ctx.call(entry, n, chi, chi, d, runtime.ptr(pre), runtime.ptr(lam), runtime.ptr(G), runtime.ptr(T))
        timings = {}
        for form in ("blocks", "gauss"):
            ctx.call("bsmps_set_contract_form", prec, int(form == "gauss"))
            timings[form] = _time_call(run)
        best[key] = min(timings, key=timings.get)
        ctx.call("bsmps_set_contract_form", prec, int(best[key] == "gauss"))
        if verbose:
            print(f"{key}: " + " ".join(f"{k}:{v:.2f}ms" for k, v in timings.items()), flush=True)
        del pre, lam, G, T
    return best


def apply(best):
    """Install a tuning {n: "lane" | "group:G" | "split", "contract_*": "blocks" | "gauss"} on
    the process-wide context."""
    ctx = runtime.context()
    for key, prec in CONTRACT_KEYS.items():
        if key in best:
            ctx.call("bsmps_set_contract_form", prec, int(best[key] == "gauss"))
    best = {n: v for n, v in best.items() if not isinstance(n, str)}
    lane_sizes = [n for n, v in best.items() if v == "lane"]
    lane_max_m = 2
    for n in sorted(lane_sizes):
        if all(best.get(k) == "lane" for k in range(6, n + 1, 2) if k in best):
            lane_max_m = max(lane_max_m, n // 2)
    ctx.set_tuning(lane_max_m, 64)
    for n, v in best.items():
        n = int(n)
        if v == "lane":
            continue
        if n in SPLIT_SIZES:
            ctx.call("bsmps_set_split", n, int(v == "split"))
        if v.startswith("group:"):
            ctx.call("bsmps_set_group_lanes", n, int(v.split(":")[1]))


def load_or_tune(verbose=False):
    """Apply the cached tuning for this device, tuning first if there is none."""
    name = torch.cuda.get_device_name(runtime.context().device)
    path = _cache_path()
    cache = {}
    if os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    entry = cache.get(f"{name}/{CACHE_VERSION}")
    if entry is None:
        best = tune(verbose=verbose)
        best.update(tune_contract(verbose=verbose))
        entry = {str(k): v for k, v in best.items()}
        cache[f"{name}/{CACHE_VERSION}"] = entry
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(cache, f, indent=1)
        except OSError:
            pass
    best = {int(k) if k.isdigit() else k: v for k, v in entry.items()}
    apply(best)
    return best


def _flops(n):
    m = n // 2
    return {0: 1, 1: 1, 2: 22, 3: 336, 4: 1520, 5: 6792, 6: 26128, 7: 90888, 8: 291064, 9: 891632,
            10: 2.6e6, 11: 7.5e6, 12: 2.1e7}.get(m, 1e8)
