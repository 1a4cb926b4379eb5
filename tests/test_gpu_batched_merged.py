"""Cross-instance construction: bsmps_gamma_sites builds the site tensors of several instances in
one call (one hafnian launch per photon total over all their sites) with the same bits as one
bsmps_gamma_site call per site, and the merged batched driver samples them within the 1e-6 bar."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def small_batch(n, seed, chi_cap=128):
    from bosonsupremacy_amd import instances
    insts = instances.random_batch(n, seed=seed, n_samples=500)
    for inst in insts:  # keep the test small
        inst.chi = min(inst.chi, chi_cap)
    return insts


def test_gamma_sites_matches_per_site_calls():
    """Every site of three instances (uneven M, d, chi and hafnian sizes) built in one merged call
    equals the same site built alone, bit for bit; the calls run back to back on one stream."""
    from bosonsupremacy_amd import mps
    insts = small_batch(3, seed=21)
    merged, alone = [], []
    for inst in insts:
        calls, kron = mps.site_calls(inst.sq_cov, inst.d, inst.chi, pruned=True)
        merged.append(calls)
        G, _, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=kron, pruned=True)
        alone.append(G)
    mps.gamma_sites([c for calls in merged for c in calls])
    torch.cuda.synchronize()
    n_sites = 0
    for inst, calls, G in zip(insts, merged, alone):
        print(f"M={inst.M} d={inst.d} chi={inst.chi}", flush=True)
        assert len(calls) == len(G) == inst.M
        for k, (c, g) in enumerate(zip(calls, G)):
            assert c.gamma.shape == g.shape, (inst.M, k)
            assert torch.equal(c.gamma, g), (inst.M, k)
            n_sites += 1
    assert n_sites == sum(inst.M for inst in insts)


def test_gamma_sites_empty_and_invalid():
    from bosonsupremacy_amd import _lib, mps, runtime
    mps.gamma_sites([])  # nothing to do
    ctx = runtime.context()
    with pytest.raises(_lib.BsmpsError):
        ctx.call("bsmps_gamma_sites", 1, None, None)


def test_run_batch_merged_small():
    """Three instances in two groups (a group budget below the tensors of all three): instances/s
    for every form, relative marginal errors of every reduced form within the 1e-6 bar along the
    FP64 sample paths, no draw off those paths."""
    from bosonsupremacy_amd import batched
    insts = small_batch(3, seed=7)
    out = batched.run_batch_merged(insts, n_samples=500, group_bytes=64 << 20)
    print({k: v for k, v in out.items() if k != "per_instance"}, flush=True)
    assert out["instances"] + len(out["skipped"]) == 3 and out["instances"] >= 1
    assert out["groups"] >= 2 and sum(out["instances_per_group"]) == out["instances"]
    for f in ("fp32", "bf16x6", "fp16x3"):
        assert out[f"{f}_marginal_abs_err_max"] < 1e-6
        assert out[f"{f}_first_order_rel_err_max"] < 1e-6 and out[f"{f}_second_order_rel_err_max"] < 1e-6
        assert out[f"{f}_off_fp64_path_total"] == 0
        assert out[f"instances_per_s_{f}"] > 0


def test_run_batch_merged_learned_form():
    """The learned-cutoff plan walked as one more form on the leading blocks of the pruned
    tensors, within the 1e-6 bar along the plain FP64 sample paths."""
    from bosonsupremacy_amd import batched, learned
    insts = small_batch(2, seed=11)
    out = batched.run_batch_merged(insts, n_samples=500, forms=("fp16x3",), learned_model=learned.load_or_train())
    print({k: v for k, v in out.items() if k != "per_instance"}, flush=True)
    assert out["instances"] >= 1
    assert out["instances_per_s_learned"] > 0
    assert out["learned_marginal_abs_err_max"] < 1e-6
    assert out["learned_off_fp64_path_total"] <= out["instances"] * 500 // 100


def test_run_batch_merged_eager_passes_match_graphs():
    """Passes run launch by launch give the same draws and errors as the graph-replayed ones."""
    from bosonsupremacy_amd import batched
    insts = small_batch(2, seed=5)
    a = batched.run_batch_merged(insts, n_samples=400, forms=("fp16x3",))
    b = batched.run_batch_merged(insts, n_samples=400, forms=("fp16x3",), graphs=False)
    assert a["instances"] == b["instances"] >= 1
    for ra, rb in zip(a["per_instance"], b["per_instance"]):
        for key in ("fp16x3_marginal_abs_err", "fp16x3_first_order_rel_err", "fp16x3_second_order_rel_err",
                    "fp16x3_off_fp64_path"):
            assert ra[key] == rb[key], key
