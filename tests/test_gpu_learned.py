"""Learned-cutoff form on the GPU: the MLP kernels against the CPU oracle, training, and the
learned form against the plain method on an instance the network never saw."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float32):
    from bosonsupremacy_amd import runtime
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).to(runtime.context().device)


def _run(n, n_in, width, n_out, seed):
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    rng = np.random.default_rng(seed)
    f = np.float32
    W1 = MO.bf16(rng.normal(size=(width, n_in)).astype(f) * 0.5)
    W2 = MO.bf16(rng.normal(size=(width, width)).astype(f) * 0.15)
    W3 = MO.bf16(rng.normal(size=(n_out, width)).astype(f) * 0.15)
    b1, b2, b3 = (rng.normal(size=s).astype(f) * 0.1 for s in (width, width, n_out))
    X = rng.normal(size=(n, n_in)).astype(f)
    dY = rng.normal(size=(n, n_out)).astype(f) / n
    d = {k: _dev(v) for k, v in dict(X=X, b1=b1, b2=b2, b3=b3, dY=dY).items()}
    Wd = [_dev(W, torch.bfloat16) for W in (W1, W2, W3)]  # exact: the values are BF16 already
    H1 = torch.empty((n, width), device=ctx.device)
    H2 = torch.empty((n, width), device=ctx.device)
    Y = torch.empty((n, n_out), device=ctx.device)
    ctx.call("bsmps_mlp_forward", n, n_in, width, n_out, runtime.ptr(d["X"]), runtime.ptr(Wd[0]),
             runtime.ptr(d["b1"]), runtime.ptr(Wd[1]), runtime.ptr(d["b2"]), runtime.ptr(Wd[2]),
             runtime.ptr(d["b3"]), runtime.ptr(H1), runtime.ptr(H2), runtime.ptr(Y))
    grads = [torch.empty(s, device=ctx.device) for s in ((width, n_in), (width,), (width, width), (width,),
                                                         (n_out, width), (n_out,))]
    scratch = torch.empty((2 * n * width,), device=ctx.device)
    ctx.call("bsmps_mlp_backward", n, n_in, width, n_out, runtime.ptr(d["X"]), runtime.ptr(Wd[1]),
             runtime.ptr(Wd[2]), runtime.ptr(H1), runtime.ptr(H2), runtime.ptr(d["dY"]), runtime.ptr(scratch),
             *[runtime.ptr(g) for g in grads])
    want_h1, want_h2, want_y = MO.forward_bf16(X, W1, b1, W2, b2, W3, b3)
    got = [runtime.to_host(t) for t in (H1, H2, Y)]
    want_g = MO.backward_bf16(X, W2, W3, want_h1, want_h2, dY)
    return got, (want_h1, want_h2, want_y), [runtime.to_host(g) for g in grads], want_g


@pytest.mark.parametrize("shape", [(300, 16, 128, 2), (37, 5, 40, 3), (1, 16, 128, 2)])
def test_mlp_forward_backward_match_oracle(shape):
    got, want, g_got, g_want = _run(*shape, seed=sum(shape))
    for a, b in zip(got, want):
        assert np.max(np.abs(a - b)) <= 1e-5 * max(1.0, np.max(np.abs(b))), shape
    for i, (a, b) in enumerate(zip(g_got, g_want)):
        assert np.max(np.abs(a - b)) <= 1e-4 * max(1e-3, np.max(np.abs(b))), (shape, i)


def test_adam_and_mse_match_oracle():
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    rng = np.random.default_rng(3)
    p, g, m, v = (rng.normal(size=1000).astype(np.float32) for _ in range(4))
    v = np.abs(v)
    pd, gd, md, vd = (_dev(a) for a in (p, g, m, v))
    ctx.call("bsmps_adam_step", 1000, runtime.ptr(pd), runtime.ptr(gd), runtime.ptr(md), runtime.ptr(vd),
             1e-3, 0.9, 0.999, 1e-8, 7)
    wp, wm, wv = MO.adam_step(p, g, m, v, 1e-3, 0.9, 0.999, 1e-8, 7)
    # FP32 arithmetic on both sides; pow and division may differ by an ulp
    assert np.allclose(runtime.to_host(pd), wp, rtol=1e-6, atol=1e-7)
    assert np.allclose(runtime.to_host(md), wm, rtol=1e-6, atol=1e-7)
    assert np.allclose(runtime.to_host(vd), wv, rtol=1e-6, atol=1e-7)
    Y, T = rng.normal(size=(50, 2)).astype(np.float32), rng.normal(size=(50, 2)).astype(np.float32)
    dY, err = torch.empty((50, 2), device=ctx.device), torch.empty((50, 2), device=ctx.device)
    Yd, Td = _dev(Y), _dev(T)  # keep both alive while the kernel runs
    ctx.call("bsmps_mse_grad", 50, 2, runtime.ptr(Yd), runtime.ptr(Td), runtime.ptr(dY), runtime.ptr(err))
    assert np.allclose(runtime.to_host(dY), 2 * (Y - T) / 100, rtol=1e-6)
    assert np.allclose(runtime.to_host(err), (Y - T) ** 2, rtol=1e-6)


def test_training_fits_a_smooth_function():
    from bosonsupremacy_amd import learned
    rng = np.random.default_rng(4)
    X = rng.normal(size=(800, 16)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 1] ** 2, np.tanh(X[:, 2] * X[:, 3])], axis=1).astype(np.float32) / 3
    model = learned.MLP(seed=1)
    hist = model.fit(X, T, steps=1500, log_every=100)
    print("loss", hist[0], hist[-1], flush=True)
    assert hist[-1][1] < 0.05 * hist[0][1]
    pred = model.predict(X[:5])
    assert pred.shape == (5, 2) and np.all(np.isfinite(pred))


def test_learned_form_gain_on_unseen_instance():
    """Train on plain-method runs of other seeds, plan the M=32 workload (seed 0), and compare with
    the pruned plain method -- the plain method without the work a zero Lambda multiplies away,
    which already keeps exactly the labels of Schmidt weight above 1e-10.

    * At the 1e-6 bar (the default tolerances) every kept prediction passes the exact check and
      the first-order marginals move by less than 1e-6; the network's own saving is reported (the
      bar leaves it almost nothing to drop: a truncation moves the sampled conditionals by the
      square root of the weight it drops).
    * At the looser tolerance the network is trained for (1e-7 of Schmidt weight per cut) its
      saving and the marginal change it costs are reported."""
    from bosonsupremacy_amd import instances, learned, mps, runtime, sampling
    from bench import site_hafnian_work_by_site
    model = learned.MLP(seed=0)
    X, T = learned.training_pairs(learned.training_instances(seeds=range(1, 5)))
    hist = model.fit(X, T, steps=2000, log_every=500)
    print("training loss", hist, flush=True)
    inst = instances.workload("m32_d6_chi512")
    Gp, Lp, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi, pruned=True)
    kres, knum = kron_out[0], kron_out[1]
    kept = [learned.kept_rank(r) for r in kres]
    w_pruned = sum(w[1] for w in site_hafnian_work_by_site(kres, knum, inst.d, site_chi=kept))
    c_pruned = sum(kept[i - 1] * kept[i] * inst.d for i in range(1, inst.M - 1)) + kept[-1] * inst.d
    sqrtW = sampling.thermal_sqrtW(inst.cov, inst.sq_cov)
    normals, uniforms = sampling.random_inputs(inst.M, 20000, seed=9)
    _, pa = sampling.Sampler(Gp, Lp, sqrtW, inst.dd).run(normals, uniforms, return_probs=True)
    ma = (runtime.to_host(pa) * np.arange(inst.dd)).sum(axis=2).mean(axis=1)
    del pa
    results = {}
    for tau in (learned.TAU_CHI, learned.TAU_CHI_TRAIN):
        site_d, site_chi, report = learned.plan_sites(model, inst.sq_cov, inst.cov, kres, inst.d, inst.chi,
                                                      tau_chi=tau)
        print(f"tau {tau:g}: plan {site_chi} {report}", flush=True)
        assert report["max_extra_photon_weight"] <= learned.TAU_D
        assert report["max_extra_schmidt_weight"] <= tau
        assert all(1 <= r <= k for r, k in zip(site_chi, kept)) and all(1 <= x <= inst.d for x in site_d)
        G2, L2, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=kron_out, site_d=site_d, site_chi=site_chi)
        # the learned tensors are leading blocks of the pruned tensors (same entries)
        for k in (1, inst.M // 2, inst.M - 2):
            g2 = runtime.to_host(G2[k])
            assert np.array_equal(g2, runtime.to_host(Gp[k])[:g2.shape[0], :g2.shape[1], :g2.shape[2]])
        w_learned = sum(w[1] for w in site_hafnian_work_by_site(kres, knum, inst.d, site_d=site_d, site_chi=site_chi))
        c_learned = sum(site_chi[i - 1] * site_chi[i] * site_d[i] for i in range(1, inst.M - 1)) + \
            site_chi[-1] * site_d[-1]
        assert w_learned <= w_pruned and c_learned <= c_pruned
        _, pb = sampling.Sampler(G2, L2, sqrtW, inst.dd).run(normals, uniforms, return_probs=True)
        mb = (runtime.to_host(pb) * np.arange(inst.dd)).sum(axis=2).mean(axis=1)
        rel = float(np.max(np.abs(ma - mb) / ma))
        results[tau] = (1 - w_learned / w_pruned, 1 - c_learned / c_pruned, rel)
        print(f"tau {tau:g}: network's saving over the pruned plain form: hafnian work {results[tau][0]:.3%}, "
              f"contraction work {results[tau][1]:.3%}, ranks below kept at "
              f"{sum(r < k for r, k in zip(site_chi, kept))} of {len(kept)} cuts; first-order marginal "
              f"change {rel:.2e}", flush=True)
        del G2, L2, pb
    assert results[learned.TAU_CHI][2] < 1e-6            # the 1e-6 bar holds at the default tolerance
    assert results[learned.TAU_CHI_TRAIN][2] < 1e-4      # the looser tolerance costs ~1e-5
