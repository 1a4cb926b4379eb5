"""Whole-tensor checks of the GPU site tensors against the mathematics (no oracle involved):
each Gamma_k Lambda_k is a truncated block of a unitary matrix (tests/closed_forms.py), and the
pruned site tensors equal the plain tensors' live blocks."""
import numpy as np
import pytest

import closed_forms as CF
from conftest import same_bits

pytestmark = pytest.mark.gpu


def _canonical_report(G, L, kres, sq_cov, d):
    """Per site k >= 1: (max_b s_b - 1, weighted right deficit, its truncation bound).

    The deficit sum_b lam_b^2 (1 - s_b) is the weight of the cut-(k-1) state outside the kept
    (photon number j < d, cut-k label) pairs.  With P the projector onto those pairs and Q the
    projector onto the kept labels of cut k-1, ||(1-P) Q psi|| <= ||(1-P) psi|| + ||(1-Q) psi||, so
    deficit <= (sqrt(tail_k + disc_k) + sqrt(disc_{k-1}))^2, with tail_k = P(n_k >= d) of the pure
    state's mode k and disc = 1 - (retained Schmidt weight) of each cut."""
    from bosonsupremacy_amd import gaussian, runtime
    M = len(G)
    disc = [1.0 - float(np.sum(r[r > 1e-10])) for r in kres]
    out = []
    for k in range(1, M):
        g = runtime.to_host(G[k])
        lam_prev = np.asarray(L[k - 1])
        lam_next = np.asarray(L[k]) if k < M - 1 else np.eye(1, g.shape[1])[0]
        live = lam_prev > 0
        s = CF.right_canonical_sums(g, lam_next)[:len(lam_prev)][live]
        deficit = float((lam_prev[live] ** 2 * (1 - s)).sum())
        Vk = sq_cov[np.ix_([k, k + M], [k, k + M])]
        tail = float(gaussian.photon_tail(Vk, d))
        dk = disc[k] if k < M - 1 else 0.0
        bound = (np.sqrt(max(tail + dk, 0.0)) + np.sqrt(max(disc[k - 1], 0.0))) ** 2
        out.append((k, float(s.max() - 1), deficit, bound))
    return out


@pytest.mark.parametrize("name", ["m8_d4_chi64", "m32_d6_chi512"])
def test_gpu_site_tensors_are_canonical(name):
    from bosonsupremacy_amd import instances, mps
    inst = instances.workload(name)
    G, L, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi)
    for k, over, deficit, bound in _canonical_report(G, L, kron_out[0], inst.sq_cov, inst.d):
        print(f"{name} site {k}: max s_b - 1 = {over:.2e}, deficit {deficit:.3e} <= bound {bound:.3e}", flush=True)
        assert over <= 1e-12, k             # a block of a unitary: no row norm above 1
        assert -1e-12 <= deficit <= bound * (1 + 1e-9) + 1e-13, k


def test_pruned_site_tensors_are_the_live_blocks():
    """build_mps(pruned=True) computes only the kept labels of both cuts; its entries are the
    plain tensors' entries, bit for bit."""
    from bosonsupremacy_amd import instances, mps, runtime
    inst = instances.workload("m32_d6_chi512")
    G, L, kron_out = mps.build_mps(inst.sq_cov, inst.d, inst.chi)
    Gp, Lp, _ = mps.build_mps(inst.sq_cov, inst.d, inst.chi, kron=kron_out, pruned=True)
    for k in range(inst.M):
        gp = runtime.to_host(Gp[k])
        rows, cols = gp.shape[:2]
        if k < inst.M - 1:
            assert cols == int((kron_out[0][k] > 1e-10).sum()) and same_bits(Lp[k], L[k][:cols]), k
        assert same_bits(gp, runtime.to_host(G[k])[:rows, :cols]), k
