"""GPU hafnians vs the reference's bits (golden vectors) and the C oracle."""
import numpy as np
import pytest
import torch

from conftest import same_bits
from oracle import c_oracle

pytestmark = pytest.mark.gpu

TUNINGS = [(4, 64), (5, 64), (4, 32), (4, 16), (2, 16)]


@pytest.fixture(scope="module")
def ctx():
    from bosonsupremacy_amd import runtime
    c = runtime.context()
    yield c
    c.set_tuning(4, 64)
    for n in range(12, 31, 2):
        c.call("bsmps_set_split", n, 1)
    c.call("bsmps_set_split_budget", 16 << 30)


@pytest.mark.parametrize("tuning", TUNINGS)
def test_dense_matches_reference_bits(ctx, golden_haf, tuning):
    from bosonsupremacy_amd import mps
    ctx.set_tuning(*tuning)
    for n in range(0, 17):
        A = golden_haf[f"A_{n}"]
        print(f"dense n={n} tuning={tuning}", flush=True)
        assert same_bits(mps.hafnian(A), golden_haf[f"haf_{n}"]), n


@pytest.mark.parametrize("tuning", TUNINGS)
def test_targets_match_reference_bits(ctx, golden_m8, tuning):
    from bosonsupremacy_amd import mps
    ctx.set_tuning(*tuning)
    S = golden_m8["site_sigma128"]
    for n in (2, 4, 6, 8, 10, 12, 14):
        tg = golden_m8[f"site_targets_{n}"]
        vals, _, _ = mps.A_elem(S, tg, np.ones(len(tg)))
        assert same_bits(vals, golden_m8[f"site_haf_{n}"]), n


@pytest.mark.parametrize("n", [16, 18, 20, 22, 24, 26, 28, 30, 32, 34])
def test_large_sizes_match_c_oracle(ctx, n):
    from bosonsupremacy_amd import mps
    ctx.set_tuning(4, 64)
    rng = np.random.default_rng(n)
    ld = 30
    S = (rng.normal(size=(ld, ld)) + 1j * rng.normal(size=(ld, ld))) * 0.4
    tg = rng.integers(0, ld, size=(5, n)).astype(np.int32)
    den = rng.uniform(0.5, 2.0, size=5)
    print(f"targets n={n}", flush=True)
    vals, _, _ = mps.A_elem(S, tg, den)
    want = c_oracle.hafnian_targets(S, tg) / den
    assert same_bits(vals, want)


def test_many_items_all_paths_agree(ctx):
    # a batch much larger than one launch tile, through every tuning, must give identical bits
    from bosonsupremacy_amd import mps
    rng = np.random.default_rng(1)
    ld = 24
    S = (rng.normal(size=(ld, ld)) + 1j * rng.normal(size=(ld, ld))) * 0.5
    for n in (6, 8, 10, 12):
        tg = rng.integers(0, ld, size=(3000, n)).astype(np.int32)
        want = c_oracle.hafnian_targets(S, tg[:200])
        outs = []
        for t in TUNINGS:
            ctx.set_tuning(*t)
            vals, _, _ = mps.A_elem(S, tg, np.ones(len(tg)))
            outs.append(vals)
            assert same_bits(vals[:200], want), (n, t)
        for o in outs[1:]:
            assert same_bits(o, outs[0])


@pytest.mark.parametrize("n", [12, 14, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34])
def test_split_matches_unsplit_and_oracle(ctx, n):
    """The frontier-split path (upper tree on cooperating lanes, subtrees one per lane) gives
    the same bits as the unsplit cooperative path and the C oracle, for every lane count and
    when the batch is cut into many chunks."""
    from bosonsupremacy_amd import _lib, mps
    ctx.set_tuning(4, 64)
    rng = np.random.default_rng(100 + n)
    ld = 30
    S = (rng.normal(size=(ld, ld)) + 1j * rng.normal(size=(ld, ld))) * 0.4
    count = {12: 3000, 14: 1500, 16: 700, 18: 300, 20: 130, 22: 70, 24: 40, 26: 20, 28: 10, 30: 6, 32: 4, 34: 3}[n]
    tg = rng.integers(0, ld, size=(count, n)).astype(np.int32)
    den = rng.uniform(0.5, 2.0, size=count)
    k = min(12, count)
    want = c_oracle.hafnian_targets(S, tg[:k]) / den[:k]
    outs = {}
    for split in (0, 1):
        ctx.call("bsmps_set_split", n, split)
        for g in (16, 32, 64):
            ctx.call("bsmps_set_group_lanes", n, g)
            print(f"n={n} split={split} lanes={g}", flush=True)
            try:
                vals, _, _ = mps.A_elem(S, tg, den)
            except _lib.BsmpsError as e:  # LDS of 64/g matrices per workgroup too large
                assert e.status == _lib.BSMPS_ERR_UNSUPPORTED
                continue
            outs[(split, g)] = vals
            assert same_bits(vals[:k], want), (split, g)
    assert (1, 64) in outs and (n > 24 or (0, 64) in outs)  # the unsplit path stops at 24
    ctx.call("bsmps_set_split", n, 1)
    ctx.call("bsmps_set_group_lanes", n, 64)
    ctx.call("bsmps_set_split_budget", 1 << 20)
    print(f"n={n} split, 1 MiB budget", flush=True)
    chunked, _, _ = mps.A_elem(S, tg, den)
    ctx.call("bsmps_set_split_budget", 16 << 30)
    ref = outs.get((0, 64), outs[(1, 64)])
    for key, vals in outs.items():
        assert same_bits(vals, ref), key
    assert same_bits(chunked, ref)


def test_sigma_select_and_odd_sizes(ctx):
    from bosonsupremacy_amd import mps
    rng = np.random.default_rng(2)
    S = rng.normal(size=(10, 10)) + 1j * rng.normal(size=(10, 10))
    tg = rng.integers(0, 10, size=(7, 5)).astype(np.int32)
    got = mps.Sigma_select(S, tg)
    assert same_bits(got, S[tg[:, :, None], tg[:, None, :]])
    vals, _, _ = mps.A_elem(S, tg, np.ones(7))
    assert same_bits(vals, np.zeros(7, complex))
    assert same_bits(mps.hafnian(np.zeros((3, 0, 0))), np.ones(3, complex))


def test_invalid_sizes_raise(ctx):
    from bosonsupremacy_amd import _lib, mps
    with pytest.raises(_lib.BsmpsError):
        mps.A_elem(np.eye(4), np.zeros((2, 36), np.int32), np.ones(2))
    with pytest.raises(_lib.BsmpsError):
        ctx.set_tuning(6, 64)
    with pytest.raises(_lib.BsmpsError):
        ctx.call("bsmps_set_split", 36, 1)
    with pytest.raises(_lib.BsmpsError):
        ctx.call("bsmps_set_split_budget", 1000)
