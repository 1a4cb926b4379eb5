"""Loop-hafnian entry points on the GPU against the oracle's loop recursion: bit for bit at sizes
1-24 (odd sizes padded), in the dense and the targets form; NULL loops take the plain hafnian
path; zero loops give the plain hafnian's value; involution counts as a closed form."""
import numpy as np
import pytest
import torch

from oracle import c_oracle

pytestmark = pytest.mark.gpu


def _random_symmetric(rng, b, n):
    A = rng.normal(size=(b, n, n)) + 1j * rng.normal(size=(b, n, n))
    return A + A.transpose(0, 2, 1)


def _batch(n):
    # 130 = two full waves and a partial one (the oracle takes ~0.03 s per matrix at size 24)
    return 130 if n <= 18 else 65


@pytest.mark.parametrize("n", list(range(1, 25)))
def test_loop_hafnian_dense_matches_oracle(n):
    from bosonsupremacy_amd import mps
    rng = np.random.default_rng(200 + n)
    B = _batch(n)
    A = _random_symmetric(rng, B, n) * 0.5
    v = (rng.normal(size=(B, n)) + 1j * rng.normal(size=(B, n))) * 0.5
    print(f"loop hafnian dense n={n} batch={B}", flush=True)
    got = mps.hafnian(A, loops=v)
    want = c_oracle.loop_hafnian_dense(A, v)
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("n", [2, 5, 8, 13, 20])
def test_loop_hafnian_targets_matches_oracle(n):
    from bosonsupremacy_amd import mps
    rng = np.random.default_rng(300 + n)
    ld = 11
    S = _random_symmetric(rng, 1, ld)[0] * 0.5
    loops = (rng.normal(size=ld) + 1j * rng.normal(size=ld)) * 0.5
    B = _batch(n)
    T = rng.integers(0, ld, size=(B, n)).astype(np.int32)
    den = rng.uniform(0.5, 2.0, size=B)
    print(f"loop hafnian targets n={n} batch={B}", flush=True)
    got, _, _ = mps.A_elem(S, T, den, loops=loops)
    h = c_oracle.loop_hafnian_targets(S, loops, T)
    r = 1.0 / den
    want = h.real * r + 1j * (h.imag * r)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [0, 2, 4, 6, 7, 10, 16, 22])
def test_null_and_zero_loops_give_the_plain_hafnian(n):
    from bosonsupremacy_amd import mps
    rng = np.random.default_rng(400 + n)
    A = _random_symmetric(rng, 70, n)
    plain = mps.hafnian(A)
    # NULL loops: the plain entry point itself
    from bosonsupremacy_amd import runtime
    ctx = runtime.context()
    Ad = runtime.to_device(A)
    out = torch.empty((70,), dtype=torch.complex128, device=ctx.device)
    ctx.call("bsmps_loop_hafnian_dense", runtime.ptr(Ad), None, 70, n, runtime.ptr(out))
    assert np.array_equal(runtime.to_host(out), plain)
    # zero loops: the loop recursion adds exact zeros (size 4 uses a closed formula in the plain path)
    zero = mps.hafnian(A, loops=np.zeros((70, n)))
    if n == 4:
        assert np.allclose(zero, plain, rtol=1e-14, atol=0)
    else:
        assert np.array_equal(zero, plain), n


def test_loop_hafnian_counts_involutions():
    from bosonsupremacy_amd import mps
    T = [1, 1]
    for n in range(2, 25):
        T.append(T[-1] + (n - 1) * T[-2])
    for n in range(1, 25):
        got = mps.hafnian(np.ones((2, n, n)), loops=np.ones((2, n)))
        assert np.all(np.abs(got - T[n]) <= 1e-12 * T[n]), n


def test_loop_hafnian_rejects_too_large():
    from bosonsupremacy_amd import mps
    from bosonsupremacy_amd._lib import BsmpsError
    with pytest.raises(BsmpsError):
        mps.hafnian(np.ones((1, 26, 26)), loops=np.ones((1, 26)))
