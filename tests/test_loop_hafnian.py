"""The oracle's loop-hafnian recursion (the reference's plain recursion, MPS_cpu.py:143-187,
extended by one loop polynomial per row).  The reference has no loop hafnian, so the restatement
is pinned by what the mathematics fixes: brute-force sums over matchings with loops, involution
counts, rank-one matrices, and the plain recursion bit for bit when every loop weight is zero."""
import numpy as np
import pytest

from oracle import c_oracle


def _brute_loop_hafnian(A, v):
    """Sum over all matchings with loops: each vertex is paired with another (weight A[i, j])
    or matched to itself (weight v[i])."""
    def rec(rest):
        if not rest:
            return 1.0 + 0j
        i, others = rest[0], rest[1:]
        total = v[i] * rec(others)
        for k, j in enumerate(others):
            total += A[i, j] * rec(others[:k] + others[k + 1:])
        return total
    return rec(list(range(len(A))))


def _random_symmetric(rng, n):
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    return A + A.T


INVOLUTIONS = [1, 1, 2, 4, 10, 26, 76, 232, 764, 2620, 9496, 35696, 140152, 568504]


@pytest.mark.parametrize("n", range(0, 10))
def test_loop_hafnian_matches_matchings_with_loops(n):
    rng = np.random.default_rng(100 + n)
    A = np.stack([_random_symmetric(rng, n) for _ in range(3)])
    v = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    got = c_oracle.loop_hafnian_dense(A, v)
    want = np.array([_brute_loop_hafnian(A[b], v[b]) for b in range(3)])
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_loop_hafnian_counts_involutions():
    """All-ones matrix with unit loops: the number of matchings with loops of n vertices is the
    number of involutions of n elements."""
    for n, count in enumerate(INVOLUTIONS):
        assert c_oracle.loop_hafnian_dense(np.ones((1, n, n)), np.ones((1, n)))[0] == count


def test_loop_hafnian_of_rank_one():
    """A = a a^T with loop weights a: every matching with loops weighs prod(a)."""
    rng = np.random.default_rng(7)
    for n in range(1, 12):
        a = rng.normal(size=n) + 1j * rng.normal(size=n)
        got = c_oracle.loop_hafnian_dense(np.outer(a, a)[None], a[None])[0]
        want = np.prod(a) * INVOLUTIONS[n]
        assert abs(got - want) <= 1e-12 * abs(want)


def test_zero_loops_give_the_plain_recursion():
    """With zero loop weights every added term is an exact zero: the value is the plain
    hafnian's, bit for bit where the reference uses the recursion (sizes 2 and >= 6; size 4 has
    its own closed formula there), and 0 for odd sizes."""
    rng = np.random.default_rng(3)
    for n in range(0, 15):
        A = np.stack([_random_symmetric(rng, n) for _ in range(4)])
        got = c_oracle.loop_hafnian_dense(A, np.zeros((4, n)))
        want = c_oracle.hafnian_dense(A)
        if n == 4:
            assert np.allclose(got, want, rtol=1e-14, atol=0)
        else:
            assert np.array_equal(got, want), n


def test_targets_form_gathers_loops_like_sigma():
    """Row r of a target matrix has loop weight loops[t_r], repeated indices included."""
    rng = np.random.default_rng(5)
    S = _random_symmetric(rng, 9)
    loops = rng.normal(size=9) + 1j * rng.normal(size=9)
    T = rng.integers(0, 9, size=(6, 7)).astype(np.int32)
    got = c_oracle.loop_hafnian_targets(S, loops, T)
    for b in range(6):
        t = T[b]
        want = c_oracle.loop_hafnian_dense(S[np.ix_(t, t)][None], loops[t][None])[0]
        assert got[b] == want
