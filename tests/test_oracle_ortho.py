"""CPU checks that pin the oracle's OrthonormalizationMultigrid restatement
(oracle/ortho.py) by what the mathematics fixes, independently of its code:

* after every correction the iterate minimises |b - A x| over x0 + the span of all
  raw corrections so far (a dense least-squares solve in numpy): a wrong
  coefficient, a dropped orthogonalisation term or a transposed product fails it;
* the stored A-images are orthonormal;
* a correction that repeats an earlier one triggers the severe-cancellation path,
  and the iterate is still the least-squares minimiser;
* a zero correction takes the random direction, which is the documented hash;
* the termination is the reference's DefaultTermination, and a post-smoother
  cancellation skips the next pre-smoother.
"""
import numpy as np
import pytest

from oracle import ortho as OT


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return Q @ np.diag(np.logspace(0, 3, n)) @ Q.T


def _components(A, seed):
    """Jacobi smoother and a Galerkin coarse correction V (V^T A V)^-1 V^T r on a
    random third of the space."""
    rng = np.random.default_rng(seed)
    n = A.shape[0]
    D = np.diag(1.0 / np.diag(A))
    V = rng.standard_normal((n, n // 3))
    Ac = V.T @ A @ V
    return (lambda r: 0.7 * D @ r), (lambda r: V @ np.linalg.solve(Ac, V.T @ r)), (lambda r: 0.7 * D @ r)


def test_least_squares_optimal_after_every_correction():
    n = 40
    A = _spd(n, 1)
    pre, coarse, post = _components(A, 2)
    b = np.random.default_rng(3).standard_normal(n)
    om = OT.OrthoMG(lambda x: A @ x, pre=pre, coarse=coarse, post=post, post_sweeps=2, tol=1e-30, maxit=4, trace=True)
    om.solve(b)
    raw = np.array([c for _, c in om.raw])
    assert len(raw) == len(om.resnorms) == 4 * 4
    for k in range(1, len(raw) + 1):
        C = raw[:k].T
        y, *_ = np.linalg.lstsq(A @ C, b, rcond=None)
        best = np.linalg.norm(b - A @ C @ y)
        assert abs(om.resnorms[k - 1] - best) <= 1e-9 * np.linalg.norm(b) + 1e-9 * best
    Qh = np.array(om.Q)
    assert np.allclose(Qh @ Qh.T, np.eye(len(Qh)), atol=1e-10)


def test_converges_to_direct_solution():
    n = 30
    A = _spd(n, 4)
    D = np.diag(1.0 / np.diag(A))
    b = np.random.default_rng(5).standard_normal(n)
    om = OT.OrthoMG(lambda x: A @ x, pre=lambda r: 0.8 * D @ r, coarse=None, post=lambda r: 0.8 * D @ r,
                    tol=1e-12, maxit=200)
    x = om.solve(b)
    assert om.converged
    assert np.linalg.norm(x - np.linalg.solve(A, b)) <= 1e-9 * np.linalg.norm(np.linalg.solve(A, b))
    assert all(om.resnorms[i + 1] <= om.resnorms[i] * (1 + 1e-12) for i in range(len(om.resnorms) - 1))


def test_repeated_correction_takes_the_cancellation_path():
    n = 20
    A = _spd(n, 6)
    b = np.random.default_rng(7).standard_normal(n)
    fixed = np.random.default_rng(8).standard_normal(n)
    # the same direction twice: the second is cancelled by the first
    om = OT.OrthoMG(lambda x: A @ x, pre=lambda r: fixed, coarse=lambda r: fixed.copy(), post=None, tol=1e-30, maxit=1,
                    trace=True)
    om.solve(b)
    assert om.n_cancel >= 1
    raw = np.array([c for _, c in om.raw])
    y, *_ = np.linalg.lstsq(A @ raw[:1].T, b, rcond=None)
    assert om.resnorms[1] <= np.linalg.norm(b - A @ raw[:1].T @ y) * (1 + 1e-9)


def test_zero_correction_takes_the_random_direction():
    n = 16
    A = _spd(n, 9)
    b = np.random.default_rng(10).standard_normal(n)
    om = OT.OrthoMG(lambda x: A @ x, pre=lambda r: np.zeros(n), coarse=None, post=None, tol=1e-30, maxit=1)
    om.solve(b)
    assert om.n_random == 1
    d = OT.random_direction(n, 1)
    assert np.all((d >= 0) & (d < 1)) and len(np.unique(d)) == n
    q = A @ d
    q /= np.linalg.norm(q)
    assert np.allclose(om.Q[0], q, atol=1e-14)
    # first value of the hash sequence, computed by hand from splitmix64's finaliser
    M = (1 << 64) - 1
    z = 0x5DEECE66D + 7919
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    z ^= z >> 31
    assert d[0] == (z >> 11) / 2.0 ** 53


def test_termination_and_pre_smoother_skip():
    assert OT.keep_running(1, 1.0, 0.5, 1e-3, 10) == (True, False)
    assert OT.keep_running(3, 1.0, 1e-4, 1e-3, 10) == (False, True)      # r < r0 tol + tol
    assert OT.keep_running(11, 1.0, 0.5, 1e-3, 10) == (False, False)
    n = 12
    A = _spd(n, 11)
    b = np.random.default_rng(12).standard_normal(n)
    fixed = np.random.default_rng(13).standard_normal(n)
    calls = {"pre": 0}

    def pre(r):
        calls["pre"] += 1
        return np.diag(1.0 / np.diag(A)) @ r

    # the post-smoother repeats the coarse direction: cancelled, so iteration 2 skips pre
    om = OT.OrthoMG(lambda x: A @ x, pre=pre, coarse=lambda r: fixed, post=lambda r: fixed, tol=1e-30, maxit=3)
    om.solve(b)
    assert om.iterations == 3
    assert calls["pre"] == 1
