"""The oracle's accelerator steps pinned by what fixes them independently of
the oracle's own code:

* Anderson (accelerators/Anderson.h:267-355): the weights of the bordered
  least-squares system, solved by numpy.linalg.solve (LAPACK) and by the
  closed form of the equality-constrained least squares,
  w = G^-1 1 / (1^T G^-1 1), on random difference sets, with the history
  replacing its oldest pair once full.
* CongAcc (accelerators/CongAcc.h:300-356): the inertia x the bisection
  returns is the midpoint of a bisection bracket that holds the root of
  cos(D, x M + (1 - x) D) = min_cos (brentq on explicit vectors), one of
  whose ends meets the bisection's tolerance 0.01.
* Broyden (no counterpart in the reference): the iterates equal those of
  scipy.optimize.broyden2 (scipy 1.15: restart reduction, no line search)
  on a nonlinear map, and on a linear map with enough memory the method
  reaches the exact fixed point within 2n steps (Gay 1979).
"""
import numpy as np
import pytest
from scipy import optimize

from oracle import pyoracle


def test_anderson_weights_match_lapack_and_closed_form():
    rng = np.random.default_rng(7)
    n0, nf, rank, step = 5, 3, 4, 0.7
    acc = pyoracle.Accelerator("anderson", n0, nf, rank=rank, step=step)
    n = n0 * n0 * nf
    hx, hd, idx = [], [], []
    for it in range(9):
        x = rng.normal(size=n)
        g = x + rng.normal(size=n) * (0.5 + it)
        got = acc.step(x, g)
        d = g - x
        if len(hd) == rank:
            k = int(np.argmin(idx))
            hx[k], hd[k], idx[k] = x, d, it
        else:
            hx.append(x)
            hd.append(d)
            idx.append(it)
        D = np.array(hd)
        gram = D @ D.T
        scale = 1.0 / np.mean(np.diag(gram))
        m = len(hd)
        A = np.ones((m + 1, m + 1))
        A[:m, :m] = gram * scale
        A[m, m] = 0.0
        b = np.zeros(m + 1)
        b[m] = 1.0
        w = np.linalg.solve(A, b)[:m]
        ginv1 = np.linalg.solve(gram, np.ones(m))
        w_closed = ginv1 / ginv1.sum()
        assert np.max(np.abs(w - w_closed)) < 1e-10
        assert abs(w.sum() - 1.0) < 1e-12
        expect = sum(w[i] * (hx[i] + step * hd[i]) for i in range(m))
        err = np.max(np.abs(got - expect)) / np.max(np.abs(expect))
        assert err < 1e-11, (it, err)


@pytest.mark.parametrize("cos_target", [-0.9, -0.2, 0.1, 0.5, 0.69])
def test_congacc_inertia_brackets_min_cosine(cos_target):
    # The bisection evaluates midpoints of [0, 1] and returns the midpoint of
    # its final bracket, computed after the evaluation that met the
    # tolerance: x = (2i + 1) / 2^k with bracket [x - 2^-k, x + 2^-k].  The
    # root x* of cos(D, x M + (1 - x) D) = min_cos (brentq on explicit
    # vectors) lies in that bracket, and one of its ends is within 0.01 of
    # the minimum cosine.
    rng = np.random.default_rng(int(1000 * (cos_target + 1)))
    n, min_cos = 200, 0.7
    d = rng.normal(size=n)
    u = rng.normal(size=n)
    u -= d * (u @ d) / (d @ d)
    u /= np.linalg.norm(u)
    dh = d / np.linalg.norm(d)
    m = 3.0 * (cos_target * dh + np.sqrt(1 - cos_target ** 2) * u)

    def cos_of(x):
        mean = x * m + (1.0 - x) * d
        return (d @ mean) / (np.linalg.norm(d) * np.linalg.norm(mean))

    x = pyoracle.congacc_inertia(np.linalg.norm(d), np.linalg.norm(m), d @ m, min_cos)
    assert 0.0 < x < 1.0
    k = 1
    while (x * 2 ** k) % 1.0 != 0.0:
        k += 1
    half = 2.0 ** -k
    root = optimize.brentq(lambda t: cos_of(t) - min_cos, 0.0, 1.0, xtol=1e-15)
    assert x - half <= root <= x + half, (x, half, root)
    ends = [abs(cos_of(e) - min_cos) for e in (x - half, x + half)]
    assert min(ends) < 0.01, (x, half, ends)


def _scipy_broyden2_iterates(F, x0, alpha, rank, count):
    out = []
    optimize.broyden2(F, x0, iter=count, alpha=alpha, reduction_method="restart",
                      max_rank=rank, line_search=None,
                      callback=lambda x, f: out.append(x.copy()))
    return out


@pytest.mark.parametrize("rank", [3, 16])
def test_broyden_matches_scipy_broyden2(rank):
    rng = np.random.default_rng(3)
    n0, nf = 3, 2
    n = n0 * n0 * nf
    M = rng.normal(size=(n, n)) * 0.25 / np.sqrt(n)
    b = rng.normal(size=n)

    def g(x):
        return np.tanh(M @ x + b)

    def F(x):
        return g(x) - x

    x0 = rng.normal(size=n)
    alpha, count = 0.8, 12
    ref = _scipy_broyden2_iterates(F, x0, alpha, rank, count)
    acc = pyoracle.Accelerator("broyden", n0, nf, rank=rank, step=alpha)
    x = x0.copy()
    for k in range(count):
        x = acc.step(x, g(x))
        err = np.max(np.abs(x - ref[k])) / np.max(np.abs(ref[k]))
        assert err < 1e-12, (k, err)


def test_broyden_linear_fixed_point_within_2n_steps():
    rng = np.random.default_rng(5)
    n0, nf = 2, 2
    n = n0 * n0 * nf
    M = rng.normal(size=(n, n)) * 0.6 / np.sqrt(n)
    b = rng.normal(size=n)
    exact = np.linalg.solve(np.eye(n) - M, b)
    acc = pyoracle.Accelerator("broyden", n0, nf, rank=16, step=1.0)
    x = np.zeros(n)
    for _ in range(2 * n + 1):
        x = acc.step(x, M @ x + b)
    assert np.max(np.abs(x - exact)) < 1e-10 * np.max(np.abs(exact))
