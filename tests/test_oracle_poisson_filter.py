"""The eigen-truncated vector-potential kernel (vector_potential_error > 0,
ComputeVectorPotential.h:296-417) pinned independently of the oracle:

* the oracle's kernel (cyclic Jacobi) equals the one built from LAPACK's
  symmetric eigen-decomposition (numpy.linalg.eigh, the reference uses
  arma::eig_sym) with the reference's truncation rule, filter count included;
* the reference's own cross-check (tests/legacy_test.cu:224-279): a uniform
  current on a 32 x 32 grid, the plain 2 ln r kernel truncated at 1e-6,
  against the legacy direct sum of the Green function, within 1e-8 in the
  relative L-infinity norm;
* the product's host-side kernel (Householder + QL) equals the oracle's,
  through the C ABI (host only, no device needed).

Truncations below ~1e-8 are left out: there 1 - sum(kept)/sum(all) is at the
rounding level of the sum and the stopping point is set by rounding, as the
reference's own warning in initializeFilters says.
"""
import numpy as np
import pytest

from oracle import pyoracle

CASES = [(16, 0.25), (40, 0.1), (101, 0.05)]


def lapack_kernel(n, h, max_error):
    g0, _ = pyoracle.poisson_kernel(n, h, 0.0)
    c = g0.mean()          # G = Q + C with the mean of Q moved into C
    q = g0 - c
    w, v = np.linalg.eigh(q)
    order = np.argsort(-np.abs(w), kind="stable")
    total = np.sum(w * w)
    partial, kept = 0.0, []
    for k in order:
        kept.append(k)
        partial += w[k] ** 2
        if np.sqrt(max(1.0 - partial / total, 0.0)) < max_error:
            break
    return c + (v[:, kept] * w[kept]) @ v[:, kept].T, len(kept) + 1


@pytest.mark.parametrize("n,h", CASES)
@pytest.mark.parametrize("max_error", [1e-3, 1e-6])
def test_oracle_filter_matches_lapack(n, h, max_error):
    g, k = pyoracle.poisson_kernel(n, h, max_error)
    ref, kref = lapack_kernel(n, h, max_error)
    assert k == kref
    assert np.max(np.abs(g - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_truncation_error_is_as_requested():
    n, h = 40, 0.1
    exact, _ = pyoracle.poisson_kernel(n, h, 0.0)
    q = exact - exact.mean()
    for err in (1e-2, 1e-4, 1e-6):
        g, _ = pyoracle.poisson_kernel(n, h, err)
        rel = np.linalg.norm(g - exact) / np.linalg.norm(q)
        assert rel < err
        # and not far below it: the last filter was needed
        assert rel > 1e-3 * err


def test_legacy_direct_sum_cross_check():
    # tests/legacy_test.cu:224-279: N = 32, h = 0.1, j = 1 everywhere, every
    # site inside, no mean removal, analytical = false, error 1e-6.
    N, h = 32, 0.1
    j = np.ones((2, N, N))
    got = pyoracle.vector_potential(N, h, N * N, j, enforce=False, max_error=1e-6,
                                    analytical=False)
    # legacy::computeVectorPotential: G(0) = h^2 (pi - 6 + 2 ln(h^2 / 2)) / 4,
    # G(r) = h^2 ln(h r), A = -G * j / (2 pi)
    y, x = np.mgrid[0:N, 0:N]
    out = np.zeros((N, N))
    for py in range(N):
        for px in range(N):
            r = np.hypot(x - px, y - py)
            with np.errstate(divide="ignore"):
                f = h * h * np.log(h * r)
            f[py, px] = 0.25 * h * h * (np.pi - 6.0 + 2.0 * np.log(h * h * 0.5))
            out += f
    legacy = -out / (2.0 * np.pi)
    for d in range(2):
        res = np.max(np.abs(legacy - got[d])) / np.max(np.abs(legacy))
        assert res < 1e-8, res


def test_exact_kernel_is_the_limit():
    n, h = 24, 0.2
    exact, k0 = pyoracle.poisson_kernel(n, h, 0.0)
    tight, k = pyoracle.poisson_kernel(n, h, 1e-7)
    assert k0 == 0 and k > 1
    assert np.max(np.abs(tight - exact)) < 1e-6 * np.max(np.abs(exact))


@pytest.mark.parametrize("n,h", CASES + [(150, 1.0 / 20.0)])
@pytest.mark.parametrize("max_error", [0.0, 1e-3, 1e-6])
def test_product_kernel_matches_oracle(product, n, h, max_error):
    g, k = product.poisson_kernel(n, h, max_error)
    ref, kref = pyoracle.poisson_kernel(n, h, max_error)
    assert k == kref
    if max_error == 0.0:
        assert np.array_equal(g, ref)
    else:
        assert np.max(np.abs(g - ref)) <= 1e-12 * np.max(np.abs(ref))
