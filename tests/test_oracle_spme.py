"""Pins of the FP64 SPME oracle (oracle/fp64.py: bspline, spme_fp64,
ewald_recip_fp64), the checker behind every SPME test of the engine.

The reference takes SPME from the `ewald` crate (src/non_bonded.rs:5,
:681-875), which is not vendored; the oracle restates smooth PME (Essmann et
al., J. Chem. Phys. 103, 8577 (1995)).  It is pinned here by what the
mathematics fixes, each check chosen so that a plausible slip fails it:

* B-splines: the closed-form cubic M4 and its derivative; partition of unity
  and the convolution identity M5(x) = ∫ M4 over [x-1, x] for order 5
  (a wrong recursion index or derivative sign fails these).
* spme_fp64 against the plain reciprocal Ewald sum ewald_recip_fp64 on a
  converged order-5 grid of a non-cubic box (a transposed axis, a dropped
  B-spline modulus, a factor in the influence function or a wrong virial
  term fails it), and ewald_recip_fp64 itself against the neutral-pair limit
  of the reference's tests/spme.rs (E_rec + E_self + E_direct → -q²/r).
* Forces = -dE/dx by central differences (a sign or derivative slip).
* k-space virial = -dE/dε under uniform scaling of the box and positions.
"""

import math

import numpy as np
import pytest
from scipy.special import erfc

from oracle import fp64


def _m4_closed(x):
    x = np.asarray(x, float)
    out = np.zeros_like(x)
    a = (x >= 0) & (x < 1)
    b = (x >= 1) & (x < 2)
    c = (x >= 2) & (x < 3)
    d = (x >= 3) & (x <= 4)
    out[a] = x[a] ** 3 / 6
    out[b] = (-3 * x[b] ** 3 + 12 * x[b] ** 2 - 12 * x[b] + 4) / 6
    out[c] = (3 * x[c] ** 3 - 24 * x[c] ** 2 + 60 * x[c] - 44) / 6
    out[d] = (4 - x[d]) ** 3 / 6
    return out


def _dm4_closed(x):
    x = np.asarray(x, float)
    out = np.zeros_like(x)
    a = (x >= 0) & (x < 1)
    b = (x >= 1) & (x < 2)
    c = (x >= 2) & (x < 3)
    d = (x >= 3) & (x <= 4)
    out[a] = x[a] ** 2 / 2
    out[b] = (-9 * x[b] ** 2 + 24 * x[b] - 12) / 6
    out[c] = (9 * x[c] ** 2 - 48 * x[c] + 60) / 6
    out[d] = -(4 - x[d]) ** 2 / 2
    return out


def test_bspline_order4_closed_form():
    x = np.linspace(-0.5, 4.5, 2001)
    m, dm = fp64.bspline(x, 4)
    np.testing.assert_allclose(m, _m4_closed(x), atol=1e-14)
    np.testing.assert_allclose(dm, _dm4_closed(x), atol=1e-13)


@pytest.mark.parametrize("n", [4, 5, 6])
def test_bspline_partition_of_unity_and_derivative(n):
    w = np.linspace(0.0, 0.999, 97)
    tot = sum(fp64.bspline(w + k, n)[0] for k in range(n))
    np.testing.assert_allclose(tot, 1.0, atol=1e-14)
    dtot = sum(fp64.bspline(w + k, n)[1] for k in range(n))
    np.testing.assert_allclose(dtot, 0.0, atol=1e-13)
    # derivative against a central difference of M_n itself
    x = np.linspace(0.05, n - 0.05, 301)
    h = 1e-6
    fd = (fp64.bspline(x + h, n)[0] - fp64.bspline(x - h, n)[0]) / (2 * h)
    np.testing.assert_allclose(fp64.bspline(x, n)[1], fd, atol=1e-8)


def test_bspline_order5_is_running_integral_of_order4():
    # M_{n+1}(x) = ∫_{x-1}^{x} M_n(t) dt (Essmann eq. 4.1's recursion in integral form)
    x = np.linspace(0.0, 5.0, 51)
    t = np.linspace(0.0, 1.0, 4001)
    for xv in x:
        tt = xv - 1 + t
        integ = np.trapezoid(_m4_closed(tt), tt)
        assert abs(fp64.bspline(np.array([xv]), 5)[0][0] - integ) < 1e-7


def _charges(n, box, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 1, (n, 3)) * box
    q = rng.normal(0, 1.0, n)
    q -= q.mean()
    return pos, q


BOX = np.array([18.0, 20.0, 22.0])
BETA = 0.3123


@pytest.fixture(scope="module")
def converged():
    pos, q = _charges(40, BOX, 11)
    K = (64, 64, 72)
    f_s, e_s, w_s = fp64.spme_fp64(pos, q, BOX, K, 5, BETA)
    f_e, e_e, w_e = fp64.ewald_recip_fp64(pos, q, BOX, BETA, 14)
    return pos, q, (f_s, e_s, w_s), (f_e, e_e, w_e)


def test_spme_converges_to_plain_ewald(converged):
    _, _, (f_s, e_s, w_s), (f_e, e_e, w_e) = converged
    assert abs(e_s - e_e) <= 1e-6 * abs(e_e)
    assert abs(w_s - w_e) <= 2e-5 * abs(w_e)
    assert np.linalg.norm(f_s - f_e) <= 2e-5 * np.linalg.norm(f_e)


def test_ewald_sum_is_converged_in_m():
    pos, q = _charges(12, BOX, 3)
    f1, e1, w1 = fp64.ewald_recip_fp64(pos, q, BOX, BETA, 14)
    f2, e2, w2 = fp64.ewald_recip_fp64(pos, q, BOX, BETA, 18)
    assert abs(e1 - e2) <= 1e-12 * abs(e2)
    assert np.linalg.norm(f1 - f2) <= 1e-11 * np.linalg.norm(f2)


def test_ewald_neutral_pair_vacuum_limit():
    """tests/spme.rs: a neutral pair far from its images interacts as -q²/r
    once the direct erfc term and the self energy are added."""
    box = np.array([60.0, 60.0, 60.0])
    q = np.array([1.0, -1.0])
    beta = 0.3123
    for r in (2.0, 3.5, 5.0):
        pos = np.array([[30.0, 30.0, 30.0], [30.0 + r, 30.0, 30.0]])
        _, e_rec, _ = fp64.ewald_recip_fp64(pos, q, box, beta, 24)
        e_tot = e_rec + fp64.self_energy(q, beta) + (-1.0) * erfc(beta * r) / r
        # images at 60 Å: dipole-dipole corrections of order (r/60)³/r
        assert abs(e_tot - (-1.0 / r)) <= 0.002 * (1.0 / r)


def test_spme_forces_are_minus_gradient():
    pos, q = _charges(16, BOX, 5)
    K = (24, 27, 30)
    f, _, _ = fp64.spme_fp64(pos, q, BOX, K, 4, BETA)
    h = 1e-5
    for i, a in [(0, 0), (3, 1), (7, 2), (15, 0), (9, 2)]:
        p1 = pos.copy()
        p1[i, a] += h
        p2 = pos.copy()
        p2[i, a] -= h
        e1 = fp64.spme_fp64(p1, q, BOX, K, 4, BETA)[1]
        e2 = fp64.spme_fp64(p2, q, BOX, K, 4, BETA)[1]
        fd = -(e1 - e2) / (2 * h)
        assert abs(fd - f[i, a]) <= 1e-6 * max(1.0, np.abs(f).max()), (i, a, fd, f[i, a])


def test_spme_virial_is_minus_dE_dscale():
    pos, q = _charges(16, BOX, 6)
    K = (24, 27, 30)
    _, e0, w = fp64.spme_fp64(pos, q, BOX, K, 4, BETA)
    eps = 1e-6
    ep = fp64.spme_fp64(pos * (1 + eps), q, BOX * (1 + eps), K, 4, BETA)[1]
    em = fp64.spme_fp64(pos * (1 - eps), q, BOX * (1 - eps), K, 4, BETA)[1]
    fd = -(ep - em) / (2 * eps)
    assert abs(fd - w) <= 1e-6 * max(abs(w), abs(e0))


def test_spme_planted_axis_swap_fails():
    """The Ewald comparison is sensitive to the axis order of a non-cubic box:
    evaluating SPME with the box's axes permuted must miss it by far."""
    pos, q = _charges(40, BOX, 11)
    _, e_e, _ = fp64.ewald_recip_fp64(pos, q, BOX, BETA, 14)
    perm = [1, 2, 0]
    _, e_bad, _ = fp64.spme_fp64(pos, q, BOX[perm], (64, 72, 64), 5, BETA)
    assert abs(e_bad - e_e) > 1e-3 * abs(e_e)
