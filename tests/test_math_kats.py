"""Known-answer tests of the pair arithmetic the kernels use (dyn_math.h,
compiled for the host in the oracle's replay library).  Each test mirrors one
of the reference's closed-form tests in src/tests/forces.rs, plus accuracy
bounds of the IEEE-only rsqrt/exp/erfc against scipy."""

import ctypes as C
import math

import numpy as np
import pytest
from scipy.special import erfc, erfcx

from conftest import oracle_lib

CHARGE_UNIT_SCALER = np.float32(18.2223)
EWALD_ALPHA = 0.35
COULOMB_CUTOFF = 10.0


def lj(r, sigma, eps):
    """LJ via the engine's c6/c12 form: returns (Fx along +x for tgt at +x, E)."""
    L = oracle_lib()
    c6 = 4 * eps * sigma ** 6
    c12 = 4 * eps * sigma ** 12
    s, e = C.c_float(), C.c_float()
    L.orc_lj(np.float32(r * r), np.float32(c6), np.float32(c12), C.byref(s), C.byref(e))
    return s.value * r, e.value


# the two erfc Coulomb forms: closed form (exp · erfcx polynomial; the
# corrections and the reference-list kernel) and the table of the pair tiles
COUL_FORMS = ["orc_coul_erfc", "orc_coul_erfc_tab"]


def coul(r, q1, q2, alpha=EWALD_ALPHA, cutoff=COULOMB_CUTOFF, form="orc_coul_erfc"):
    L = oracle_lib()
    if r >= cutoff:
        return 0.0, 0.0
    s, e = C.c_float(), C.c_float()
    getattr(L, form)(np.float32(r * r), np.float32(q1 * q2), np.float32(alpha), C.byref(s),
                     C.byref(e))
    return s.value * r, e.value


def soft(r, sigma, eps, lam):
    L = oracle_lib()
    s, e, d = C.c_float(), C.c_float(), C.c_float()
    L.orc_soft_core(np.float32(r * r), np.float32(sigma), np.float32(eps), np.float32(lam),
                    C.byref(s), C.byref(e), C.byref(d))
    return s.value * r, e.value, d.value


def test_rsqrt_accuracy():
    L = oracle_lib()
    x = np.geomspace(1e-10, 1e10, 20001).astype(np.float32)
    y = np.array([L.orc_rsqrtf(float(v)) for v in x])
    rel = np.abs(y * np.sqrt(x.astype(np.float64)) - 1)
    assert rel.max() < 2e-7


def test_exp_and_erfc_accuracy():
    L = oracle_lib()
    ys = np.linspace(-80, 0, 20001).astype(np.float32)
    e = np.array([L.orc_exp_neg(float(v)) for v in ys])
    assert np.max(np.abs(e / np.exp(ys.astype(np.float64)) - 1)) < 2e-7
    xs = np.linspace(0, 4, 20001).astype(np.float32)
    g = np.array([L.orc_erfcx(float(v)) for v in xs])
    assert np.max(np.abs(g / erfcx(xs.astype(np.float64)) - 1)) < 3e-7
    ec = np.array([L.orc_erfcf(float(v)) for v in xs])
    # f32 input rounding alone costs ~2x·|x|·6e-8 relative at the top of the range
    assert np.max(np.abs(ec / erfc(xs.astype(np.float64)) - 1)) < 1e-6


def test_fixed_point_conversion():
    L = oracle_lib()
    for v in [0.0, 1.0, -1.0, 3.5, -1234.5678, 1e-9, 2.0 ** -33, 262144.0]:
        want = int(np.rint(np.float64(np.float32(v)) * 2.0 ** 32))
        assert L.orc_fx(v) == want
    # clamped, not wrapped
    assert L.orc_fx(1e9) == int(262144.0 * 2 ** 32)
    assert L.orc_fx(-1e9) == -int(262144.0 * 2 ** 32)
    assert L.orc_fx(float("inf")) == int(262144.0 * 2 ** 32)


def test_lennard_jones():
    """tests/forces.rs:222-274: force and energy against the analytic 12-6
    formulas, repulsive below r_min, attractive above, E(r_min) = −ε."""
    sigma, eps = 3.4, 0.086
    r_min = 2 ** (1 / 6) * sigma
    for r in [2.0, 2.5, 3.0, 4.5, 5.0, 7.0]:
        fx, e = lj(r, sigma, eps)
        sr6 = (sigma / r) ** 6
        f_an = 24 * eps * (2 * sr6 * sr6 - sr6) / r
        e_an = 4 * eps * (sr6 * sr6 - sr6)
        assert abs(fx - f_an) <= 1e-4 * max(1.0, abs(f_an))
        assert abs(e - e_an) <= 1e-4 * max(1.0, abs(e_an))
        assert (fx > 0) == (r < r_min)
    _, e_min = lj(r_min, sigma, eps)
    assert abs(e_min + eps) < 1e-4


def test_lennard_jones_newton3():
    """tests/forces.rs:278-294: the pair force depends on r only, F_src = −F_tgt."""
    for r in [2.0, 3.817, 5.5]:
        f1, _ = lj(r, 3.4, 0.086)
        f2, _ = lj(r, 3.4, 0.086)
        assert f1 - f2 == 0.0


def test_alchemical_lj_soft_core_endpoints():
    """tests/forces.rs:296-322."""
    sigma, eps, r = 3.4, 0.086, 2.4
    sr6 = (sigma / r) ** 6
    hard_e = 4 * eps * (sr6 * sr6 - sr6)
    hard_f = 24 * eps * (2 * sr6 * sr6 - sr6) / r
    f0, e0, _ = soft(r, sigma, eps, 0.0)
    f1, e1, d1 = soft(r, sigma, eps, 1.0)
    assert abs(e0 - hard_e) < 1e-5 * max(1, abs(hard_e))
    assert abs(f0 - hard_f) < 1e-5 * max(1, abs(hard_f))
    assert abs(e1) < 1e-6 and abs(f1) < 1e-6
    assert math.isfinite(d1)


def test_alchemical_lj_soft_core_dhdl_matches_finite_difference():
    """tests/forces.rs:324-346 (2 %)."""
    sigma, eps, r, lam, h = 3.4, 0.086, 2.2, 0.37, 1e-3
    _, _, d = soft(r, sigma, eps, lam)
    _, ep, _ = soft(r, sigma, eps, lam + h)
    _, em, _ = soft(r, sigma, eps, lam - h)
    num = (ep - em) / (2 * h)
    assert abs(d - num) <= 0.02 * abs(num)


def test_alchemical_lj_soft_core_force_matches_energy_gradient():
    """The FP64 soft-core oracle satisfies F = −dE/dr and dH/dλ = dE/dλ (f64
    central differences); the f32 arithmetic matches that oracle."""
    from oracle import fp64
    sigma, eps, lam, h = 3.4, 0.086, 0.3, 1e-6
    r = np.array([1.0, 2.2, 3.5, 5.0])
    e, f, d = fp64.soft_core_fp64(r, sigma, eps, lam)
    ep = fp64.soft_core_fp64(r + h, sigma, eps, lam)[0]
    em = fp64.soft_core_fp64(r - h, sigma, eps, lam)[0]
    assert np.allclose(f, -(ep - em) / (2 * h), rtol=1e-6, atol=1e-9)
    lp = fp64.soft_core_fp64(r, sigma, eps, lam + h)[0]
    lm = fp64.soft_core_fp64(r, sigma, eps, lam - h)[0]
    assert np.allclose(d, (lp - lm) / (2 * h), rtol=1e-6, atol=1e-9)
    for k, rr in enumerate(r):
        f32, e32, d32 = soft(rr, sigma, eps, lam)
        assert abs(f32 - f[k]) <= 1e-5 * max(abs(f[k]), 1e-2)
        assert abs(e32 - e[k]) <= 1e-5 * max(abs(e[k]), 1e-2)
        assert abs(d32 - d[k]) <= 1e-5 * max(abs(d[k]), 1e-2)


def test_alchemical_lj_soft_core_overlap_is_finite():
    """tests/forces.rs:348-367."""
    f, e, d = soft(1e-4, 3.4, 0.086, 0.5)
    assert all(math.isfinite(v) for v in (f, e, d))


@pytest.mark.parametrize("form", COUL_FORMS)
@pytest.mark.parametrize("q1e,q2e", [(1.0, -1.0), (1.0, 1.0), (0.5, -0.3)])
def test_coulomb_short_range_force_energy_consistency(q1e, q2e, form):
    """tests/forces.rs:378-423: F = −dE/dr by central difference (2 %)."""
    q1, q2 = q1e * CHARGE_UNIT_SCALER, q2e * CHARGE_UNIT_SCALER
    for r in [3.0, 5.0, 8.0]:
        f, _ = coul(r, q1, q2, form=form)
        _, ep = coul(r + 1e-3, q1, q2, form=form)
        _, em = coul(r - 1e-3, q1, q2, form=form)
        num = -(ep - em) / 2e-3
        assert abs(f - num) <= 0.02 * abs(num)


@pytest.mark.parametrize("form", COUL_FORMS)
def test_coulomb_short_range_matches_closed_form(form):
    """The erfc form of src/cuda/cuda.cu:147-180 against scipy in f64."""
    b = 0.3123
    for r in np.linspace(0.8, 9.9, 40):
        q1 = q2 = float(CHARGE_UNIT_SCALER)
        f, e = coul(r, q1, q2, alpha=b, form=form)
        e_an = q1 * q2 * erfc(b * r) / r
        f_an = q1 * q2 * (erfc(b * r) / r ** 2 + 2 * b / math.sqrt(math.pi) * math.exp(-(b * r) ** 2) / r)
        # f32 rounding of βr (~1e-7) is amplified by |d ln f/d ln x| ≈ 2x² ≈ 20
        # at the top of the range
        assert abs(e - e_an) <= 3e-6 * abs(e_an)
        assert abs(f - f_an) <= 3e-6 * abs(f_an)


def test_coulomb_table_accuracy():
    """The tiles' table (cubic Hermite pieces of φ(x) = erfc(x) + (2/√π)x e^{−x²}
    and erfc(x) on [0, 4]) against scipy at 20,001 points: the relative error
    stays at f32 rounding over the whole range, near the cutoff's small
    values included (a planted error in a derivative or a coefficient of
    the Hermite pieces moves it by orders of magnitude)."""
    L = oracle_lib()
    L.orc_coul_tab_eval.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    xs = np.linspace(1e-3, 3.999, 20001)
    p, e = C.c_float(), C.c_float()
    rs, re = [], []
    for x in xs:
        xf = float(np.float32(x))
        L.orc_coul_tab_eval(np.float32(xf), C.byref(p), C.byref(e))
        phi = erfc(xf) + 2 / math.sqrt(math.pi) * xf * math.exp(-xf * xf)
        rs.append(abs(p.value / phi - 1))
        re.append(abs(e.value / erfc(xf) - 1))
    assert max(rs) < 3e-7, max(rs)
    assert max(re) < 3e-7, max(re)


@pytest.mark.parametrize("form", COUL_FORMS)
def test_coulomb_short_range_direction(form):
    """tests/forces.rs:428-462: like charges repel, opposite attract."""
    q = float(CHARGE_UNIT_SCALER)
    for r in [3.0, 5.0, 8.0]:
        f, e = coul(r, q, q, form=form)
        assert f > 0 and e > 0
        f, e = coul(r, q, -q, form=form)
        assert f < 0 and e < 0


@pytest.mark.parametrize("form", COUL_FORMS)
def test_coulomb_short_range_cutoff(form):
    """tests/forces.rs:490-507: exactly zero at and beyond the cutoff."""
    q = float(CHARGE_UNIT_SCALER)
    for r in [COULOMB_CUTOFF, COULOMB_CUTOFF + 0.5, COULOMB_CUTOFF + 2.0]:
        assert coul(r, q, -q, form=form) == (0.0, 0.0)


def test_erf_pair_matches_closed_form():
    L = oracle_lib()
    from scipy.special import erf
    b = 0.3123
    for r in np.linspace(0.1, 9.0, 60):
        s, e = C.c_float(), C.c_float()
        L.orc_coul_erf(np.float32(r * r), np.float32(100.0), np.float32(b), C.byref(s), C.byref(e))
        e_an = 100.0 * erf(b * r) / r
        f_an = 100.0 * (erf(b * r) / r ** 2 - 2 * b / math.sqrt(math.pi) * math.exp(-(b * r) ** 2) / r)
        assert abs(e.value - e_an) <= 1e-6 * abs(e_an)
        assert abs(s.value * r - f_an) <= 2e-6 * abs(e_an) / r


@pytest.mark.parametrize("order", [4, 5])
def test_bspline_weights(order):
    """B-spline weights against the recursive definition (f64), partition of
    unity and zero-sum derivatives."""
    from oracle import fp64
    L = oracle_lib()
    fn = L.orc_bspline4 if order == 4 else L.orc_bspline5
    for w in np.linspace(0, 0.999, 37):
        th = np.zeros(order, np.float32)
        dth = np.zeros(order, np.float32)
        fn(np.float32(w), th.ctypes.data_as(C.c_void_p), dth.ctypes.data_as(C.c_void_p))
        x = w + (order - 1 - np.arange(order))
        m, dm = fp64.bspline(x, order)
        assert np.max(np.abs(th - m)) < 3e-7
        assert np.max(np.abs(dth - dm)) < 3e-7
        assert abs(th.astype(np.float64).sum() - 1) < 3e-7
        assert abs(dth.astype(np.float64).sum()) < 3e-7


def test_min_image_reference_rounding():
    """The neighbor build's min image rounds half away from zero, as Rust's
    f32::round (src/barostat.rs:165-168), unlike rint's half-to-even."""
    L = oracle_lib()
    ext = 10.0
    assert L.orc_min_image_ref(5.0, ext) == -5.0    # 0.5 → 1 (away), rint would give 0
    assert L.orc_min_image_ref(-5.0, ext) == 5.0
    assert L.orc_min_image_ref(15.0, ext) == -5.0   # 1.5 → 2
    assert L.orc_min_image_ref(25.0, ext) == -5.0   # 2.5 → 3, rint gives 2 (→ +5)
    assert L.orc_norm2_ref(1.0, 2.0, 3.0) == 14.0
