"""SPME reciprocal space on the GPU (replaces ewald::PmeRecip): fixed-point
charge grid bit-exact against the host replay, forces/energy/virial against
the FP64 SPME oracle of identical discretisation, and the reference's own
SPME tests (src/tests/spme.rs) with the short-range term from the FP64
closed form."""

import math

import numpy as np
import pytest
from scipy.special import erfc

from conftest import rel_rms, replay_pme_spread
from dynamics_amd import api, systems
from oracle import fp64

pytestmark = pytest.mark.gpu

K_ELEC = float(np.float32(18.2223) * np.float32(18.2223))
CHARGE_UNIT_SCALER = float(np.float32(18.2223))


def _random_system(n, box, seed):
    rng = np.random.default_rng(seed)
    pos = (rng.uniform(0, 1, (n, 3)) * box).astype(np.float32)
    q = rng.normal(0, 8.0, n)
    q -= q.mean()
    return pos, q.astype(np.float32)


@pytest.mark.parametrize("order,K", [(4, (32, 36, 40)), (5, (30, 32, 48)), (4, (64, 64, 64))])
def test_pme_matches_fp64(gpu, order, K):
    box = np.array([30.0, 33.0, 37.0], np.float32)
    pos, q = _random_system(2000, box, order)
    pme = api.PmeRecip(K, box, 0.3123, order=order)
    f, e, w = pme.forces_and_virial(pos, q)
    f64, e64, w64 = fp64.spme_fp64(pos, q, box, K, order, 0.3123)
    e64 += fp64.self_energy(q, 0.3123)
    err = rel_rms(f, f64)
    assert err <= 1e-5, f"force rel RMS {err:.2e}"
    assert abs(e - e64) <= 1e-6 * abs(e64), (e, e64)
    assert abs(w - w64) <= 1e-5 * abs(w64), (w, w64)
    grid = pme.fixed_grid()
    xyzq = np.concatenate([pos, q[:, None]], axis=1)
    assert np.array_equal(grid, replay_pme_spread(xyzq, K, box, order))


def test_pme_water_config0(gpu):
    s = systems.config0()
    pos, qs, *_ = s.site_arrays()
    keep = qs != 0
    pos = np.mod(pos[keep], s.box).astype(np.float32)
    q = qs[keep]
    pme = api.PmeRecip((40, 40, 40), s.box, 0.3123, order=4)
    f, e, w = pme.forces_and_virial(pos, q)
    f64, e64, w64 = fp64.spme_fp64(pos, q, s.box, (40, 40, 40), 4, 0.3123)
    e64 += fp64.self_energy(q, 0.3123)
    assert rel_rms(f, f64) <= 1e-5
    assert abs(e - e64) <= 1e-6 * abs(e64)


def _pair(r1, r2, q1e, q2e, box_len, alpha=0.35, cutoff=10.0):
    q1, q2 = q1e * CHARGE_UNIT_SCALER, q2e * CHARGE_UNIT_SCALER
    d = np.array(r1, np.float64) - np.array(r2, np.float64)
    d -= box_len * np.round(d / box_len)
    r = float(np.linalg.norm(d))
    e_sr = q1 * q2 * erfc(alpha * r) / r if r < cutoff else 0.0
    fmag = q1 * q2 * (erfc(alpha * r) / r ** 2 + 2 * alpha / math.sqrt(math.pi) *
                      math.exp(-(alpha * r) ** 2) / r) if r < cutoff else 0.0
    f_sr = d / r * fmag
    l = np.full(3, box_len, np.float32)
    n = api.pme_grid_n(l, 1.0)
    pme = api.PmeRecip(n, l, alpha)
    pos = np.mod(np.array([r1, r2], np.float32), box_len)
    f, e = pme.forces(pos, np.array([q1, q2], np.float32))
    return f_sr, f, e_sr, e


@pytest.mark.parametrize("dist", [3.0, 5.0, 8.0])
def test_spme_energy_opposite_charges(gpu, dist):
    """src/tests/spme.rs:147-168: a neutral pair in a 50 Å box gives the
    vacuum Coulomb energy within 1 %."""
    c = 25.0
    _, _, e_sr, e_lr = _pair([c - dist / 2, c, c], [c + dist / 2, c, c], 1.0, -1.0, 50.0)
    e_vac = -K_ELEC / dist
    assert abs((e_sr + e_lr - e_vac) / e_vac) < 0.01


@pytest.mark.parametrize("dist", [3.0, 5.0, 8.0])
def test_spme_force_magnitude(gpu, dist):
    """src/tests/spme.rs:226-270 (1 %, 2 % at 8 Å)."""
    c = 25.0
    f_sr, f_lr, _, _ = _pair([c - dist / 2, c, c], [c + dist / 2, c, c], 1.0, -1.0, 50.0)
    fx = f_sr[0] + f_lr[0, 0]
    f_vac = K_ELEC / dist ** 2
    tol = 0.02 if dist >= 7 else 0.01
    assert abs((fx - f_vac) / f_vac) < tol
    assert abs(f_sr[1] + f_lr[0, 1]) < 0.01 * f_vac


@pytest.mark.parametrize("dist", [3.0, 5.0, 8.0])
def test_spme_force_like_charges(gpu, dist):
    """src/tests/spme.rs:397-440."""
    c = 25.0
    f_sr, f_lr, _, _ = _pair([c - dist / 2, c, c], [c + dist / 2, c, c], 1.0, 1.0, 50.0)
    fx = f_sr[0] + f_lr[0, 0]
    f_vac = -K_ELEC / dist ** 2
    tol = 0.02 if dist >= 7 else 0.01
    assert abs((fx - f_vac) / f_vac) < tol


def test_spme_energy_non_cubic_box(gpu):
    """src/tests/spme.rs:349-381."""
    l = np.array([60.0, 30.0, 30.0], np.float32)
    r1 = np.array([27.5, 15.0, 15.0])
    r2 = np.array([32.5, 15.0, 15.0])
    q1, q2 = CHARGE_UNIT_SCALER, -CHARGE_UNIT_SCALER
    pme = api.PmeRecip(api.pme_grid_n(l, 1.0), l, 0.35)
    _, e_lr = pme.forces(np.array([r1, r2], np.float32), np.array([q1, q2], np.float32))
    e_sr = q1 * q2 * erfc(0.35 * 5.0) / 5.0
    e_vac = -K_ELEC / 5.0
    assert abs((e_sr + e_lr - e_vac) / e_vac) < 0.01


@pytest.mark.parametrize("dist", [3.0, 5.0, 8.0])
def test_spme_force_matches_energy_gradient(gpu, dist):
    """src/tests/spme.rs:455-491: F = −dE/dx by central difference (2 %)."""
    c, h = 25.0, 0.01
    r2 = [c + dist / 2, c, c]
    _, _, ep_sr, ep_lr = _pair([c - dist / 2 + h, c, c], r2, 1.0, -1.0, 50.0)
    _, _, em_sr, em_lr = _pair([c - dist / 2 - h, c, c], r2, 1.0, -1.0, 50.0)
    f_sr, f_lr, _, _ = _pair([c - dist / 2, c, c], r2, 1.0, -1.0, 50.0)
    num = -((ep_sr + ep_lr) - (em_sr + em_lr)) / (2 * h)
    assert abs((f_sr[0] + f_lr[0, 0] - num) / num) < 0.02


def test_pme_errors(gpu):
    with pytest.raises(api.DynError):
        api.PmeRecip((7, 32, 32), np.full(3, 30.0, np.float32), 0.3)   # 7 is not 2,3,5-smooth
    with pytest.raises(api.DynError):
        api.PmeRecip((32, 32, 32), np.full(3, 30.0, np.float32), 0.3, order=6)
    pme = api.PmeRecip((32, 32, 32), np.full(3, 30.0, np.float32), 0.3)
    with pytest.raises(api.DynError):
        pme.forces(np.full((2, 3), np.nan, np.float32), np.ones(2, np.float32))
    f, e = pme.forces(np.zeros((0, 3), np.float32), np.zeros(0, np.float32))
    assert f.shape == (0, 3) and e == 0.0
