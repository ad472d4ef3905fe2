"""The FP64 MD-step oracle, pinned by facts it does not restate: the
published SplitMix64 outputs, −∇U finite differences of the bonded terms,
rigid-body invariants of the water drift, constraint residuals after SHAKE
and RATTLE, and energy conservation of the γ = 0 step."""

import copy

import numpy as np
import pytest

from dynamics_amd import systems
from oracle import md_fp64 as md


def test_splitmix64_known_answer():
    # SplitMix64 from state 0 (Steele, Lea & Flood 2014; the reference
    # generator of xoshiro's seeding): first three outputs
    want = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = [int(md.mix64(np.uint64((k * 0x9E3779B97F4A7C15) & md.MASK))) for k in range(3)]
    assert got == want
    r = systems.SplitMix64(0)
    assert [int(v) for v in r.u64(3)] == want


def test_normal_deviates():
    z = md.normal(11, 5, np.arange(400000))
    assert abs(z.mean()) < 5e-3 and abs(z.std() - 1) < 5e-3
    z2 = md.normal(11, 6, np.arange(400000))
    assert abs(np.corrcoef(z, z2)[0, 1]) < 5e-3


@pytest.fixture(scope="module")
def chain():
    s = systems.small_solvated()
    ff = systems.chain_force_field(s, constrain_h=False)
    rng = np.random.default_rng(3)
    pos = s.pos_dyn.astype(np.float64) + rng.normal(0, 0.05, s.pos_dyn.shape)
    return s, ff, pos


def test_bonded_forces_are_gradient(chain):
    s, ff, pos = chain
    box = s.box.astype(np.float64)
    F, e = md.bonded_fp64(pos, ff, box)
    assert all(v > 0 for v in e)
    h = 1e-6
    for atom in [0, 5, 17, 29, 45]:
        for c in range(3):
            p, m = pos.copy(), pos.copy()
            p[atom, c] += h
            m[atom, c] -= h
            fd = -(sum(md.bonded_fp64(p, ff, box)[1]) - sum(md.bonded_fp64(m, ff, box)[1])) / (2 * h)
            assert abs(fd - F[atom, c]) <= 1e-5 * max(1.0, abs(F[atom, c])), (atom, c, fd, F[atom, c])
    assert np.allclose(F.sum(0), 0, atol=1e-8)


def test_rigid_drift_invariants():
    s = systems.small_water(n_water=64, box=12.4, lattice_n=4)
    xw = systems.water_positions(s).astype(np.float64)
    rng = np.random.default_rng(5)
    vw = rng.normal(0, 5, (64, 3, 3))
    box = s.box.astype(np.float64)
    x2, v2 = md.rigid_drift(xw, vw, 0.002, box)

    def geom(x):
        o, mm, h0, h1 = (x[:, k] for k in range(4))
        return (np.linalg.norm(md.min_image(h0 - o, box), axis=1),
                np.linalg.norm(md.min_image(h1 - o, box), axis=1),
                np.linalg.norm(md.min_image(h1 - h0, box), axis=1),
                np.linalg.norm(md.min_image(mm - o, box), axis=1))

    for a, b in zip(geom(xw)[:3], geom(x2)[:3]):
        assert np.allclose(a, b, atol=1e-9)
    assert np.allclose(geom(x2)[3], md.O_EP_R, atol=1e-9)
    m = np.array([md.M_O, md.M_H, md.M_H])
    p_before = (m[None, :, None] * vw).sum(1)
    p_after = (m[None, :, None] * v2).sum(1)
    assert np.allclose(p_before, p_after, atol=1e-9)
    # the velocities after the drift are those of a rigid body: every
    # intramolecular distance has zero rate of change
    o, h0, h1 = x2[:, 0], x2[:, 2], x2[:, 3]
    for (pa, va), (pb, vb) in [((o, v2[:, 0]), (h0, v2[:, 1])), ((o, v2[:, 0]), (h1, v2[:, 2])),
                               ((h0, v2[:, 1]), (h1, v2[:, 2]))]:
        r = md.min_image(pb - pa, box)
        assert np.abs(((vb - va) * r).sum(1)).max() < 1e-8


def test_shake_and_rattle():
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    model = md.Model(s, ff, 10.0, 0.3123, None, shake_tol=1e-8)
    rng = np.random.default_rng(2)
    st = md.MdState(s.pos_dyn.astype(np.float64) + rng.normal(0, 0.02, s.pos_dyn.shape),
                    rng.normal(0, 5, s.pos_dyn.shape), systems.water_positions(s).astype(np.float64),
                    np.zeros((s.n_water, 3, 3)))
    model.shake(st)
    model.rattle(st)
    i, j = ff.constraints[:, 0].astype(int), ff.constraints[:, 1].astype(int)
    d = st.x_dyn[j] - st.x_dyn[i]
    assert np.abs(np.linalg.norm(d, axis=1) - ff.constraint_r0).max() < 1e-6
    # RATTLE leaves every constrained bond's relative velocity perpendicular
    # to it only approximately (one pass); it must be small against |v|
    dv = st.v_dyn[j] - st.v_dyn[i]
    along = np.abs((dv * d).sum(1)) / np.linalg.norm(d, axis=1)
    assert along.max() < 0.5


def test_nve_energy_conservation():
    """γ = 0, standard-mode forces (conservative): the total energy drifts
    far less than the kinetic energy fluctuates."""
    s = systems.small_water(n_water=216, box=18.6, lattice_n=6)
    ff = systems.water_force_field(s)
    K = (20, 20, 20)
    model = md.Model(s, ff, 8.0, 0.35, K, mode="standard", dt=0.001, gamma=0.0)
    _, vw = systems.initial_velocities(s, ff, 300.0)
    st = md.MdState(np.zeros((0, 3)), np.zeros((0, 3)), systems.water_positions(s).astype(np.float64), vw.astype(np.float64))
    e, kin = [], []
    for _ in range(30):
        model.step(st)
        k = md.kinetic(st, model.mass)
        kin.append(k)
        e.append(k + model.energy["nonbonded"])
    e, kin = np.array(e), np.array(kin)
    assert np.ptp(e) < 0.05 * np.ptp(kin), (np.ptp(e), np.ptp(kin))
