"""The FP64 oracle's pressure and C-rescale barostat (oracle/md_fp64.py).

Pinned by the reference's own pressure tests (src/tests/therm_baro.rs):
no forces → the ideal-gas pressure 2K/(3V) (:191-291); two atoms with LJ
only and no velocity → W = r·F(r) (:305-417); and, for everything those
leave open (the SPME and bonded virials, the group correction), by the
identity W = −dU/dλ, U's derivative when every pressure group's centre of
mass and the box scale by λ, in standard mode (a Hamiltonian whose forces
are U's gradient)."""

import copy

import numpy as np

from dynamics_amd import systems
from oracle import fp64, md_fp64 as md

F32 = np.float32
BAR = 69476.95457055373  # bar per kcal/mol/Å³ (src/barostat.rs:22)


def _two_atoms(r, sigma=3.4, eps=0.1, box=60.0):
    c = box / 2
    z3 = np.zeros((0, 3), F32)
    z2 = np.zeros((0, 2), np.int32)
    s = systems.System(
        box=np.full(3, box, F32), pos_dyn=np.array([[c - r / 2, c, c], [c + r / 2, c, c]], F32),
        q_dyn=np.zeros(2, F32), sigma_dyn=np.full(2, sigma, F32), eps_dyn=np.full(2, eps, F32),
        type_dyn=np.zeros(2, np.int32), lj_types=np.array([[sigma, eps]], F32), pos_o=z3,
        pos_m=z3, pos_h0=z3, pos_h1=z3, bonds=z2, excl_12_13=z2, pairs_14=z2,
        mol_start=np.array([0, 1]))
    z = lambda w, dt: np.zeros((0, w), dt)
    ff = systems.ForceField(mass=np.full(2, 12.011, F32), bonds=z(2, np.uint32),
                            bond_k_r0=z(2, F32), angles=z(3, np.uint32), angle_k_t0=z(2, F32),
                            dihedrals=z(4, np.uint32), dihe_k_n_phase=z(3, F32),
                            constraints=z(2, np.uint32), constraint_r0=np.zeros(0, F32))
    return s, ff


def _state(s, v):
    return md.MdState(s.pos_dyn.astype(np.float64), np.asarray(v, np.float64),
                      np.zeros((0, 4, 3)), np.zeros((0, 3, 3)))


def test_ideal_gas_pressure():
    s, ff = _two_atoms(20.0, eps=0.0)
    v = np.array([[1.0, -2.0, 0.5], [0.3, 0.0, -1.5]])
    model = md.Model(s, ff, 10.0, 0.3123, None)
    st = _state(s, v)
    model.forces(st)
    P, W, K = model.pressure(st)
    ke = 0.5 * float(F32(12.011)) * (v ** 2).sum() / 418.4
    assert W == 0.0
    assert abs(K - ke) <= 1e-12 * ke
    assert abs(P - 2 * ke / (3 * 60.0 ** 3) * BAR) <= 1e-9 * abs(P)


def test_lj_pair_virial_pressure():
    r, sigma, eps = 4.0, 3.4, 0.1
    s, ff = _two_atoms(r, sigma, eps)
    model = md.Model(s, ff, 10.0, 0.3123, None)
    st = _state(s, np.zeros((2, 3)))
    model.forces(st)
    P, W, K = model.pressure(st)
    e = float(np.sqrt(F32(eps) * F32(eps)))
    sr6 = (sigma / r) ** 6
    f_r = 24 * e / r * (2 * sr6 * sr6 - sr6)   # > 0: repulsive
    assert K == 0.0
    assert abs(W - r * f_r) <= 1e-6 * abs(r * f_r), (W, r * f_r)
    assert abs(P - r * f_r / (3 * 60.0 ** 3) * BAR) <= 1e-6 * abs(P)


def _scaled_energy(model, st, lam):
    """U with every group's centre of mass and the box scaled by lam, and
    the number of pairs inside the cutoff."""
    m2 = copy.copy(model)
    st2 = copy.deepcopy(st)
    m2.rescale(st2, lam)
    m2.forces(st2)
    e = m2.energy
    sy = md._system_at(m2.system, st2)
    sy.box = m2.box.copy()
    n = fp64.pair_sum_fp64(sy, 10.0, 10.0, 0.3123)["pair_count"]
    return e["nonbonded"] + e["bond"] + e["angle"] + e["dihedral"], n


def test_group_virial_is_volume_derivative():
    """−dU/dλ by central differences equals the group virial W (nonbonded
    pairs, SPME k-space and exclusion corrections, bonded terms, less the
    intra-group part of constraint clusters and rigid waters)."""
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    model = md.Model(s, ff, 10.0, 0.3123, (27, 27, 27), mode="standard")
    vd, vw = systems.initial_velocities(s, ff)
    # bonds and angles away from their rest values, so the bonded virial counts
    x = s.pos_dyn.astype(np.float64) + np.random.default_rng(3).normal(0, 0.05, s.pos_dyn.shape)
    st = md.MdState(x, vd.astype(np.float64), systems.water_positions(s).astype(np.float64),
                    vw.astype(np.float64))
    model.forces(st)
    _, W, _ = model.pressure(st)
    _, _, wb = md.bonded_fp64(x, ff, model.box, virial=True)
    assert abs(wb) > 10.0
    # a step small enough that no pair crosses the cutoff (U jumps there)
    h = 1e-8
    up, n_up = _scaled_energy(model, st, 1 + h)
    dn, n_dn = _scaled_energy(model, st, 1 - h)
    assert n_up == n_dn
    fd = -(up - dn) / (2 * h)
    print(f"W {W:.6f}  -dU/dlambda {fd:.6f}")
    assert abs(fd - W) <= 1e-6 * abs(W) + 0.01, (fd, W)


def test_c_rescale_factor():
    s, ff = _two_atoms(20.0)
    model = md.Model(s, ff, 10.0, 0.3123, None, dt=0.002, temperature=0.0, nstpcouple=10,
                     pressure_target=1.0, tau_p=5.0, compressibility=4.5e-5)
    # at T = 0 the noise vanishes: μ = exp((κ/τ)(P − P₀)·nstpcouple·dt / 3)
    for P in (1.0, 500.0, -300.0):
        mu = model.c_rescale_mu(P, 7)
        assert abs(np.log(mu) * 3 - 4.5e-5 / 5.0 * (P - 1.0) * 0.02) < 1e-15
    assert model.c_rescale_mu(1.0, 3) == 1.0
    assert model.c_rescale_mu(1e12, 3) == np.exp(0.1 / 3)      # ΔlnV capped at 10 %
    assert model.c_rescale_mu(-1e12, 3) == np.exp(-0.1 / 3)
    # with temperature: the same deviate for the same step, another for the next
    model.T = 300.0
    a, b, c = model.c_rescale_mu(1.0, 3), model.c_rescale_mu(1.0, 3), model.c_rescale_mu(1.0, 4)
    assert a == b and a != c and a != 1.0


def test_rescale_keeps_groups_rigid():
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    model = md.Model(s, ff, 10.0, 0.3123, None)
    vd, vw = systems.initial_velocities(s, ff)
    st = md.MdState(s.pos_dyn.astype(np.float64), vd.astype(np.float64),
                    systems.water_positions(s).astype(np.float64), vw.astype(np.float64))
    before = copy.deepcopy(st)
    model.rescale(st, 1.01)
    i, j = ff.constraints[:, 0].astype(int), ff.constraints[:, 1].astype(int)
    assert np.allclose(st.x_dyn[j] - st.x_dyn[i], before.x_dyn[j] - before.x_dyn[i], atol=1e-12)
    assert np.allclose(st.x_w - st.x_w[:, :1], before.x_w - before.x_w[:, :1], atol=1e-12)
    assert np.allclose(model.box, s.box * 1.01)
    assert np.allclose(st.v_w, before.v_w / 1.01)
