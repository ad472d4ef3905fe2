"""Full-size GPU runs of the named workloads that no other test covers:

* config 3: 16 alchemical λ-windows (λ = k/15, solute atoms [0, 40)
  decoupled) of the 37,000-site solvated chain in one batch, list rebuilt
  every step, 2 steps: bit-exact against three windows stepped alone, and
  the λ = 7/15 window's forces and dH/dλ against the FP64 oracle;
* config 4: 128 uneven solvated boxes in one batch, 2 steps: bit-exact
  against three members alone, two members against the FP64 oracle;
* config 0: 2,048 OPC waters, 100 velocity-Verlet steps at 2 fs (rigid
  water, γ = 0): the rigid geometry holds and the total energy is conserved;
* config 2 with FP64 pair arithmetic and FP64 SPME at full size (279,600
  sites) against the FP64 oracle.

Each member of a batch is checked against its own run (the plain
single-system path, src/alchemical.rs:336-361 steps one window per MdState)
and against FP64 of the same Hamiltonian.
"""

import copy

import numpy as np
import pytest

from conftest import rel_rms
from dynamics_amd import api, systems
from oracle import fp64

pytestmark = pytest.mark.gpu

BETA = float(np.float32(0.3123))
RC = 10.0


def _with_state(s, st):
    r = copy.copy(s)
    r.pos_dyn = np.ascontiguousarray(st[0])
    r.pos_o, r.pos_m, r.pos_h0, r.pos_h1 = (np.ascontiguousarray(st[2][:, k]) for k in range(4))
    return r


def _relaxed(s, ff, K, steps):
    eng = api.Engine(s, cutoff=RC, lj_cutoff=RC, beta=BETA, grid=K, order=4, list_buffer=1.0)
    pd, pw = api.relax(eng, ff, s.pos_dyn, systems.water_positions(s), steps=steps)
    eng.close()
    return pd, pw


def _alch_member(s, ff, K, lam, pd, pw, v, vw, seed):
    e = api.Engine(s, cutoff=RC, lj_cutoff=RC, beta=BETA, grid=K, order=4,
                   mode=api.MODE_REFERENCE, list_buffer=0.0)
    a0, na = systems.CONFIG3_ALCH
    e.set_alchemical(a0, na, lam)
    m = api.Md(e, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=seed)
    m.set_state(pd, v, pw, vw)
    return e, m


def test_config3_lambda_windows_batch(gpu):
    s = systems.config1()
    ff = systems.chain_force_field(s)
    K = (64, 64, 64)
    pd, pw = _relaxed(s, ff, K, 300)
    v, vw = systems.initial_velocities(s, ff)
    lams = systems.config3_lambdas(16)
    print(f"config 3: {len(lams)} windows of {s.n_sites} sites, list rebuilt every step",
          flush=True)
    made = [_alch_member(s, ff, K, lam, pd, pw, v, vw, 100 + k) for k, lam in enumerate(lams)]
    b = api.Batch([m for _, m in made])
    b.step(2)
    b.close()
    for k in (0, 7, 15):
        e1, m1 = _alch_member(s, ff, K, lams[k], pd, pw, v, vw, 100 + k)
        m1.step(2)
        for x, y in zip(m1.state(), made[k][1].state()):
            assert np.array_equal(x, y), f"window {k}: state differs from its own run"
        ea, eb = e1.energies(), made[k][0].energies()
        for key in ("total", "reference", "dh_dl", "virial_sr", "virial_recip"):
            assert ea[key] == eb[key], (k, key)
        m1.close()
        e1.close()
    # λ = 7/15 against FP64 at the window's final positions
    e7, m7 = made[7]
    st = _with_state(s, m7.state())
    F = e7.site_forces()
    E = e7.energies()
    alch = np.zeros(s.n_dyn, bool)
    alch[slice(systems.CONFIG3_ALCH[0], sum(systems.CONFIG3_ALCH))] = True
    ref = fp64.reference_energy_and_forces(st, RC, RC, BETA, K, 4, alch=alch, lam=lams[7])
    err = rel_rms(F, ref["forces"])
    print(f"window 7 (lambda {lams[7]:.4f}): force rel RMS {err:.2e}, dH/dl {E['dh_dl']:.6f} "
          f"vs FP64 {ref['dh_dl']:.6f}", flush=True)
    assert err <= 1e-5
    # dH/dλ is a sum over the decoupled set's pairs and the SPME cross term
    assert abs(E["dh_dl"] - ref["dh_dl"]) <= 1e-5 * (abs(ref["dh_dl"]) + 1.0)
    for e, m in made:
        m.close()
        e.close()


def _uneven_member(s, seed):
    ff = systems.chain_force_field(s)
    e = api.Engine(s, cutoff=RC, lj_cutoff=RC, beta=BETA, grid=api.pme_grid_n(s.box, 1.0),
                   order=4, mode=api.MODE_REFERENCE, list_buffer=0.0)
    m = api.Md(e, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=seed)
    v, vw = systems.initial_velocities(s, ff)
    m.set_state(s.pos_dyn, v, systems.water_positions(s), vw)
    return e, m


def test_config4_uneven_batch(gpu):
    S = systems.config4(128)
    print(f"config 4: {len(S)} boxes, {sum(x.n_sites for x in S)} sites", flush=True)
    made = [_uneven_member(s, 200 + k) for k, s in enumerate(S)]
    b = api.Batch([m for _, m in made])
    b.step(2)
    b.close()
    for k in (0, 57, 127):
        e1, m1 = _uneven_member(S[k], 200 + k)
        m1.step(2)
        for x, y in zip(m1.state(), made[k][1].state()):
            assert np.array_equal(x, y), f"box {k}: state differs from its own run"
        assert e1.energies()["total"] == made[k][0].energies()["total"]
        m1.close()
        e1.close()
    for k in (3, 90):
        e, m = made[k]
        st = _with_state(S[k], m.state())
        F = e.site_forces()
        ref = fp64.reference_energy_and_forces(st, RC, RC, BETA, tuple(e.pme_grid()), 4)
        err = rel_rms(F, ref["forces"])
        print(f"box {k} ({S[k].n_sites} sites, grid {tuple(e.pme_grid())}): force rel RMS "
              f"{err:.2e}", flush=True)
        assert err <= 1e-5
    for e, m in made:
        m.close()
        e.close()


def test_config0_velocity_verlet_100_steps(gpu):
    """One evaluation, then 100 velocity-Verlet steps at dt 2 fs with rigid
    OPC water (SETTLE in the reference) from Maxwell–Boltzmann velocities at
    300 K, standard mode (the conservative Hamiltonian)."""
    s = systems.config0()
    ff = systems.water_force_field(s)
    K = (40, 40, 40)
    eng = api.Engine(s, cutoff=RC, lj_cutoff=RC, beta=BETA, grid=K, order=4,
                     mode=api.MODE_STANDARD, list_buffer=1.0)
    m = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=0.0, shake_tol=1e-5, seed=1)
    v, vw = systems.initial_velocities(s, ff)
    m.set_state(s.pos_dyn, v, systems.water_positions(s), vw)
    e0 = m.energies()
    tot, kin = [], []
    for _ in range(10):
        m.step(10)
        e = m.energies()
        tot.append(e["kinetic_rigid"] + eng.energies()["total"])
        kin.append(e["kinetic_rigid"])
    assert e["steps"] == 100
    pd, vd, pw, vwg = m.state()
    box = s.box.astype(np.float64)
    mi = lambda d: d - box * np.round(d / box)
    o, mm, h0, h1 = (pw[:, k].astype(np.float64) for k in range(4))
    for h in (h0, h1):
        assert np.abs(np.linalg.norm(mi(h - o), axis=1) - 0.87243313).max() < 2e-4
    assert np.abs(np.linalg.norm(mi(mm - o), axis=1) - 0.15939833).max() < 2e-4
    tot, kin = np.array(tot), np.array(kin)
    print(f"config 0 VV: E_total drift {tot[-1] - tot[0]:.3f}, ptp {np.ptp(tot):.3f}, "
          f"KE ptp {np.ptp(kin):.3f} kcal/mol, T {e['temperature']:.1f} K "
          f"(start {e0['temperature']:.1f} K)", flush=True)
    assert np.isfinite(tot).all()
    # the total energy moves by a small fraction of what the kinetic energy
    # exchanges with the potential over the 100 steps
    assert np.ptp(tot) < 0.1 * np.ptp(kin)
    m.close()
    eng.close()


def test_config2_fp64_precision_full_size(gpu):
    s = systems.config2()
    K = (128, 128, 128)
    print(f"config 2 FP64 pair + FP64 SPME: {s.n_sites} sites", flush=True)
    eng = api.Engine(s, cutoff=9.0, lj_cutoff=9.0, beta=BETA, grid=K, order=4,
                     mode=api.MODE_REFERENCE, precision=api.PREC_PAIR_F64 | api.PREC_PME_F64)
    eng.compute(api.COMPUTE_ALL | api.COMPUTE_REBUILD)
    F = eng.site_forces()
    E = eng.energies()
    eng.close()
    ref = fp64.reference_energy_and_forces(s, 9.0, 9.0, BETA, K, 4, mode="reference")
    err = rel_rms(F, ref["forces"])
    print(f"config 2 FP64: force rel RMS {err:.2e}", flush=True)
    assert err <= 1e-9
    scale = abs(ref["e_lj"]) + abs(ref["e_coul"]) + abs(ref["e_recip"]) + abs(ref["e_self"])
    assert abs(E["total"] - ref["e_total"]) <= 1e-9 * scale
