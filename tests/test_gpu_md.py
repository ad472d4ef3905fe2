"""Device-resident MD step on the GPU against the FP64 restatement of the
reference step: positions and velocities after 1 and 3 Langevin-middle steps
of the solvated chain (bonded terms, SHAKE/RATTLE on bonds to hydrogen,
rigid water); rigid geometry and constraints over many steps; energy
conservation of the γ = 0 step; the Langevin temperature."""

import numpy as np
import pytest

from dynamics_amd import api, systems
from oracle import md_fp64

pytestmark = pytest.mark.gpu

BETA = float(np.float32(0.3123))
F64 = np.float64


def _setup(s, ff, mode=api.MODE_REFERENCE, K=(27, 27, 27), rc=10.0, gamma=1.0, dt=0.002,
           buffer=1.0, seed=5, staged=False):
    eng = api.Engine(s, cutoff=rc, lj_cutoff=rc, beta=BETA, grid=K, order=4, mode=mode,
                     list_buffer=buffer)
    m = api.Md(eng, ff, dt=dt, temperature=300.0, gamma=gamma, shake_tol=1e-5, seed=seed,
               staged_builds=staged)
    v, vw = systems.initial_velocities(s, ff)
    m.set_state(s.pos_dyn, v, systems.water_positions(s), vw)
    return eng, m, v, vw


def _min_image(d, box):
    return d - box * np.round(d / box)


@pytest.mark.parametrize("staged", [False, True])
def test_steps_vs_fp64(gpu, staged):
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    eng, m, v, vw = _setup(s, ff, staged=staged)
    model = md_fp64.Model(s, ff, 10.0, BETA, (27, 27, 27), dt=0.002, gamma=1.0, seed=5)
    st = md_fp64.MdState(s.pos_dyn.astype(F64), v.astype(F64),
                         systems.water_positions(s).astype(F64), vw.astype(F64))
    box = s.box.astype(F64)
    for n in (1, 2):
        m.step(n)
        for _ in range(n):
            model.step(st)
        pd, vd, pw, vwg = m.state()
        dx_dyn = np.abs(pd - st.x_dyn).max()
        dx_w = np.abs(_min_image(pw - st.x_w, box)).max()
        dv = max(np.abs(vd - st.v_dyn).max(), np.abs(vwg - st.v_w).max())
        print(f"after {st.step} steps: |dx| solute {dx_dyn:.2e} water {dx_w:.2e} A, "
              f"|dv| {dv:.2e} A/ps", flush=True)
        assert dx_dyn < 1e-4 and dx_w < 1e-4
        assert dv < 2e-3
    # the last step of each call evaluates nonbonded energies: a fresh
    # evaluation at the same positions agrees — it may use the list staged
    # during that step (the same pairs, summed in another order), so counts
    # agree exactly and energies to f32 summation order
    e_step = eng.energies()
    eng.compute(api.COMPUTE_ALL)
    e_again = eng.energies()
    for k in ("list_builds", "list_entries"):
        e_step.pop(k), e_again.pop(k)
    for k, v in e_step.items():
        if isinstance(v, int):
            assert e_again[k] == v, k
        else:
            a, b = np.asarray(e_again[k], np.float64), np.asarray(v, np.float64)
            assert np.all(np.abs(a - b) <= 1e-7 * (np.abs(b) + 1.0)), (k, v, e_again[k])
    m.close()
    eng.close()


def test_geometry_and_constraints_hold(gpu):
    """200 Langevin steps from a relaxed start: rigid water keeps its
    geometry, constrained bonds their lengths, the temperature stays sane."""
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    eng, m, v, vw = _setup(s, ff)
    m.close()
    pd0, pw0 = api.relax(eng, ff, s.pos_dyn, systems.water_positions(s), steps=300)
    m = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=5)
    m.set_state(pd0, v, pw0, vw)
    m.step(200)
    pd, vd, pw, vwg = m.state()
    box = s.box.astype(F64)
    o, mm, h0, h1 = (pw[:, k].astype(F64) for k in range(4))
    oh0 = np.linalg.norm(_min_image(h0 - o, box), axis=1)
    oh1 = np.linalg.norm(_min_image(h1 - o, box), axis=1)
    om = np.linalg.norm(_min_image(mm - o, box), axis=1)
    assert np.abs(oh0 - 0.87243313).max() < 2e-4 and np.abs(oh1 - 0.87243313).max() < 2e-4
    assert np.abs(om - 0.15939833).max() < 1e-4
    i, j = ff.constraints[:, 0].astype(int), ff.constraints[:, 1].astype(int)
    r = np.linalg.norm(pd[j].astype(F64) - pd[i].astype(F64), axis=1)
    assert np.abs(r - ff.constraint_r0).max() < 1e-4
    E = m.energies()
    print(f"after 200 steps: T {E['temperature']:.1f} K", flush=True)
    assert E["steps"] == 200 and 150 < E["temperature"] < 450
    assert np.isfinite(pd).all() and np.isfinite(vd).all()
    m.close()
    eng.close()


def test_nve_energy_conservation(gpu):
    """γ = 0 with standard-mode (conservative) forces: the total energy drifts
    far less than the kinetic energy fluctuates."""
    s = systems.small_water()
    ff = systems.water_force_field(s)
    K = api.pme_grid_n(s.box, 1.0)
    eng, m, _, _ = _setup(s, ff, mode=api.MODE_STANDARD, K=K, gamma=0.0, dt=0.001)
    tot, kin = [], []
    for _ in range(20):
        m.step(20)
        e = m.energies()
        pot = eng.energies()["total"]
        kin.append(e["kinetic"])
        tot.append(e["kinetic"] + pot)
    tot, kin = np.array(tot), np.array(kin)
    print(f"NVE: ptp(E) {np.ptp(tot):.4f}, ptp(KE) {np.ptp(kin):.4f} kcal/mol", flush=True)
    assert np.ptp(tot) < 0.1 * np.ptp(kin)
    m.close()
    eng.close()


def test_langevin_temperature(gpu):
    s = systems.small_water()
    ff = systems.water_force_field(s)
    K = api.pme_grid_n(s.box, 1.0)
    eng, m, _, _ = _setup(s, ff, K=K, gamma=10.0)
    m.step(1000)
    T = []
    for _ in range(20):
        m.step(50)
        T.append(m.energies()["temperature"])
    print(f"Langevin: mean T {np.mean(T):.1f} K (rigid-body kinetic energy, 6 dof per water)",
          flush=True)
    assert abs(np.mean(T) - 300.0) < 12.0
    m.close()
    eng.close()


def test_md_errors(gpu):
    s = systems.small_water()
    ff = systems.water_force_field(s)
    eng = api.Engine(s, cutoff=10.0, beta=BETA, grid=(25, 25, 25), order=4)
    with pytest.raises(api.DynError):
        api.Md(eng, ff, dt=0.05)
    with pytest.raises(api.DynError):
        api.Md(eng, ff, gamma=-1.0)
    m = api.Md(eng, ff)
    with pytest.raises(api.DynError):
        m.step(1)  # state not set


def test_spme_cadence(gpu):
    """spme_ratio = 2 (the reference's SPME_RATIO): SPME on every other force
    evaluation, the kept SPME forces in between.  From a relaxed start the
    trajectory stays close to the every-step one over a few steps, and
    differs from it."""
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=BETA, grid=(27, 27, 27), order=4,
                     mode=api.MODE_REFERENCE, list_buffer=1.0)
    pd0, pw0 = api.relax(eng, ff, s.pos_dyn, systems.water_positions(s), steps=300)
    eng.close()
    out = []
    for ratio in (1, 2):
        eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=BETA, grid=(27, 27, 27), order=4,
                         mode=api.MODE_REFERENCE, list_buffer=1.0)
        m = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=5,
                   spme_ratio=ratio)
        v, vw = systems.initial_velocities(s, ff)
        m.set_state(pd0, v, pw0, vw)
        m.step(4)
        out.append(m.state())
        m.close()
        eng.close()
    dx = max(np.abs(out[0][0] - out[1][0]).max(), np.abs(out[0][2] - out[1][2]).max())
    print(f"spme_ratio 2 vs 1 after 4 steps: |dx| {dx:.2e} A", flush=True)
    # kept steps lack the 1-4 vacuum correction (as the reference's do): up
    # to a few kcal/mol/Å on charged chain atoms, a few 0.01 Å in 4 steps
    assert 0.0 < dx < 5e-2


@pytest.mark.parametrize("gamma,nst", [(1.0, 0), (0.0, 0), (1.0, 2), (0.0, 2)])
def test_joined_steps_bit_exact(gpu, gamma, nst):
    """Inside a call of several steps, one launch ends a step and begins the
    next (k_md_step_ba); one-step calls never join.  Both give the same
    trajectory bit for bit: the Langevin noise of each step keeps its index,
    and a pressure-coupling step of velocity Verlet (the barostat after the
    last half-kick) is not joined."""
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=BETA, grid=(27, 27, 27), order=4,
                     mode=api.MODE_REFERENCE, list_buffer=1.0)
    pd0, pw0 = api.relax(eng, ff, s.pos_dyn, systems.water_positions(s), steps=100)
    eng.close()
    out = []
    for split in ((5,), (1, 1, 1, 1, 1), (2, 3)):
        # a 2 Å buffer: no list rebuild inside the 5 steps, whichever call
        # reads the displacements
        eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=BETA, grid=(27, 27, 27), order=4,
                         mode=api.MODE_REFERENCE, list_buffer=2.0)
        m = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=gamma, shake_tol=1e-5, seed=5,
                   nstpcouple=nst)
        v, vw = systems.initial_velocities(s, ff)
        m.set_state(pd0, v, pw0, vw)
        for n in split:
            m.step(n)
        e = m.energies()
        out.append((m.state(), eng.energies()["list_builds"], e["pcouple_steps"], e["box"]))
        m.close()
        eng.close()
    ref = out[0]
    print(f"gamma {gamma} nstpcouple {nst}: list builds {[o[1] for o in out]}, "
          f"pressure couplings {[o[2] for o in out]}", flush=True)
    for o in out[1:]:
        assert o[1] == ref[1] and o[2] == ref[2] and o[3] == ref[3]
        for a, b in zip(o[0], ref[0]):
            assert np.array_equal(a, b)
    if nst:
        assert ref[2] == 2


def _unlisted_pairs(s, dbg, rc=10.0):
    """Interacting site pairs within rc, at the slot positions the exported
    list was last used with, that the list lacks in either direction."""
    from conftest import NULL_ENTRY
    from test_oracle_list import INTERACT, _classes_gids
    slot_site = dbg["slot_site"].astype(np.int64)
    ns = s.n_dyn + 4 * s.n_water
    cls, gid = _classes_gids(s)
    ok = slot_site >= 0
    slot_of = np.empty(ns, np.int64)
    slot_of[slot_site[ok]] = np.nonzero(ok)[0]
    p = np.empty((ns, 3))
    p[slot_site[ok]] = dbg["xyzq"][ok, :3]
    row_ptr = dbg["row_ptr"].astype(np.int64)
    cj = dbg["cj"][:8 * row_ptr[-1]]
    jslot = (cj & NULL_ENTRY).astype(np.int64)
    ent_row = np.repeat(np.repeat(np.arange(len(row_ptr) - 1), np.diff(row_ptr)), 8)
    listed = set(zip(ent_row.tolist(), jslot.tolist()))
    box = np.asarray(s.box, np.float64)
    missing = []
    for a in range(ns):
        d = p[a + 1:] - p[a]
        d -= box * np.round(d / box)
        r2 = (d * d).sum(1)
        b = np.nonzero(r2 < rc * rc)[0] + a + 1
        b = b[INTERACT[cls[a], cls[b]] & (gid[b] != gid[a])]
        for bb in b:
            if (slot_of[a] // 8, slot_of[bb]) not in listed or (slot_of[bb] // 8, slot_of[a]) not in listed:
                missing.append((a, int(bb)))
    return missing


def test_staged_lists_cover_cutoff(gpu):
    """Staged list builds (beside the pair kernel, current from the next
    evaluation on): after each call the list the last evaluation used holds
    every interacting pair within the cutoff at the positions it was used
    with, and its energies equal a fresh evaluation's at those positions."""
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    eng, m, _, _ = _setup(s, ff, buffer=1.0, staged=True)
    for call, n in enumerate((1, 2, 3, 5)):
        m.step(n)
        dbg = eng.debug_export()
        miss = _unlisted_pairs(s, dbg)
        e = eng.energies()
        print(f"call {call}: {n} steps, builds {e['list_builds']}, pairs {e['pair_count']}, "
              f"unlisted {len(miss)} {miss[:4]}", flush=True)
        assert not miss
    m.close()
    eng.close()
