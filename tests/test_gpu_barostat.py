"""Pressure and C-rescale pressure coupling of the device-resident step
(dyn_md_params.nstpcouple; the reference's Barostat, src/barostat.rs:324-431)
against the FP64 oracle's: the group virial and centre-of-mass kinetic
energy at a configuration, box and state after coupled steps, batches;
and the SPME grid that follows the box (regen_pme, src/non_bonded.rs:850-875)."""

import numpy as np
import pytest

from dynamics_amd import api, systems
from oracle import fp64, md_fp64

pytestmark = pytest.mark.gpu

BETA = float(np.float32(0.3123))
F64 = np.float64
K27 = (27, 27, 27)


def _setup(s, ff, seed=5, buffer=1.0, **kw):
    eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=BETA, grid=K27, order=4,
                     mode=api.MODE_REFERENCE, list_buffer=buffer)
    m = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=seed, **kw)
    v, vw = systems.initial_velocities(s, ff)
    m.set_state(s.pos_dyn, v, systems.water_positions(s), vw)
    return eng, m, v, vw


def _oracle(s, ff, v, vw, **kw):
    model = md_fp64.Model(s, ff, 10.0, BETA, K27, dt=0.002, gamma=1.0, seed=5, **kw)
    st = md_fp64.MdState(s.pos_dyn.astype(F64), v.astype(F64),
                         systems.water_positions(s).astype(F64), vw.astype(F64))
    return model, st


def _min_image(d, box):
    return d - box * np.round(d / box)


def test_pressure_vs_oracle(gpu):
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    eng, m, v, vw = _setup(s, ff)
    E = m.energies()
    model, st = _oracle(s, ff, v, vw)
    model.forces(st)
    P, W, K = model.pressure(st)
    r = fp64.reference_energy_and_forces(s, 10.0, 10.0, BETA, K27, 4)
    scale = abs(r["virial_sr"]) + abs(r["virial_kspace"]) + abs(r["virial_corr"]) + \
        abs(model.w_atomic - W)
    print(f"P {E['pressure']:.3f} vs {P:.3f} bar; W {E['virial']:.4f} vs {W:.4f} "
          f"(scale {scale:.1f}); K {E['kinetic_com']:.6f} vs {K:.6f}", flush=True)
    assert abs(E["kinetic_com"] - K) <= 1e-7 * K   # water masses are f32 on the device
    assert abs(E["virial"] - W) <= 2e-6 * scale
    V = float(np.prod(s.box.astype(F64)))
    assert abs(E["pressure"] - P) <= 2e-6 * scale / (3 * V) * md_fp64.BAR_PER_KCAL_MOL_A3 + 1e-9 * abs(P)
    assert E["volume"] == pytest.approx(V, rel=1e-7) and E["pcouple_steps"] == 0
    m.close()
    eng.close()


@pytest.mark.parametrize("nst", [1, 2])
def test_barostat_steps_vs_oracle(gpu, nst):
    """The unrelaxed box starts near 10⁵ bar: a short τ_p makes each
    coupling move the box by a measurable amount."""
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    kw = dict(nstpcouple=nst, pressure_target=1.0, tau_p=0.5, compressibility=4.5e-5)
    eng, m, v, vw = _setup(s, ff, **kw)
    model, st = _oracle(s, ff, v, vw, **kw)
    m.step(4)
    for _ in range(4):
        model.step(st)
    E = m.energies()
    box_g = np.array(E["box"])
    print(f"nstpcouple {nst}: box {box_g} vs {model.box}, couplings {E['pcouple_steps']}, "
          f"mu log {[round(x[4], 8) for x in model.pressure_log]}", flush=True)
    assert E["pcouple_steps"] == 4 // nst == len(model.pressure_log)
    assert np.abs(box_g / model.box - 1).max() < 2e-6
    assert abs(model.box[0] / s.box[0] - 1) > 1e-4          # the box did move
    pd, vd, pw, vwg = m.state()
    dx_dyn = np.abs(pd - st.x_dyn).max()
    dx_w = np.abs(_min_image(pw - st.x_w, model.box)).max()
    dv = max(np.abs(vd - st.v_dyn).max(), np.abs(vwg - st.v_w).max())
    print(f"  |dx| solute {dx_dyn:.2e} water {dx_w:.2e} A, |dv| {dv:.2e} A/ps", flush=True)
    assert dx_dyn < 2e-4 and dx_w < 2e-4 and dv < 4e-3
    m.close()
    eng.close()


def test_barostat_relaxes_pressure(gpu):
    """60 steps coupled every 5th: the over-compressed box expands and its
    pressure falls."""
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    eng, m, _, _ = _setup(s, ff, nstpcouple=5, tau_p=0.5)
    p0 = m.energies()["pressure"]
    vols = []
    for _ in range(6):
        m.step(10)
        E = m.energies()
        vols.append(E["volume"])
    print(f"pressure {p0:.0f} -> {E['pressure']:.0f} bar, volume {vols}", flush=True)
    assert E["pcouple_steps"] == 12
    assert vols[-1] > vols[0] > float(np.prod(s.box.astype(F64)))
    assert E["pressure"] < p0
    m.close()
    eng.close()


def test_batch_with_barostat(gpu):
    """Members of a coupled batch step as they do alone (lists rebuilt every
    step: the same lists either way)."""
    def run(batched):
        made = []
        for k, s in enumerate([systems.small_solvated(), systems.small_water()]):
            ff = systems.chain_force_field(s) if s.n_dyn else systems.water_force_field(s)
            made.append(_setup(s, ff, seed=3 + k, buffer=0.0, nstpcouple=1, tau_p=0.5)[:2])
        if batched:
            b = api.Batch([m for _, m in made])
            b.step(3)
            b.close()
        else:
            for _, m in made:
                m.step(3)
        out = [(m.state(), m.energies()) for _, m in made]
        for e, m in made:
            m.close()
            e.close()
        return out
    a, b = run(False), run(True)
    for (sa, ea), (sb, eb) in zip(a, b):
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y)
        assert ea["box"] == eb["box"] and ea["pcouple_steps"] == eb["pcouple_steps"] == 3


def test_mesh_spacing_follows_box(gpu):
    """With a mesh spacing the grid is re-planned when the box needs another
    size, and the result is that of an engine made for the new box."""
    s = systems.small_solvated()
    e1 = api.Engine(s, cutoff=10.0, beta=BETA, mesh_spacing=1.0)
    assert e1.pme_grid() == (27, 27, 27)
    big = np.full(3, 28.3, np.float32)
    e1.set_box(big)
    assert e1.pme_grid() == (30, 30, 30)
    e1.compute(api.COMPUTE_ALL)
    e2 = api.Engine(s, cutoff=10.0, beta=BETA, grid=(30, 30, 30))
    e2.set_box(big)
    e2.compute(api.COMPUTE_ALL)
    assert np.array_equal(e1.fixed_forces(), e2.fixed_forces())
    assert e1.energies()["recip"] == e2.energies()["recip"]
    e1.set_box(s.box)      # back: the 27³ grid again, the first result
    e1.compute(api.COMPUTE_ALL)
    e3 = api.Engine(s, cutoff=10.0, beta=BETA, grid=K27)
    e3.compute(api.COMPUTE_ALL)
    assert e1.pme_grid() == K27
    assert np.array_equal(e1.fixed_forces(), e3.fixed_forces())
    with pytest.raises(api.DynError):
        api.Engine(s, cutoff=10.0, beta=BETA, grid=K27, mesh_spacing=1.0)
    for e in (e1, e2, e3):
        e.close()


def test_barostat_errors(gpu):
    s = systems.small_water()
    ff = systems.water_force_field(s)
    eng = api.Engine(s, cutoff=10.0, beta=BETA, grid=(25, 25, 25))
    with pytest.raises(api.DynError):
        api.Md(eng, ff, nstpcouple=1, tau_p=0.0)
    with pytest.raises(api.DynError):
        api.Md(eng, ff, nstpcouple=1, compressibility=-1.0)
    eng2 = api.Engine(s, cutoff=10.0, beta=BETA, grid=(25, 25, 25))
    m1 = api.Md(eng, ff, nstpcouple=2)
    m2 = api.Md(eng2, ff, nstpcouple=3)
    with pytest.raises(api.DynError):
        api.Batch([m1, m2])                  # members share nstpcouple
    for x in (m1, m2, eng, eng2):
        x.close()
