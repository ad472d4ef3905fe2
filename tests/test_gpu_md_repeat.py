"""The repeat path of dyn_md_step, run deterministically through the test
hook dyn_engine_debug_inject_status (it sets the status bit of a pair list
whose work items outgrew their buffer, as the list build itself would; the
buffers are then enlarged and the call runs again):

* one flagged evaluation inside a call: the call repeats from its start and
  returns OK with the expected step count, one recorded repeat, and a state
  that agrees with the same call run without the flag; its final forces are
  a fresh engine's at the final positions;
* every evaluation flagged: the call gives up with DYN_ERR_CAPACITY and leaves
  the MD exactly at the state it started from (positions, velocities, step
  index), after which the next call runs normally;
* a repeat across a C-rescale pressure coupling restores the start's box.
"""

import dataclasses

import numpy as np
import pytest

from dynamics_amd import api, systems

pytestmark = pytest.mark.gpu

BETA = float(np.float32(0.3123))
OVERFLOW = 8  # the status bit of a pair list whose work items outgrew their buffer


def _setup(s, ff, pd, pw, v, vw, **kw):
    eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=BETA, grid=(27, 27, 27), order=4,
                     mode=api.MODE_REFERENCE, list_buffer=1.0)
    m = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=9, **kw)
    m.set_state(pd, v, pw, vw)
    return eng, m


@pytest.fixture(scope="module")
def start():
    s = systems.small_solvated()
    ff = systems.chain_force_field(s)
    eng, m = _setup(s, ff, s.pos_dyn, systems.water_positions(s),
                    *systems.initial_velocities(s, ff))
    m.close()
    pd, pw = api.relax(eng, ff, s.pos_dyn, systems.water_positions(s), steps=200)
    eng.close()
    v, vw = systems.initial_velocities(s, ff)
    return s, ff, pd, pw, v, vw


def _state_close(a, b, tol):
    for x, y in zip(a, b):
        assert np.abs(x.astype(np.float64) - y.astype(np.float64)).max() <= tol


def test_flagged_evaluation_repeats_call(gpu, start):
    s, ff, pd, pw, v, vw = start
    eng_a, m_a = _setup(s, ff, pd, pw, v, vw)
    m_a.step(1)
    m_a.step(6)
    eng_b, m_b = _setup(s, ff, pd, pw, v, vw)
    m_b.step(1)
    assert m_b.energies()["repeats"] == 0
    print("inject: overflow flag on the 3rd evaluation of a 6-step call", flush=True)
    eng_b.debug_inject_status(OVERFLOW, skip=2, count=1)
    m_b.step(6)
    E = m_b.energies()
    assert E["steps"] == 7
    assert E["repeats"] == 1
    # the repeat rebuilt its list at another step: only list-order rounding differs
    _state_close(m_a.state(), m_b.state(), 2e-3)
    # the final forces are those of the final positions
    pd_b, _, pw_b, _ = m_b.state()
    s2 = dataclasses.replace(s, pos_dyn=pd_b, pos_o=pw_b[:, 0].copy(), pos_m=pw_b[:, 1].copy(),
                             pos_h0=pw_b[:, 2].copy(), pos_h1=pw_b[:, 3].copy())
    eng_c = api.Engine(s2, cutoff=10.0, lj_cutoff=10.0, beta=BETA, grid=(27, 27, 27), order=4,
                       mode=api.MODE_REFERENCE, list_buffer=1.0)
    eng_c.compute(api.COMPUTE_ALL)
    fc, fb = eng_c.site_forces(), eng_b.site_forces()
    assert np.linalg.norm(fb - fc) <= 1e-6 * np.linalg.norm(fc)
    for h in (m_a, m_b):
        h.close()
    for e in (eng_a, eng_b, eng_c):
        e.close()


def test_give_up_restores_start_state(gpu, start):
    s, ff, pd, pw, v, vw = start
    eng, m = _setup(s, ff, pd, pw, v, vw)
    m.step(2)
    before = m.state()
    e_before = m.energies()
    print("inject: overflow flag on every evaluation (the call gives up)", flush=True)
    eng.debug_inject_status(OVERFLOW, skip=0, count=1000)
    with pytest.raises(api.DynError) as ei:
        m.step(4)
    assert ei.value.code == api.DYN_ERR_CAPACITY
    after = m.state()
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    e_after = m.energies()
    assert e_after["steps"] == e_before["steps"] == 2
    assert e_after["repeats"] == 0
    # the forces of the start were put back: the energies read the same
    assert e_after["potential_nonbonded"] == pytest.approx(e_before["potential_nonbonded"],
                                                           rel=1e-12)
    eng.debug_inject_status(0, 0, 0)
    m.step(4)
    # the same four steps from the same start without any failure
    eng_r, m_r = _setup(s, ff, pd, pw, v, vw)
    m_r.step(2)
    m_r.step(4)
    _state_close(m.state(), m_r.state(), 2e-3)
    assert m.energies()["steps"] == 6
    for h in (m, m_r):
        h.close()
    for e in (eng, eng_r):
        e.close()


def test_repeat_across_pressure_coupling(gpu, start):
    s, ff, pd, pw, v, vw = start
    kw = dict(nstpcouple=2, pressure_target=1.0, tau_p=0.5)
    eng_a, m_a = _setup(s, ff, pd, pw, v, vw, **kw)
    m_a.step(1)
    m_a.step(5)
    eng_b, m_b = _setup(s, ff, pd, pw, v, vw, **kw)
    m_b.step(1)
    print("inject: overflow flag after the first coupling of a 5-step call", flush=True)
    eng_b.debug_inject_status(OVERFLOW, skip=3, count=1)
    m_b.step(5)
    Ea, Eb = m_a.energies(), m_b.energies()
    assert Eb["repeats"] == 1
    assert Eb["pcouple_steps"] == Ea["pcouple_steps"] == 3
    np.testing.assert_allclose(Eb["box"], Ea["box"], rtol=1e-6)
    _state_close(m_a.state(), m_b.state(), 2e-3)
    for h in (m_a, m_b):
        h.close()
    for e in (eng_a, eng_b):
        e.close()
