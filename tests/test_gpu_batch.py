"""Batches of independent systems (dyn_batch_*): every kernel of a step is one
launch for all members.  With the list rebuilt every step (the same lists
either way) a batch reproduces each member's own run bit for bit — systems
of uneven sizes and PME grids, alchemical λ-windows (two SPME evaluations
each) mixed with plain systems.  Each member decides on its own list builds (on the
device), so with a list buffer too a batch reproduces its members' own runs
bit for bit, also when the steps run as captured graphs."""

import numpy as np
import pytest

from dynamics_amd import api, systems

pytestmark = pytest.mark.gpu

BETA = float(np.float32(0.3123))


def _members():
    """(system, force field, alchemical range or None, λ) of an uneven set."""
    solv = systems.small_solvated()
    return [
        (systems.small_water(), None, 0.0),
        (solv, None, 0.0),
        (systems.small_water(n_water=343, box=23.0, lattice_n=7, seed=5), None, 0.0),
        (solv, (4, 12), 0.3),
        (solv, (4, 12), 0.8),
    ]


def _make(s, alch, lam, buffer, seed, staged=False):
    ff = systems.chain_force_field(s) if s.n_dyn else systems.water_force_field(s)
    K = api.pme_grid_n(s.box, 1.0)
    eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=BETA, grid=K, order=4,
                     mode=api.MODE_REFERENCE, list_buffer=buffer)
    # quench the built structures' close contacts first, as a real run does
    pd, pw = api.relax(eng, ff, s.pos_dyn, systems.water_positions(s), steps=200)
    if alch:
        eng.set_alchemical(alch[0], alch[1], lam)
    m = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=seed,
               staged_builds=staged)
    v, vw = systems.initial_velocities(s, ff)
    m.set_state(pd, v, pw, vw)
    return eng, m


def _run(buffer, steps, batched, profile=False, staged=False):
    """steps: an int, or a sequence of step-call lengths."""
    calls = [steps] if isinstance(steps, int) else list(steps)
    made = [_make(s, a, lam, buffer, seed=11 + k, staged=staged)
            for k, (s, a, lam) in enumerate(_members())]
    if batched:
        b = api.Batch([m for _, m in made])
        if profile:
            b.profile(True)
        for n in calls:
            g0, u0 = api.layout_generation(), api.arg_uploads()
            b.step(n)
            print(f"batch call of {n} steps: layout generation +{api.layout_generation() - g0}, "
                  f"argument uploads +{api.arg_uploads() - u0}, capture refusal "
                  f"{api.capture_refusal()!r}", flush=True)
        b.close()
    else:
        for e, m in made:
            if profile:
                e.profile(True)
            for n in calls:
                m.step(n)
    out = [(m.state(), m.energies(), e.energies()) for e, m in made]
    for e, m in made:
        m.close()
        e.close()
    return out


def test_batch_matches_members_bit_exact(gpu):
    print("batch of 5 (uneven sizes, two λ-windows), list rebuilt every step", flush=True)
    a = _run(0.0, 3, batched=False)
    b = _run(0.0, 3, batched=True)
    for k, ((sa, ma, ea), (sb, mb, eb)) in enumerate(zip(a, b)):
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y), f"member {k}: state differs from its own run"
        assert ma == mb, f"member {k}: MD energies differ"
        for key in ("total", "reference", "virial_sr", "virial_recip", "dh_dl"):
            assert ea[key] == eb[key], (k, key, ea[key], eb[key])


def _same_states(a, b):
    for k, ((sa, ma, ea), (sb, mb, eb)) in enumerate(zip(a, b)):
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y), f"member {k}: states differ"
        assert ea["list_builds"] == eb["list_builds"], k
        assert ea["total"] == eb["total"], k


@pytest.mark.parametrize("staged", [False, True])
def test_batch_with_list_buffer(gpu, staged):
    """With a 1 Å list buffer and three calls of 20 steps (the second records
    a graph of 16 steady steps, the third replays it), the batch and each
    member alone take the same list builds and reach the same state bit for
    bit; so does a profiled run (direct launches on one stream)."""
    print(f"batch of 5 with a 1 A list buffer, 3 x 20 steps, staged builds {staged}", flush=True)
    a = _run(1.0, (20, 20, 20), batched=False, staged=staged)
    b = _run(1.0, (20, 20, 20), batched=True, staged=staged)
    for name, r in (("alone", a), ("batch", b)):
        print(name, "repeats", [m["repeats"] for _, m, _ in r], "graph steps",
              [m["graph_steps"] for _, m, _ in r], "builds", [e["list_builds"] for _, _, e in r],
              flush=True)
    _same_states(a, b)
    assert [m["graph_steps"] for _, m, _ in a] == [32] * len(a)
    assert [m["graph_steps"] for _, m, _ in b] == [32] * len(b)
    assert all(e["list_builds"] > 1 for _, _, e in a)
    print("the same members stepped with direct launches on one stream", flush=True)
    c = _run(1.0, (20, 20, 20), batched=True, profile=True, staged=staged)
    _same_states(a, c)
    assert all(m["graph_steps"] == 0 for _, m, _ in c)


def test_batch_errors(gpu):
    s = systems.small_water()
    e1, m1 = _make(s, None, 0.0, 1.0, 1)
    e2, m2 = _make(s, None, 0.0, 1.0, 2)
    with pytest.raises(api.DynError):
        api.Batch([m1, m1])                  # one engine twice
    ff = systems.water_force_field(s)
    m3 = api.Md(e2, ff, gamma=0.0)
    with pytest.raises(api.DynError):
        api.Batch([m1, m3])                  # Langevin with plain velocity Verlet
    m3.close()
    b = api.Batch([m1, m2])
    with pytest.raises(api.DynError):
        m1.step(1)                           # stepped by the batch
    b.step(2)
    b.close()
    m1.step(1)                               # its own again
    assert m1.energies()["steps"] == 3
    for x in (m1, m2, e1, e2):
        x.close()
