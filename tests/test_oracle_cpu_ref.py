"""Pin of the C++ restatement of the reference CPU path (oracle/cpu/cpu_ref.cpp:
build_neighbors, setup_pairs, calc_force_cpu / f_nonbonded_cpu and the
substitute SPME, src/neighbors.rs:180-281, src/non_bonded.rs:258-412,
:541-677, :950-1057), which is the timed CPU baseline.

It is held to the FP64 evaluation of the same Hamiltonian
(fp64.reference_energy_and_forces, reference mode: LJ cut at lj_cutoff,
energy of solute-involving pairs, no reciprocal exclusion correction, the
force-only 1-4 vacuum correction), itself pinned by tests/test_math_kats.py
and tests/test_oracle_spme.py.  The reference's own GPU-vs-CPU bar is 2e-4
(src/tests/force_components.rs:460-490); the f32 pair math of the CPU path
lands far inside 5e-6.
"""

import numpy as np
import pytest

from dynamics_amd import systems
from oracle import cpu_baseline, fp64

RC = 10.0
BETA = float(np.float32(0.3123))
K = (30, 30, 30)


@pytest.fixture(scope="module")
def solvated():
    s = systems.small_solvated()
    st, F = cpu_baseline.run(s, steps=1, threads=4, cutoff=RC, lj_cutoff=RC, skin=4.0,
                             alpha=BETA, K=K, order=4, want_forces=True)
    ref = fp64.reference_energy_and_forces(s, RC, RC, BETA, K, 4, mode="reference")
    sr = fp64.pair_sum_fp64(s, RC, RC, BETA)
    return s, st, F, ref, sr


def test_cpu_ref_forces_vs_fp64(solvated):
    _, _, F, ref, _ = solvated
    err = np.linalg.norm(F - ref["forces"]) / np.linalg.norm(ref["forces"])
    assert err <= 5e-6, err


def test_cpu_ref_energies_vs_fp64(solvated):
    _, st, _, ref, sr = solvated
    e_sr = sr["e_lj_solute"] + sr["e_coul_solute"]
    scale = abs(sr["e_lj_solute"]) + abs(sr["e_coul_solute"])
    assert abs(st.energy_short - e_sr) <= 1e-6 * scale
    e_rec = ref["e_recip"] + ref["e_self"]
    assert abs(st.energy_recip - e_rec) <= 1e-6 * abs(e_rec)


def test_cpu_ref_counts(solvated):
    """Interacting pairs: every non-excluded site pair within rc with a
    nonzero ε or charge product, counted once (src/non_bonded.rs:541-677
    lists each unordered pair once)."""
    _, st, _, _, sr = solvated
    assert st.n_interacting == sr["pair_count"]


def test_cpu_ref_virial_vs_fp64(solvated):
    _, st, _, _, sr = solvated
    assert abs(st.virial_short - sr["virial"]) <= 5e-6 * abs(sr["virial"])


def test_cpu_ref_planted_sign_fails(solvated):
    """The force comparison is sensitive: flipping the sign of the 1-4
    vacuum correction (a plausible slip) moves the forces by far more than
    the bar."""
    s, _, F, ref, _ = solvated
    f14, _ = fp64.corr_14_vacuum(s.pos_dyn, s.q_dyn, s.pairs_14.astype(np.int64), s.box)
    bad = F.copy()
    bad[:s.n_dyn] -= 2 * f14
    err = np.linalg.norm(bad - ref["forces"]) / np.linalg.norm(ref["forces"])
    assert err > 1e-4


def _frames(s, shifts):
    pos, _, _, _, _, _ = s.site_arrays()
    return np.stack([pos.astype(np.float32) + np.float32(d) for d in shifts])


def test_cpu_ref_trajectory_cadence(solvated):
    """The trajectory mode (the timed baseline at the reference's cadence):
    a list kept across frames that moved less than half the skin gives the
    forces of a fresh evaluation (to summation order), a move of half the
    skin rebuilds, and with SPME every second step the second of two equal
    frames reuses the first one's SPME forces exactly."""
    s, _, F, _, _ = solvated
    kw = dict(threads=4, cutoff=RC, lj_cutoff=RC, skin=4.0, alpha=BETA, K=K, order=4)
    # a uniform shift of every site keeps all pair distances: list kept
    out = cpu_baseline.run_trajectory(s, _frames(s, [0.0, 0.6, 1.1]), spme_ratio=1, **kw)
    assert out["rebuilds"] == 1 and out["steps"] == 3  # |(1.1, 1.1, 1.1)| < skin/2
    out = cpu_baseline.run_trajectory(s, _frames(s, [0.0, 0.6, 1.2]), spme_ratio=1, **kw)
    assert out["rebuilds"] == 2  # |(1.2, 1.2, 1.2)| ≥ skin/2
    # forces of the last frame, list reused and SPME kept from the first
    L = cpu_baseline._lib()
    inp, keep = cpu_baseline._input(s, RC, RC, 4.0, BETA, K, 4)
    fr = np.ascontiguousarray(_frames(s, [0.0, 0.0]))
    st = cpu_baseline.Stats()
    G = np.zeros_like(F)
    rb = cpu_baseline.C.c_int()
    assert L.cpuref_run_traj(cpu_baseline.C.byref(inp), 2, fr.ctypes.data_as(cpu_baseline.C.c_void_p),
                             2, 4, cpu_baseline.C.byref(st), G.ctypes.data_as(cpu_baseline.C.c_void_p),
                             cpu_baseline.C.byref(rb)) == 0
    assert rb.value == 1
    assert np.linalg.norm(G - F) <= 1e-12 * np.linalg.norm(F)
