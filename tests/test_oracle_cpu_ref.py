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
