"""Molecule-pair energy matrix (dyn_engine_set_molecules /
dyn_engine_get_mol_energies; the reference's energy_between_mols,
src/non_bonded.rs:350-361) against the FP64 oracle's matrix.  Entries are
sums of f32 pair energies in 2⁻³² fixed point: held to 5e-6 of the largest
entry (the bar of the engine's energy components), plus the energy of a
pair that f32 rounding of r² may put on either side of the cutoff."""

import numpy as np
import pytest

from dynamics_amd import api, systems
from oracle import fp64

pytestmark = pytest.mark.gpu

RC = 10.0
BETA = float(np.float32(0.3123))
# LJ energy of one pair at the 10 Å cutoff (σ 3.4 Å, ε 0.2 kcal/mol) bounds
# what a pair within rounding of the cutoff can move an entry by
EDGE = 4 * 0.2 * (3.4 / RC) ** 6


def _check(s, starts, K, alch=None, lam=0.0):
    eng = api.Engine(s, cutoff=RC, lj_cutoff=RC, beta=BETA, grid=K, order=4,
                     mode=api.MODE_REFERENCE)
    if alch:
        eng.set_alchemical(alch[0], alch[1], lam)
    eng.set_molecules(starts)
    eng.compute(api.COMPUTE_ALL)
    M = eng.mol_energies()
    a = None
    if alch:
        a = np.zeros(s.n_dyn, bool)
        a[alch[0]:alch[0] + alch[1]] = True
    ref = fp64.pair_sum_fp64(s, RC, RC, BETA, alch=a, lam=lam, mol_start=starts)["mol_energies"]
    err = np.abs(M - ref).max()
    tol = 5e-6 * np.abs(ref).max() + EDGE
    print(f"{len(starts)} molecules: max |dM| {err:.3e} (tol {tol:.3e}), "
          f"max |M| {np.abs(ref).max():.3e}", flush=True)
    assert err <= tol
    assert np.allclose(M, M.T, rtol=0, atol=1e-9 * max(1.0, np.abs(M).max()))
    return eng, M


@pytest.mark.parametrize("starts", [[0], [0, 12, 30, 47], [0, 1, 69]])
def test_mol_energies_small(gpu, starts):
    eng, _ = _check(systems.small_solvated(), starts, (27, 27, 27))
    eng.close()


def test_mol_energies_alchemical(gpu):
    eng, _ = _check(systems.small_solvated(), [0, 4, 16, 40], (27, 27, 27), alch=(4, 12), lam=0.3)
    eng.close()


def test_mol_energies_bench_system(gpu):
    """The 2,600-atom chain of the 37,000-site box cut into 26 molecules."""
    from test_gpu_engine import relaxed_config1
    s = relaxed_config1()
    eng, M = _check(s, list(range(0, s.n_dyn, 100)), (64, 64, 64))
    # the same evaluation again: bit for bit (fixed-point sums)
    eng.compute(api.COMPUTE_ALL | api.COMPUTE_REBUILD)
    assert np.array_equal(eng.mol_energies(), M)
    eng.close()


def test_mol_energies_errors(gpu):
    s = systems.small_solvated()
    eng = api.Engine(s, cutoff=RC, beta=BETA, grid=(27, 27, 27), order=4)
    with pytest.raises(api.DynError, match="no molecules"):
        eng.mol_energies()
    for bad in ([1, 5], [0, 5, 5], [0, 9, 3], [0, s.n_dyn]):
        with pytest.raises(api.DynError):
            eng.set_molecules(bad)
    eng.set_molecules([0, 30])
    with pytest.raises(api.DynError, match="no compute"):
        eng.mol_energies()
    eng.compute(api.COMPUTE_ALL | api.COMPUTE_NO_ENERGY)
    with pytest.raises(api.DynError, match="no compute"):
        eng.mol_energies()
    eng.compute(api.COMPUTE_ALL)
    M = eng.mol_energies()
    assert M.shape == (2, 2) and np.all(np.isfinite(M))
    eng.set_molecules([])                     # off
    with pytest.raises(api.DynError, match="no molecules"):
        eng.mol_energies()
    eng.compute(api.COMPUTE_ALL)              # and evaluations go on without it
    eng.close()
    w = systems.small_water()
    weng = api.Engine(w, cutoff=RC, beta=BETA)
    with pytest.raises(api.DynError):
        weng.set_molecules([0])               # no solute atoms
    weng.close()
