"""The FP64 oracle's molecule-pair energy matrix (src/non_bonded.rs:350-361),
pinned by separability: pair energies are additive over pairs, so with the
solute cut into molecules A, B, … the diagonal M_AA is the pair energy of a
sub-system holding only A's atoms, and M_AB = E(A ∪ B) − E(A) − E(B).  Each
E comes from the oracle's plain pair sum over that sub-system (its own
exclusions and 1-4 pairs, no water), which is checked against the
reference's closed-form tests elsewhere."""

import dataclasses

import numpy as np
import pytest

from dynamics_amd import systems
from oracle import fp64

BETA = float(np.float32(0.3123))
RC = 10.0


def _subsystem(s, atoms):
    """The solute atoms `atoms` (sorted indices) alone, without water."""
    atoms = np.asarray(atoms, np.int64)
    remap = -np.ones(s.n_dyn, np.int64)
    remap[atoms] = np.arange(atoms.size)

    def pairs(p):
        p = np.asarray(p, np.int64).reshape(-1, 2)
        keep = (remap[p[:, 0]] >= 0) & (remap[p[:, 1]] >= 0)
        return remap[p[keep]].astype(np.int32).reshape(-1, 2)

    empty = np.zeros((0, 3), np.float32)
    return dataclasses.replace(
        s, pos_dyn=s.pos_dyn[atoms], q_dyn=s.q_dyn[atoms], sigma_dyn=s.sigma_dyn[atoms],
        eps_dyn=s.eps_dyn[atoms], type_dyn=s.type_dyn[atoms], pos_o=empty, pos_m=empty,
        pos_h0=empty, pos_h1=empty, bonds=pairs(s.bonds), excl_12_13=pairs(s.excl_12_13),
        pairs_14=pairs(s.pairs_14), mol_start=np.zeros(1, np.int64))


def _energy(s, atoms):
    r = fp64.pair_sum_fp64(_subsystem(s, atoms), RC, RC, BETA)
    return r["e_lj"] + r["e_coul"]


@pytest.mark.parametrize("starts", [[0], [0, 12, 30, 47], [0, 1, 69]])
def test_mol_energies_separable(starts):
    s = systems.small_solvated()
    M = fp64.pair_sum_fp64(s, RC, RC, BETA, mol_start=starts)["mol_energies"]
    bounds = list(starts) + [s.n_dyn]
    mols = [np.arange(bounds[k], bounds[k + 1]) for k in range(len(starts))]
    assert M.shape == (len(starts), len(starts))
    assert np.array_equal(M, M.T)
    E = [_energy(s, a) for a in mols]
    scale = max(1.0, np.abs(M).max())
    for a in range(len(mols)):
        assert abs(M[a, a] - E[a]) <= 1e-12 * scale, (a, M[a, a], E[a])
        for b in range(a + 1, len(mols)):
            e_ab = _energy(s, np.concatenate([mols[a], mols[b]])) - E[a] - E[b]
            assert abs(M[a, b] - e_ab) <= 1e-11 * scale, (a, b, M[a, b], e_ab)
    # the chain's 1-4 pairs and solute–solute pairs are really in there
    assert s.pairs_14.shape[0] > 0 and abs(M.sum()) > 1.0


def test_mol_energies_water_only():
    s = systems.small_water()
    M = fp64.pair_sum_fp64(s, RC, RC, BETA, mol_start=[])["mol_energies"]
    assert M.shape == (0, 0)
