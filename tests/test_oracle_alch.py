"""The FP64 oracle's alchemical path (soft-core LJ, scaled Coulomb, the
reference's three-evaluation SPME cross scaling, dH/dλ), pinned by
properties that do not restate its formulas: λ = 0 is the plain sum, dH/dλ is
the λ-derivative of the total energy, forces are −∇E, and at λ = 1 the
decoupled atoms and the rest are two non-interacting systems."""

import copy

import numpy as np
import pytest

from dynamics_amd import systems
from oracle import fp64

RC, BETA, K = 10.0, 0.3123, (27, 27, 27)


@pytest.fixture(scope="module")
def sys_alch():
    s = systems.small_solvated()
    alch = np.zeros(s.n_dyn, bool)
    alch[4:16] = True  # a 12-atom stretch of the chain (crosses bonded pairs)
    return s, alch


def _total(s, alch, lam, K=K):
    r = fp64.reference_energy_and_forces(s, RC, RC, BETA, K, 4, alch=alch, lam=lam)
    return r["e_total"], r


def test_lambda0_is_plain(sys_alch):
    s, alch = sys_alch
    e0, r0 = _total(s, alch, 0.0)
    e1, r1 = _total(s, None, 0.0)
    assert abs(e0 - e1) <= 1e-10 * abs(e1)
    assert np.allclose(r0["forces"], r1["forces"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("lam", [0.2, 0.35, 0.65, 0.8])
def test_dhdl_is_energy_derivative(sys_alch, lam):
    s, alch = sys_alch
    h = 1e-4
    e_p, _ = _total(s, alch, lam + h)
    e_m, _ = _total(s, alch, lam - h)
    _, r = _total(s, alch, lam)
    fd = (e_p - e_m) / (2 * h)
    assert abs(r["dh_dl"]) > 1e-3
    assert abs(fd - r["dh_dl"]) <= 1e-6 * abs(r["dh_dl"]) + 1e-7, (fd, r["dh_dl"])


@pytest.mark.parametrize("lam", [0.3, 0.7])
def test_forces_are_gradient(sys_alch, lam):
    s, alch = sys_alch
    _, r = _total(s, alch, lam, K=None)
    h = 1e-5
    for atom in [4, 10, 20]:   # inside the molecule, inside, outside
        for c in range(3):
            sp, sm = copy.deepcopy(s), copy.deepcopy(s)
            sp.pos_dyn = s.pos_dyn.astype(np.float64).copy()
            sm.pos_dyn = s.pos_dyn.astype(np.float64).copy()
            sp.pos_dyn[atom, c] += h
            sm.pos_dyn[atom, c] -= h
            ep = fp64.pair_sum_fp64(sp, RC, RC, BETA, alch=alch, lam=lam)
            em = fp64.pair_sum_fp64(sm, RC, RC, BETA, alch=alch, lam=lam)
            fd = -((ep["e_lj"] + ep["e_coul"]) - (em["e_lj"] + em["e_coul"])) / (2 * h)
            got = r["forces"][atom, c]
            assert abs(fd - got) <= 1e-5 * max(1.0, abs(got)), (atom, c, fd, got)


def test_lambda1_decouples(sys_alch):
    s, alch = sys_alch
    e_full, _ = _total(s, alch, 1.0)
    # the environment alone: the molecule's charges and LJ switched off
    env = copy.deepcopy(s)
    env.q_dyn = np.where(alch, 0.0, s.q_dyn).astype(np.float32)
    env.eps_dyn = np.where(alch, 0.0, s.eps_dyn).astype(np.float32)
    e_env, _ = _total(env, None, 0.0)
    # the molecule alone: no water, everything else switched off
    mol = copy.deepcopy(s)
    mol.q_dyn = np.where(alch, s.q_dyn, 0.0).astype(np.float32)
    mol.eps_dyn = np.where(alch, s.eps_dyn, 0.0).astype(np.float32)
    for k in ["pos_o", "pos_m", "pos_h0", "pos_h1"]:
        setattr(mol, k, np.zeros((0, 3), np.float32))
    e_mol, _ = _total(mol, None, 0.0)
    assert abs(e_full - (e_env + e_mol)) <= 1e-9 * (abs(e_env) + abs(e_mol)), (e_full, e_env, e_mol)
