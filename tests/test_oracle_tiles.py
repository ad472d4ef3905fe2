"""The engine's pair algorithm end to end on the host: the restated list
build (cluster pairs, class segments, exclusion masks) and the tile replay
(the same pair functions the kernel runs) against the FP64 pair sum of the
solvated chain — forces, and the energy split into pairs that involve a
solute atom (the reference's reported energy, src/non_bonded.rs:337-342) and
water-only pairs, plain and alchemical."""

import os
import re

import numpy as np
import pytest

from conftest import host_engine_debug, rel_rms, replay_engine_tiles, slot_forces_to_sites
from dynamics_amd import systems
from oracle import fp64

RC = 10.0
BETA = float(np.float32(0.3123))
CHUNK = int(re.search(r"#define DYN_CHUNK (\d+)", open(os.path.join(
    os.path.dirname(__file__), "..", "dynamics_amd", "csrc", "engine.h")).read()).group(1))


@pytest.fixture(scope="module")
def solvated():
    return systems.small_solvated()


@pytest.mark.parametrize("alch_on,lam", [(False, 0.0), (True, 0.0), (True, 0.3), (True, 0.7)])
def test_tiles_vs_fp64(solvated, alch_on, lam):
    s = solvated
    alch = None
    if alch_on:
        alch = np.zeros(s.n_dyn, bool)
        alch[4:16] = True
    dbg = host_engine_debug(s, RC, alch, chunk=CHUNK)
    fx, e = replay_engine_tiles(dbg, s, RC, RC, BETA, alch=alch_on, lam=lam, energies=True)
    e = e / 2.0  # every pair is evaluated from both of its rows
    kw = dict(alch=alch, lam=lam) if alch_on else {}
    ref = fp64.pair_sum_fp64(s, RC, RC, BETA, include_14=False, **kw)
    assert rel_rms(slot_forces_to_sites(fx, dbg, s.n_sites), ref["forces"]) <= 1e-5
    assert abs(e[1] - ref["e_lj_solute"]) <= 1e-5 * abs(ref["e_lj_solute"]) + 1e-5
    assert abs(e[3] - ref["e_coul_solute"]) <= 1e-5 * abs(ref["e_coul_solute"]) + 1e-5
    assert abs(e[0] + e[1] - ref["e_lj"]) <= 1e-6 * abs(ref["e_lj"])
    assert abs(e[2] + e[3] - ref["e_coul"]) <= 5e-6 * abs(ref["e_coul"])
    assert abs(e[4] - ref["virial"]) <= 5e-6 * abs(ref["virial"])
    assert abs(e[5] - ref["dh_dl"]) <= 1e-5 * max(1.0, abs(ref["dh_dl"]))
