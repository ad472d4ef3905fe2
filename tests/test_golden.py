"""The committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py from the FP64 oracle): the seeded systems still
produce the same inputs, the oracle still produces the same outputs, and the
config-0 pair count is the figure the reference's configuration implies
(≈1.44 M interacting site pairs: 10 Å cutoff, 2,048 OPC waters in a 39.4 Å
box)."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as mg  # noqa: E402

FIXTURES = [(name, mode) for name, (_, _, modes) in mg.CASES.items() for mode in modes]


def load(name, mode):
    with np.load(mg.fixture_path(name, mode), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name,mode", FIXTURES)
def test_inputs_unchanged(name, mode):
    s = mg.CASES[name][0]()
    assert str(load(name, mode)["digest"]) == mg.input_digest(s)


@pytest.mark.parametrize("name,mode", FIXTURES)
def test_oracle_reproduces_fixture(name, mode):
    build, grid, _ = mg.CASES[name]
    g = load(name, mode)
    r = mg.evaluate(build(), grid, mode)
    assert np.array_equal(r["forces"], g["forces"])
    for k in mg.SCALARS:
        assert abs(float(r[k]) - float(g[k])) <= 1e-12 * max(1.0, abs(float(g[k]))), k


def test_config0_pair_count():
    g = load("opc2048_box39.4", "reference")
    assert abs(float(g["pair_count"]) - 1.44e6) <= 0.005 * 1.44e6
    # water only: no solute-involving energy, so the reference energy is the
    # reciprocal sum plus the self term
    assert abs(float(g["e_reference"]) - float(g["e_recip"] + g["e_self"])) <= 1e-9 * abs(float(g["e_self"]))


def test_fixture_forces_balance():
    """Pair and correction forces cancel in pairs; SPME's net force is a
    small interpolation error: the sum over sites stays far below one site's
    typical force."""
    for name, mode in FIXTURES:
        F = load(name, mode)["forces"].astype(np.float64)
        typical = np.sqrt((F ** 2).sum(axis=1).mean())
        assert np.linalg.norm(F.sum(axis=0)) <= 0.05 * typical, name
