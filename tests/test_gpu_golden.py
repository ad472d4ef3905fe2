"""The engine on the GPU against the committed golden fixtures (FP64 oracle
outputs, tests/golden/): forces within 1e-5 relative RMS, the total and
reference energies within 1e-6 of the energy scale, the pair count within
the pairs f32 rounding may put either side of the cutoff."""

import os
import sys

import numpy as np
import pytest

from conftest import rel_rms
from dynamics_amd import api

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as mg  # noqa: E402

pytestmark = pytest.mark.gpu

FIXTURES = [(name, mode) for name, (_, _, modes) in mg.CASES.items() for mode in modes]


@pytest.mark.parametrize("name,mode", FIXTURES)
def test_engine_matches_golden(gpu, name, mode):
    build, grid, _ = mg.CASES[name]
    s = build()
    with np.load(mg.fixture_path(name, mode), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    assert str(g["digest"]) == mg.input_digest(s)
    m = api.MODE_REFERENCE if mode == "reference" else api.MODE_STANDARD
    eng = api.Engine(s, cutoff=mg.RC, lj_cutoff=mg.RC, beta=mg.BETA, grid=grid, order=4, mode=m)
    eng.compute(api.COMPUTE_ALL)
    F = eng.site_forces()
    E = eng.energies()
    eng.close()
    err = rel_rms(F, g["forces"].astype(np.float64))
    assert err <= 1e-5, f"force rel RMS {err:.2e}"
    scale = sum(abs(float(g[k])) for k in ("e_lj", "e_coul", "e_recip", "e_self", "e_corr"))
    assert abs(E["total"] - float(g["e_total"])) <= 1e-6 * scale
    if mode == "reference":
        assert abs(E["reference"] - float(g["e_reference"])) <= 1e-6 * scale
    assert abs(E["pair_count"] - float(g["pair_count"])) <= 2
