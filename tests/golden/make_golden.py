"""Writes the golden fixtures of the nonbonded evaluation (run from the
repository root: python tests/golden/make_golden.py).

The reference holds no stored input/output vectors for this path (its tests
compute closed forms in place), so the fixtures are the FP64 oracle's
evaluation of seeded synthetic systems, kept so that the GPU tests and the
oracle can be checked against fixed numbers:

* ``opc2048_box39.4`` — config 0: 2,048 OPC waters, 39.4 Å box, rc 10 Å,
  β 0.3123, SPME 40³ order 4, reference semantics;
* ``chain70_opc600_box27`` — a 30-heavy + 40-H chain in 600 waters, 27 Å
  box, SPME 30³ order 4, in reference and in standard (textbook Ewald) mode.

Each fixture holds the inputs' SHA-256 (site positions and charges as
float32, so a change of the system generator is caught), the site forces
(float32, kcal/(mol·Å), site order: solute atoms, then O, M, H0, H1 per
water) and the energies, virials and pair count as float64.
"""

from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from dynamics_amd import systems  # noqa: E402
from oracle import fp64  # noqa: E402

RC = 10.0
BETA = float(np.float32(0.3123))  # the engine takes β as f32

CASES = {
    # name: (system builder, SPME grid, modes)
    "opc2048_box39.4": (systems.config0, (40, 40, 40), ("reference",)),
    "chain70_opc600_box27": (systems.small_solvated, (30, 30, 30), ("reference", "standard")),
}

SCALARS = ("e_lj", "e_coul", "e_recip", "e_self", "e_corr", "e_total", "e_reference",
           "virial_sr", "virial_recip", "pair_count")


def input_digest(system) -> str:
    pos, q, _, _, _, _ = system.site_arrays()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(pos, dtype=np.float32).tobytes())
    h.update(np.ascontiguousarray(q, dtype=np.float32).tobytes())
    return h.hexdigest()


def fixture_path(name: str, mode: str) -> str:
    return os.path.join(HERE, f"{name}.{mode}.npz")


def evaluate(system, grid, mode) -> dict:
    r = fp64.reference_energy_and_forces(system, RC, RC, BETA, grid, 4, mode=mode)
    out = {k: np.float64(r[k]) for k in SCALARS}
    out["forces"] = r["forces"].astype(np.float32)
    return out


def main() -> None:
    for name, (build, grid, modes) in CASES.items():
        s = build()
        for mode in modes:
            r = evaluate(s, grid, mode)
            np.savez_compressed(fixture_path(name, mode), digest=np.array(input_digest(s)),
                                grid=np.array(grid, np.int32), cutoff=np.float64(RC),
                                beta=np.float64(BETA), **r)
            print(name, mode, r["e_total"], int(r["pair_count"]))


if __name__ == "__main__":
    main()
