"""Autotuner of the MD workload on one GPU.

It chooses, for a given system and force field, the knobs of the engine that
change its speed but not the Hamiltonian beyond a stated tolerance:

* the SPME B-spline order (4 or 5) and grid: among the (order, grid)
  candidates whose FP32 forces are no farther from a converged FP64 Ewald
  evaluation (FP64 pair arithmetic and SPME, order 5 on a fine grid) than
  the requested reference discretisation's, the one whose SPME evaluation
  takes the least time;
* the pair-list buffer (the rebuild interval against the pair kernel's lane
  efficiency), the work-item length of the pair kernel (j groups per wave),
  how steps are enqueued (graphs or direct launches, one stream or two) and
  where list builds run (before the pair kernel, or staged beside it): the
  combination whose MD steps take the least time, each timed over whole
  step calls after the call that records the step graphs.

The reference fixes all of these (SPME order 4 in the ewald crate, a 4 Å
neighbor skin, src/config.rs:127, :205-210); the tuner is this engine's own.
Every measurement goes through the public C ABI, so a Rust host can run the
same procedure.
"""

from __future__ import annotations

import sys
import time

import numpy as np

from . import api


def _forces(system, cutoff, beta, grid, order, precision=0):
    eng = api.Engine(system, cutoff=cutoff, lj_cutoff=cutoff, beta=beta, grid=grid, order=order,
                     mode=api.MODE_REFERENCE, list_buffer=0.0, precision=precision)
    eng.compute(api.COMPUTE_ALL)
    f = eng.site_forces()
    eng.close()
    return f


def _spme_ms(system, cutoff, beta, grid, order, reps=20):
    """Milliseconds of one SPME evaluation (spread, FFT, convolution, gather)."""
    eng = api.Engine(system, cutoff=cutoff, lj_cutoff=cutoff, beta=beta, grid=grid, order=order,
                     mode=api.MODE_REFERENCE, list_buffer=0.0)
    eng.compute(api.COMPUTE_RECIP)
    eng.profile(True)
    for _ in range(reps):
        eng.compute(api.COMPUTE_RECIP | api.COMPUTE_ASYNC)
    eng.sync()
    prof = eng.profile_read()
    eng.close()
    return sum(v[0] for k, v in prof.items() if k.startswith("pme_")) / reps


def tune_spme(system, cutoff, beta, grid, order=4, spacings=(1.25, 1.1, 1.0, 0.9),
              orders=(4, 5), fine_spacing=0.5):
    """(order, grid) of least SPME time among those at least as accurate as
    (order, grid) against a converged FP64 evaluation; returns (choice,
    table of every candidate)."""
    both = api.PREC_PAIR_F64 | api.PREC_PME_F64
    fine = api.pme_grid_n(system.box, fine_spacing)
    f_ref = _forces(system, cutoff, beta, fine, 5, precision=both)
    nrm = float(np.linalg.norm(f_ref))
    cands = {(order, tuple(grid))}
    for o in orders:
        for sp in spacings:
            cands.add((o, tuple(api.pme_grid_n(system.box, sp))))
    rows = []
    for o, K in sorted(cands):
        err = float(np.linalg.norm(_forces(system, cutoff, beta, K, o) - f_ref) / nrm)
        rows.append({"order": o, "grid": list(K), "force_rel_rms_vs_converged": err,
                     "spme_ms": _spme_ms(system, cutoff, beta, K, o)})
    base = next(r for r in rows if r["order"] == order and tuple(r["grid"]) == tuple(grid))
    ok = [r for r in rows if r["force_rel_rms_vs_converged"] <= base["force_rel_rms_vs_converged"]]
    best = min(ok, key=lambda r: r["spme_ms"])
    return best, rows


def time_md(system, ff, state, cutoff, beta, grid, order, list_buffer, item_groups, steps=40,
            seed=1, direct_launch=False, single_stream=False, staged_builds=False):
    """Milliseconds per MD step (Langevin-middle, dt 2 fs, 300 K) from a
    thermal state, timed over one call of `steps` after two untimed calls
    (the second records the step graphs)."""
    eng = api.Engine(system, cutoff=cutoff, lj_cutoff=cutoff, beta=beta, grid=grid, order=order,
                     mode=api.MODE_REFERENCE, list_buffer=list_buffer, item_groups=item_groups,
                     single_stream=single_stream)
    md = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=seed,
                direct_launch=direct_launch, staged_builds=staged_builds)
    md.set_state(*state)
    md.step(steps)
    md.step(steps)
    md.energies()
    b0 = eng.energies()["list_builds"]
    t0 = time.perf_counter()
    md.step(steps)
    md.energies()
    ms = (time.perf_counter() - t0) / steps * 1e3
    builds = eng.energies()["list_builds"] - b0
    md.close()
    eng.close()
    return ms, builds


def tune_md(system, ff, state, cutoff=10.0, beta=0.3123, grid=(64, 64, 64), order=4,
            buffers=(1.0, 1.5, 2.0, 2.5, 3.0), item_groups=(8, 16, 32), steps=40,
            spme=True, launch_modes=(False, True), stream_modes=(False, True),
            staged_modes=(False, True)):
    """The tuned settings for MD of `system` from `state` (pos_dyn, vel_dyn,
    pos_water, vel_water): a dict with the choice and the measured table."""
    table = {}
    choice = {"order": order, "grid": list(grid)}
    if spme:
        best, rows = tune_spme(system, cutoff, beta, grid, order)
        choice = {"order": best["order"], "grid": best["grid"]}
        table["spme"] = rows
    md_rows = []
    for staged in staged_modes:
        for one in stream_modes:
            for direct in launch_modes:
                for b in buffers:
                    if cutoff + b >= 0.5 * float(min(system.box)):
                        continue
                    for g in item_groups:
                        ms, builds = time_md(system, ff, state, cutoff, beta,
                                             tuple(choice["grid"]), choice["order"], b, g, steps,
                                             direct_launch=direct, single_stream=one,
                                             staged_builds=staged)
                        md_rows.append({"list_buffer": b, "item_groups": g,
                                        "direct_launch": direct, "single_stream": one,
                                        "staged_builds": staged, "ms_per_step": ms,
                                        "list_builds_per_step": builds / steps})
                        print(f"tune: buffer {b} groups {g} direct {direct} one stream {one} "
                              f"staged {staged}: {ms:.4f} ms/step", file=sys.stderr, flush=True)
    best = min(md_rows, key=lambda r: r["ms_per_step"])
    choice.update({k: best[k] for k in ("list_buffer", "item_groups", "direct_launch",
                                        "single_stream", "staged_builds", "ms_per_step")})
    table["md"] = md_rows
    return {"choice": choice, "table": table}
