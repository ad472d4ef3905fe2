#!/usr/bin/env python3
"""Throughput of the nonbonded force path on one MI355X.

Workload: the solvated-chain system — 64 Å cubic box, 8,600 OPC waters around
a 1,200-heavy-atom + 1,400-H random-walk chain (37,000 sites), 10 Å cutoff,
β = 0.3123 Å⁻¹, SPME 64³ order 4, reference semantics — with everything
resident in HBM.

--mode md (default): one step is a whole device-resident MD step at dt = 2 fs
(Langevin-middle, γ = 1/ps, 300 K; rigid water; SHAKE/RATTLE on bonds to
hydrogen; harmonic bonds and angles and threefold torsions on the chain):
nonbonded forces with the pair list reused inside a list buffer (rebuilt
on the device when a site has moved half the buffer), bonded forces,
integration.
--mode eval: one step is one nonbonded evaluation with the list rebuilt every
time (no integration).
ns/day = steps/s × dt (1 step/s = 0.1728 ns/day at 2 fs).

Prints one JSON line (rank 0) with the metric, the pair-kernel roofline from
HIP events over the timed region, and the CPU baseline (the oracle's C++
restatement of the reference CPU path, timed on this host's cores).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ns/day & pair-interactions/s at ≤1e-5 rel force err vs FP64, 1 GPU"
FP32_PEAK_TFLOPS = 157.3
# declared FLOP weights per interacting pair (FMA = 2; rsqrt, erfc, exp, rint = 1)
FLOPS_LJ_COUL, FLOPS_LJ, FLOPS_COUL = 63, 50, 51


# list buffer of the MD modes (Å): the device rebuilds the list when a site
# has moved half the buffer since the last build; with 32 j groups per work
# item of the pair kernel, the autotuner's choice for the MD workload
# (`--mode tune`)
MD_BUFFER = 2.0
MD_ITEM_GROUPS = 32


def equilibrate(eng, ff, s, pd, pw, steps, seed=7):
    """Untimed Langevin MD at 300 K with γ = 10/ps from the relaxed
    structure (Maxwell–Boltzmann start), so that the timed steps run on a
    thermal configuration.  Returns the state (pos_dyn, vel_dyn, pos_water,
    vel_water)."""
    from dynamics_amd import api, systems
    v, vw = systems.initial_velocities(s, ff, 300.0)
    m = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=10.0, shake_tol=1e-5, seed=seed)
    m.set_state(pd, v, pw, vw)
    done = 0
    while done < steps:  # in calls of 500 steps, each reported on stderr
        k = min(500, steps - done)
        m.step(k)
        done += k
        print(f"bench: equilibration {done}/{steps} steps", file=sys.stderr, flush=True)
    st = m.state()
    m.close()
    return st


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="about the wall time of the CPU-baseline sample (0 skips it)")
    ap.add_argument("--mode", choices=["md", "eval", "copy", "batch-lambda", "batch-uneven",
                                       "precision", "tune"],
                    default="md",
                    help="md: device-resident MD steps; eval: nonbonded evaluations only; "
                         "copy: a plain 256 MiB device copy, three times (the known-size "
                         "kernel that scales the HBM byte counters); batch-lambda: 16 "
                         "alchemical λ-windows of the solvated chain in one batch; "
                         "batch-uneven: 128 small solvated boxes of uneven sizes in one batch; "
                         "precision: the 128 A water box with the pair kernel and SPME each in "
                         "FP32 and FP64, errors against the engine's FP64 evaluation; "
                         "tune: the autotuner's table for the MD workload (SPME order and "
                         "grid, list buffer, work-item length)")
    ap.add_argument("--systems", type=int, default=None,
                    help="batch modes: number of systems (16 windows / 128 boxes)")
    ap.add_argument("--relax-steps", type=int, default=400,
                    help="untimed relaxation steps that quench the built chain's close contacts")
    ap.add_argument("--equil-steps", type=int, default=2000,
                    help="untimed Langevin steps (γ = 10/ps, 300 K) that thermalise the "
                         "relaxed structure before the timed MD")
    ap.add_argument("--direct-launch", action="store_true",
                    help="MD modes: launch every kernel directly (no step graphs)")
    ap.add_argument("--staged-builds", action="store_true",
                    help="MD modes: list builds beside the pair kernel, serving from the next step")
    ap.add_argument("--single-stream", action="store_true",
                    help="every kernel on the engine's stream (no second stream for SPME, "
                         "corrections and bonded forces)")
    ap.add_argument("--item-groups", type=int, default=MD_ITEM_GROUPS,
                    help="j groups (of 8 j slots) per work item of the pair kernel (MD modes)")
    ap.add_argument("--list-buffer", type=float, default=None,
                    help="list buffer in Å (MD modes default %.1f; eval default 0: rebuild "
                         "every step)" % MD_BUFFER)
    args = ap.parse_args()
    if args.list_buffer is None:
        # the smallest uneven box (24 A) allows rc + buffer < 12 A only
        args.list_buffer = {"eval": 0.0, "batch-uneven": 1.5}.get(args.mode, MD_BUFFER)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        return 2

    if args.mode == "copy":
        import torch
        n = 64 * 1024 * 1024
        x = torch.ones(n, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(f"copied {4 * n} bytes, 3 times", flush=True)
        return 0

    if args.mode == "batch-lambda":
        return run_batch_lambda(args)
    if args.mode == "batch-uneven":
        return run_batch_uneven(args)
    if args.mode == "precision":
        return run_precision(args)
    if args.mode == "tune":
        return run_tune(args)

    import numpy as np
    from dynamics_amd import api, systems

    s = systems.config1()
    K = (64, 64, 64)
    eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=0.3123, grid=K, order=4,
                     mode=api.MODE_REFERENCE, list_buffer=args.list_buffer,
                     item_groups=args.item_groups, single_stream=args.single_stream)
    # the built chain has close contacts: quench them first (untimed), so that
    # both modes run on a physical configuration
    ff = systems.chain_force_field(s)
    pd0, pw0 = api.relax(eng, ff, s.pos_dyn, systems.water_positions(s), steps=args.relax_steps)
    print(f"bench: relaxed ({args.relax_steps} steps)", file=sys.stderr, flush=True)
    s.pos_dyn = pd0
    s.pos_o, s.pos_m, s.pos_h0, s.pos_h1 = (np.ascontiguousarray(pw0[:, k]) for k in range(4))
    md = None
    if args.mode == "md":
        st = equilibrate(eng, ff, s, pd0, pw0, args.equil_steps)
        print(f"bench: equilibrated ({args.equil_steps} steps)", file=sys.stderr, flush=True)
        md = api.Md(eng, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=1,
                    direct_launch=args.direct_launch, staged_builds=args.staged_builds)
        md.set_state(*st)

        def run(n):
            md.step(n)

        def wait():
            md.energies()
    else:
        eng.set_positions(s)
        flags = api.COMPUTE_ALL | api.COMPUTE_ASYNC
        if args.list_buffer <= 0:
            flags |= api.COMPUTE_REBUILD

        def run(n):
            for _ in range(n):
                eng.compute(flags)

        def wait():
            eng.sync()

    # The timed steps carry HIP events around the pair kernel (two per step,
    # on its stream; graphs of steps hold them as event nodes).  Two untimed
    # calls of the timed length first: the second records the graphs of its
    # steady steps (as the first call of a production run would), then the
    # warmup steps.
    eng.profile(2)
    if md is not None:
        run(args.steps)
        run(args.steps)
    run(args.warmup)
    wait()
    print(f"bench: warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    builds0 = eng.energies()["list_builds"]
    graph0 = md.energies()["graph_steps"] if md is not None else 0
    eng.profile(2)  # the same mode: clears the totals, keeps the graphs
    up0 = api.arg_uploads()
    t0 = time.perf_counter()
    run(args.steps)
    wait()
    t1 = time.perf_counter()
    uploads = api.arg_uploads() - up0
    E = eng.energies()
    builds = E["list_builds"] - builds0
    graph_steps = (md.energies()["graph_steps"] - graph0) if md is not None else 0
    prof_pairs = eng.profile_read()
    n_both, n_lj, n_coul = E["pair_count_kind"]
    n_pairs = E["pair_count"]  # pair kernel pairs + the 1-4 pairs within the cutoff
    md_energies = md.energies() if md is not None else None
    # accuracy at the timed steps' end (thermal state, reused list): the FP32
    # forces of the last evaluation against the engine's FP64 evaluation
    # (FP64 pair arithmetic and FP64 SPME) at the same positions
    f32 = eng.site_forces()
    if md is not None:
        pd, _, pw, _ = md.state()
        s_end = systems.with_positions(s, pd, pw)
    else:
        s_end = s
    e64 = api.Engine(s_end, cutoff=10.0, lj_cutoff=10.0, beta=0.3123, grid=K, order=4,
                     mode=api.MODE_REFERENCE, list_buffer=0.0,
                     precision=api.PREC_PAIR_F64 | api.PREC_PME_F64)
    e64.compute(api.COMPUTE_ALL)
    f64 = e64.site_forces()
    e64.close()
    force_err = float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64))
    # per-stage breakdown: a separate run with every kernel timed (one stream,
    # direct launches)
    eng.profile(1)
    n_prof = max(1, min(args.steps, 50))
    run(n_prof)
    wait()
    prof = eng.profile_read()
    eng.profile(False)

    sec_per_step = (t1 - t0) / args.steps
    steps_per_s = 1.0 / sec_per_step
    ns_day = steps_per_s * 0.1728
    pair_ms, pair_n = prof_pairs.get("pairs", (float("nan"), 1))
    pair_s = pair_ms / 1e3 / max(pair_n, 1)
    flops = FLOPS_LJ_COUL * n_both + FLOPS_LJ * n_lj + FLOPS_COUL * n_coul
    achieved = flops / pair_s / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", "pairs_traffic_chain2600_opc8600.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    out = {
        "metric": METRIC,
        "value": ns_day,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("chain2600_opc8600_box64: 37,000-site solvated chain, one device-resident "
                         "MD step per step (Langevin-middle, dt 2 fs, rigid water, SHAKE on H "
                         "bonds, bonded terms, nonbonded LJ/erfc-Coulomb pairs + 1-4 corrections "
                         "+ SPME every step)" if md is not None else
                         "chain2600_opc8600_box64: 37,000-site solvated chain, one nonbonded "
                         "evaluation per step (list rebuild + LJ/erfc-Coulomb pairs + 1-4 "
                         "corrections + SPME), integrator not included"),
            "sites": s.n_sites, "cutoff_A": 10.0, "beta_inv_A": 0.3123, "pme_grid": list(K),
            "pme_order": 4, "mode": "reference", "dt_fs": 2.0, "step": args.mode,
            "list": ("rebuilt every step" if args.list_buffer <= 0 else
                     f"buffer {args.list_buffer} A, rebuilt on the device when a site has "
                     "moved half the buffer"),
            "item_groups": args.item_groups,
            "list_builds_in_timed_steps": builds,
            "steps_in_graph_launches": graph_steps,
            "launch": ("direct" if args.direct_launch else "graphs of 16 steady steps") +
                      (", one stream" if args.single_stream else ", two streams") +
                      (", staged list builds" if args.staged_builds else ""),
            "start": (f"relaxed ({args.relax_steps} steps), then {args.equil_steps} untimed "
                      "Langevin steps at 300 K, γ 10/ps" if md is not None else
                      f"relaxed ({args.relax_steps} steps)"),
            "accumulation": "int64 fixed point (2^-32), f64 energy/virial",
        },
        "force_rel_rms_vs_fp64": force_err,
        "pair_interactions_per_s": n_pairs * steps_per_s,
        "pair_interactions_per_step": n_pairs,
        "cluster_pair_entries": E["list_entries"],
        "arg_uploads_in_timed_steps": uploads,
        "tile_lane_efficiency": 2.0 * (n_both + n_lj + n_coul) / max(64.0 * E["list_entries"], 1.0),
        "roofline": {
            "bound": "valu",
            "kernel": "k_eng_pairs",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "flops_per_launch": flops,
            "avg_launch_ms": pair_s * 1e3,
            "timed_launches": pair_n,
        },
        "kernels_ms_per_step": {k: v[0] / n_prof for k, v in prof.items()
                                if not k.startswith("host:")},
        "host_ms_per_step": {k[5:]: v[0] / n_prof for k, v in prof.items() if k.startswith("host:")},
        "timed_host_ms_per_step": {k[5:]: v[0] / args.steps for k, v in prof_pairs.items()
                                   if k.startswith("host:md_")},
    }
    if md_energies is not None:
        out["md_temperature_K"] = md_energies["temperature"]
    if args.cpu_seconds > 0 and md is not None:
        out["cpu_baseline"] = cpu_baseline_md(s, md, K, args.cpu_seconds)
    print(json.dumps(out))
    return 0


def cpu_baseline_md(s, md, K, seconds):
    """The reference CPU path (the oracle's C++ restatement: brute-force
    neighbor lists with the 4 Å skin, setup_pairs, calc_force_cpu, f64 SPME
    order 4) over a bounded sample of the same MD trajectory at the
    reference's own cadence: lists rebuilt when a solute atom or water O has
    moved 2 Å, SPME every second step (src/neighbors.rs:159-169, src/lib.rs:
    208, :1282-1341), on every host core this process may use.  The frames
    come from the GPU trajectory, one per 2 fs step."""
    try:
        import numpy as np
        from oracle import cpu_baseline
        threads = cpu_baseline.host_threads()
        # size the sample from one full evaluation's cost (about `seconds`)
        one = cpu_baseline.timed(s, seconds=0.0, threads=threads, cutoff=10.0, lj_cutoff=10.0,
                                 skin=4.0, alpha=0.3123, K=K, order=4)
        per_step = one["ms_short"] + 0.5 * one["ms_spme"]
        n = int(max(100, min(400, seconds * 1e3 / max(per_step, 1e-3))))
        frames = []
        for _ in range(n):
            md.step(1)
            pd, _, pw, _ = md.state()
            frames.append(np.concatenate([pd.reshape(-1, 3), pw.reshape(-1, 3)]))
        frames = np.ascontiguousarray(np.stack(frames), np.float32)
        cb = cpu_baseline.run_trajectory(s, frames, spme_ratio=2, threads=threads, cutoff=10.0,
                                         lj_cutoff=10.0, skin=4.0, alpha=0.3123, K=K, order=4)
        return {
            "value": cb["ns_per_day"],
            "unit": "ns/day",
            "cores": cb["threads"],
            "kind": "port",
            "sample": f"{cb['steps']} consecutive 2 fs steps of the same 37,000-site MD "
                      f"trajectory ({cb['wall_s']:.1f} s) at the reference's cadence: "
                      f"{cb['rebuilds']} neighbor rebuilds (brute force, rc + 4 A skin, at 2 A "
                      "displacement) with setup_pairs, calc_force_cpu every step (f32 pair "
                      "math, per-thread f64 buffers), f64 SPME 64^3 order 4 every second step; "
                      "integration not included",
            "pair_interactions_per_s": cb["pair_interactions_per_s"],
            "ms_components_per_step": {k: cb[k] for k in ("ms_neighbors", "ms_setup_pairs",
                                                          "ms_short", "ms_spme")},
            "ns_per_day_full_evaluation_every_step": one["ns_per_day"],
            "cpu": _cpu_model(),
        }
    except Exception as e:  # the baseline must never hide the GPU number
        return {"value": None, "error": str(e)}


def _time_steps(step, wait, steps):
    wait()
    t0 = time.perf_counter()
    step(steps)
    wait()
    return time.perf_counter() - t0


def run_batch_lambda(args) -> int:
    """16 alchemical λ-windows (λ = k/15) of the 37,000-site solvated chain,
    40 decoupled solute atoms, Langevin-middle MD at dt 2 fs: one batch step
    advances every window.  Compared with one window stepped alone."""
    import numpy as np
    from dynamics_amd import api, systems

    n = args.systems or 16
    s = systems.config1()
    K = (64, 64, 64)
    ff = systems.chain_force_field(s)

    def engine():
        return api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=0.3123, grid=K, order=4,
                          mode=api.MODE_REFERENCE, list_buffer=args.list_buffer,
                          item_groups=args.item_groups)

    e0 = engine()
    pd0, pw0 = api.relax(e0, ff, s.pos_dyn, systems.water_positions(s), steps=args.relax_steps)
    pd0, v, pw0, vw = equilibrate(e0, ff, s, pd0, pw0, args.equil_steps)
    e0.close()
    print("bench: relaxed and equilibrated", file=sys.stderr, flush=True)
    lams = [k / max(n - 1, 1) for k in range(n)]
    made = []
    for k, lam in enumerate(lams):
        e = engine()
        e.set_alchemical(0, 40, lam)
        m = api.Md(e, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=100 + k)
        m.set_state(pd0, v, pw0, vw)
        made.append((e, m))
    # one window stepped alone (the plain single-system path)
    e1, m1 = made[n // 2]
    m1.step(args.warmup)
    t_one = _time_steps(m1.step, lambda: m1.energies(), args.steps)
    for e, m in made:
        m.set_state(pd0, v, pw0, vw)
    b = api.Batch([m for _, m in made])
    b.step(args.warmup)
    print("bench: warmup done", file=sys.stderr, flush=True)
    up0 = api.arg_uploads()
    t_b = _time_steps(b.step, lambda: made[0][1].energies(), args.steps)
    uploads = api.arg_uploads() - up0
    n_prof = max(1, min(args.steps, 20))
    b.profile(True)
    b.step(n_prof)
    prof = b.profile_read()
    b.close()
    dh = [e.energies()["dh_dl"] for e, _ in made]
    temps = [m.energies()["temperature"] for _, m in made]
    one = args.steps * 0.1728 / t_one
    agg = n * args.steps * 0.1728 / t_b
    out = {
        "metric": "aggregate ns/day of batched λ-windows, 1 GPU", "value": agg, "unit": "ns/day",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_b / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} alchemical windows of chain2600_opc8600_box64 (37,000 sites "
                               "each), 40 decoupled solute atoms, λ = k/(n-1), soft-core α 0.5, "
                               "Langevin-middle dt 2 fs, SPME 64^3 order 4 twice per window "
                               f"(the cross-term scaling), list buffer {args.list_buffer:.1f} A, "
                               "thermalised start",
                   "windows": n, "sites_per_window": s.n_sites},
        "single_window_ns_per_day": one,
        "batched_over_single": agg / one,
        "dh_dl_range": [min(dh), max(dh)],
        "temperature_K_range": [min(temps), max(temps)],
        "arg_uploads_in_timed_steps": uploads,
        "kernels_ms_per_step": {k: v_[0] / n_prof for k, v_ in prof.items()
                                if not k.startswith("host:")},
        "host_ms_per_step": {k[5:]: v_[0] / n_prof for k, v_ in prof.items()
                             if k.startswith("host:")},
    }
    for e, m in made:
        m.close()
        e.close()
    print(json.dumps(out))
    return 0


def uneven_systems(n, seed=4):
    """The batched uneven-systems workload (systems.config4)."""
    from dynamics_amd import systems
    return systems.config4(n, seed)


def run_batch_uneven(args) -> int:
    """128 independent solvated boxes (edges 24–40 Å, 20–60-atom solutes, PME
    grids at ~1 Å spacing), Langevin-middle MD: one batch step advances
    every box.  Compared with the same boxes stepped one after the other."""
    import numpy as np
    from dynamics_amd import api, systems

    n = args.systems or 128
    S = uneven_systems(n)
    print(f"bench: built {n} systems", file=sys.stderr, flush=True)

    def make(s, gamma, temp, seed):
        ff = systems.chain_force_field(s)
        e = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=0.3123, grid=api.pme_grid_n(s.box, 1.0),
                       order=4, mode=api.MODE_REFERENCE, list_buffer=args.list_buffer,
                       item_groups=args.item_groups)
        m = api.Md(e, ff, dt=0.002, temperature=temp, gamma=gamma, shake_tol=1e-5, seed=seed)
        return e, m, ff

    # relax every box (batched quench), then MD from the relaxed states
    made = [make(s, 1.0e4, 0.0, 0) for s in S]
    for (e, m, ff), s in zip(made, S):
        m.set_state(s.pos_dyn, np.zeros((s.n_dyn, 3), np.float32), systems.water_positions(s),
                    np.zeros((s.n_water, 3, 3), np.float32))
    b = api.Batch([m for _, m, _ in made])
    b.step(args.relax_steps)
    b.close()
    starts = [m.state() for _, m, _ in made]
    for e, m, _ in made:
        m.close()
        e.close()
    # thermalise every box (batched Langevin, γ = 10/ps, 300 K), untimed
    made = [make(s, 10.0, 300.0, 300 + k) for k, s in enumerate(S)]
    for (e, m, ff), s, st in zip(made, S, starts):
        v, vw = systems.initial_velocities(s, ff, 300.0)
        m.set_state(st[0], v, st[2], vw)
    b = api.Batch([m for _, m, _ in made])
    b.step(args.equil_steps)
    b.close()
    starts = [m.state() for _, m, _ in made]
    for e, m, _ in made:
        m.close()
        e.close()
    print("bench: relaxed and equilibrated", file=sys.stderr, flush=True)
    made = [make(s, 1.0, 300.0, 200 + k) for k, s in enumerate(S)]
    vel = [(st[1], st[3]) for st in starts]

    def reset():
        for (e, m, ff), st, (v, vw) in zip(made, starts, vel):
            m.set_state(st[0], v, st[2], vw)

    # the same boxes one after the other (plain single-system path)
    reset()
    for _, m, _ in made:
        m.step(args.warmup)
    t_seq = 0.0
    for _, m, _ in made:
        t_seq += _time_steps(m.step, lambda: m.energies(), args.steps)
    reset()
    b = api.Batch([m for _, m, _ in made])
    b.step(args.warmup)
    print("bench: warmup done", file=sys.stderr, flush=True)
    up0 = api.arg_uploads()
    t_b = _time_steps(b.step, lambda: made[0][1].energies(), args.steps)
    uploads = api.arg_uploads() - up0
    n_prof = max(1, min(args.steps, 20))
    b.profile(True)
    b.step(n_prof)
    prof = b.profile_read()
    b.close()
    temps = [m.energies()["temperature"] for _, m, _ in made]
    sites = [s.n_sites for s in S]
    seq = n * args.steps * 0.1728 / t_seq
    agg = n * args.steps * 0.1728 / t_b
    out = {
        "metric": "aggregate ns/day of batched uneven systems, 1 GPU", "value": agg,
        "unit": "ns/day", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_b / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{n} solvated boxes, edges U(24, 40) A, 20-60-atom chains, "
                               "OPC water, rc 10 A, SPME ~1 A grids order 4, Langevin-middle "
                               f"dt 2 fs, list buffer {args.list_buffer:.1f} A, thermalised start",
                   "systems": n, "sites_total": int(sum(sites)),
                   "sites_range": [int(min(sites)), int(max(sites))]},
        "sequential_ns_per_day": seq,
        "batched_over_sequential": agg / seq,
        "temperature_K_range": [min(temps), max(temps)],
        "arg_uploads_in_timed_steps": uploads,
        "kernels_ms_per_step": {k: v_[0] / n_prof for k, v_ in prof.items()
                                if not k.startswith("host:")},
        "host_ms_per_step": {k[5:]: v_[0] / n_prof for k, v_ in prof.items()
                             if k.startswith("host:")},
    }
    for e, m, _ in made:
        m.close()
        e.close()
    print(json.dumps(out))
    return 0


def run_precision(args) -> int:
    """The 128 Å water box (69,900 OPC waters, 279,600 sites; rc 9 Å, SPME 128³
    order 4): one nonbonded evaluation per step with the pair kernel and the
    SPME each in FP32 or FP64; force and energy errors against the engine's
    all-FP64 evaluation of the same Hamiltonian (held to the FP64 oracle by
    the tests), and the time per evaluation."""
    import numpy as np
    from dynamics_amd import api, systems

    s = systems.config2()
    K = (128, 128, 128)
    beta = float(np.float32(0.3123))
    both = api.PREC_PAIR_F64 | api.PREC_PME_F64
    e64 = api.Engine(s, cutoff=9.0, lj_cutoff=9.0, beta=beta, grid=K, order=4,
                     mode=api.MODE_REFERENCE, list_buffer=0.0, precision=both)
    e64.compute(api.COMPUTE_ALL)
    F64 = e64.site_forces()
    E64 = e64.energies()
    e64.close()
    print("bench: FP64 evaluation done", file=sys.stderr, flush=True)
    nrm = float(np.linalg.norm(F64))
    scale = abs(E64["lj"]) + abs(E64["coul_sr"]) + abs(E64["recip"]) + abs(E64["self"])
    rows = []
    for prec, label in ((0, "pair f32, spme f32"), (api.PREC_PAIR_F64, "pair f64, spme f32"),
                        (api.PREC_PME_F64, "pair f32, spme f64"), (both, "pair f64, spme f64")):
        eng = api.Engine(s, cutoff=9.0, lj_cutoff=9.0, beta=beta, grid=K, order=4,
                         mode=api.MODE_REFERENCE, list_buffer=1.0, precision=prec)
        eng.set_positions(s)
        eng.compute(api.COMPUTE_ALL | api.COMPUTE_REBUILD)
        F = eng.site_forces()
        E = eng.energies()
        err = float(np.linalg.norm(F - F64) / nrm)
        e_err = abs(E["total"] - E64["total"]) / scale
        for _ in range(args.warmup):
            eng.compute(api.COMPUTE_ALL | api.COMPUTE_ASYNC)
        eng.profile(True)
        eng.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng.compute(api.COMPUTE_ALL | api.COMPUTE_ASYNC)
        eng.sync()
        t = (time.perf_counter() - t0) / args.steps
        prof = eng.profile_read()
        eng.close()
        rows.append({"precision": label, "force_rel_rms_vs_engine_fp64": err,
                     "energy_rel_err": e_err,
                     "ms_per_eval": t * 1e3,
                     "pairs_ms": prof.get("pairs", (0.0, 1))[0] / max(prof.get("pairs", (0, 1))[1], 1),
                     "spme_ms": sum(v[0] for k, v in prof.items()
                                    if k.startswith("pme_")) / args.steps})
        print(f"bench: {label}: {err:.2e}", file=sys.stderr, flush=True)
    out = {
        "metric": "force/energy error vs FP64 and time per evaluation by precision, 1 GPU",
        "value": rows[0]["force_rel_rms_vs_engine_fp64"],
        "unit": "relative RMS force error (FP32 path)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": rows[0]["ms_per_eval"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32/f64", "data": "synthetic",
        "config": {"workload": "opc69900_box128: 69,900 OPC waters (279,600 sites), rc 9 A, "
                               "SPME 128^3 order 4, reference mode, one evaluation per step "
                               "with a 1 A list buffer"},
        "precision_sweep": rows,
    }
    print(json.dumps(out))
    return 0


def run_tune(args) -> int:
    """The autotuner (dynamics_amd.tune) on the MD workload from the same
    thermal start as --mode md: SPME (order, grid) candidates against a
    converged FP64 evaluation and their SPME time, then MD ms/step over list
    buffers × work-item lengths.  Prints the choice and the table."""
    from dynamics_amd import api, systems, tune

    s = systems.config1()
    K = (64, 64, 64)
    eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=0.3123, grid=K, order=4,
                     mode=api.MODE_REFERENCE, list_buffer=MD_BUFFER)
    ff = systems.chain_force_field(s)
    pd0, pw0 = api.relax(eng, ff, s.pos_dyn, systems.water_positions(s), steps=args.relax_steps)
    s = systems.with_positions(s, pd0, pw0)
    st = equilibrate(eng, ff, s, pd0, pw0, args.equil_steps)
    eng.close()
    print("bench: relaxed and equilibrated", file=sys.stderr, flush=True)
    sys_end = systems.with_positions(s, st[0], st[2])
    res = tune.tune_md(sys_end, ff, st, grid=K, order=4, steps=max(args.steps // 5, 20))
    out = {"metric": "autotuner choice for the MD workload, 1 GPU",
           "value": res["choice"]["ms_per_step"], "unit": "ms/step",
           "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
           "config": {"workload": "chain2600_opc8600_box64 MD (Langevin-middle, dt 2 fs), "
                                  "thermalised start"},
           "choice": res["choice"], "table": res["table"]}
    print(json.dumps(out))
    return 0


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


if __name__ == "__main__":
    sys.exit(main())
