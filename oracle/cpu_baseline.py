"""Timed CPU baseline (test infrastructure only): ctypes driver of
oracle/cpu/cpu_ref.cpp, the C++ restatement of the reference's CPU
nonbonded path (see that file for the per-component citations)."""

from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libcpuref.so")


class _Input(C.Structure):
    _fields_ = [("n_std", C.c_uint32), ("n_wat", C.c_uint32),
                ("pos_std", C.c_void_p), ("q_std", C.c_void_p), ("sig_std", C.c_void_p),
                ("eps_std", C.c_void_p), ("pos_w", C.c_void_p),
                ("n_excl", C.c_uint32), ("n_14", C.c_uint32),
                ("excl", C.c_void_p), ("p14", C.c_void_p),
                ("ext", C.c_float * 3), ("cutoff", C.c_float), ("lj_cutoff", C.c_float),
                ("skin", C.c_float), ("alpha", C.c_float), ("K", C.c_uint32 * 3),
                ("order", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("us_neighbors", C.c_double), ("us_setup_pairs", C.c_double),
                ("us_short", C.c_double), ("us_spme", C.c_double),
                ("energy_short", C.c_double), ("virial_short", C.c_double),
                ("energy_recip", C.c_double), ("virial_recip", C.c_double),
                ("n_list_pairs", C.c_uint64), ("n_interacting", C.c_uint64),
                ("steps", C.c_int)]


def _lib():
    if not os.path.exists(_LIB):
        import subprocess
        subprocess.run(["make", "-s", "-C", os.path.join(os.path.dirname(_LIB), "..", "cpu")],
                       check=True)
    L = C.CDLL(_LIB)
    L.cpuref_run.argtypes = [C.POINTER(_Input), C.c_int, C.c_int, C.POINTER(Stats), C.c_void_p]
    L.cpuref_run.restype = C.c_int
    L.cpuref_run_traj.argtypes = [C.POINTER(_Input), C.c_int, C.c_void_p, C.c_int, C.c_int,
                                  C.POINTER(Stats), C.c_void_p, C.POINTER(C.c_int)]
    L.cpuref_run_traj.restype = C.c_int
    return L


def host_threads():
    """The host cores this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def _input(system, cutoff, lj_cutoff, skin, alpha, K, order):
    s = system
    keep = [np.ascontiguousarray(a, dt) for a, dt in [
        (s.pos_dyn, np.float32), (s.q_dyn, np.float32), (s.sigma_dyn, np.float32),
        (s.eps_dyn, np.float32),
        (np.stack([s.pos_o, s.pos_m, s.pos_h0, s.pos_h1], axis=1), np.float32),
        (s.excl_12_13, np.uint32), (s.pairs_14, np.uint32)]]
    ptr = lambda a: a.ctypes.data_as(C.c_void_p) if a.size else None
    inp = _Input(s.n_dyn, s.n_water, ptr(keep[0]), ptr(keep[1]), ptr(keep[2]), ptr(keep[3]),
                 ptr(keep[4]), len(s.excl_12_13), len(s.pairs_14), ptr(keep[5]), ptr(keep[6]),
                 (C.c_float * 3)(*[float(v) for v in s.box]), cutoff, lj_cutoff, skin, alpha,
                 (C.c_uint32 * 3)(*K), order)
    return inp, keep


def run(system, steps=1, threads=None, cutoff=10.0, lj_cutoff=10.0, skin=4.0, alpha=0.3123,
        K=(64, 64, 64), order=4, want_forces=False):
    """Runs `steps` evaluations; returns (Stats, forces (n_sites,3) or None).
    Sites are ordered solute, then per water O, M, H0, H1."""
    L = _lib()
    threads = threads or host_threads()
    inp, keep = _input(system, cutoff, lj_cutoff, skin, alpha, K, order)
    s = system
    st = Stats()
    F = np.zeros((s.n_sites, 3)) if want_forces else None
    rc = L.cpuref_run(C.byref(inp), steps, threads, C.byref(st),
                      F.ctypes.data_as(C.c_void_p) if want_forces else None)
    if rc != 0:
        raise RuntimeError("cpuref_run failed")
    return st, F


def run_trajectory(system, frames, spme_ratio=2, threads=None, cutoff=10.0, lj_cutoff=10.0,
                   skin=4.0, alpha=0.3123, K=(64, 64, 64), order=4):
    """The reference CPU path over a trajectory at its own cadence: frames
    (n_frames, n_sites, 3) in site order; neighbor lists and the pair list
    rebuilt when a solute atom or water O has moved half the skin, SPME every
    spme_ratio-th step (src/neighbors.rs:56-83, :159-169; src/lib.rs:
    1282-1341).  Returns a dict of wall time, steps/s, rebuilds and the
    per-component times (ms per step)."""
    L = _lib()
    threads = threads or host_threads()
    inp, keep = _input(system, cutoff, lj_cutoff, skin, alpha, K, order)
    fr = np.ascontiguousarray(frames, np.float32)
    n = fr.shape[0]
    st = Stats()
    rb = C.c_int()
    t0 = time.perf_counter()
    rc = L.cpuref_run_traj(C.byref(inp), n, fr.ctypes.data_as(C.c_void_p), int(spme_ratio),
                           threads, C.byref(st), None, C.byref(rb))
    wall = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError("cpuref_run_traj failed")
    sps = n / wall
    return {
        "steps": n, "wall_s": wall, "threads": threads, "steps_per_s": sps,
        "ns_per_day": sps * 0.1728, "rebuilds": rb.value,
        "pair_interactions_per_s": st.n_interacting * sps,
        "ms_neighbors": st.us_neighbors / n / 1e3, "ms_setup_pairs": st.us_setup_pairs / n / 1e3,
        "ms_short": st.us_short / n / 1e3, "ms_spme": st.us_spme / n / 1e3,
    }


def timed(system, seconds=12.0, threads=None, **kw):
    """Evaluations repeated until about `seconds` of wall time: returns a dict
    with steps/s, ns/day at 2 fs, pair-interactions/s and the component
    breakdown (ms per evaluation)."""
    threads = threads or host_threads()
    t0 = time.perf_counter()
    st, _ = run(system, 1, threads, **kw)
    one = time.perf_counter() - t0
    n = max(1, min(500, int(seconds / max(one, 1e-3))))
    t0 = time.perf_counter()
    st, _ = run(system, n, threads, **kw)
    wall = time.perf_counter() - t0
    sps = n / wall
    return {
        "steps": n, "wall_s": wall, "threads": threads, "steps_per_s": sps,
        "ns_per_day": sps * 0.1728, "pair_interactions_per_s": st.n_interacting * sps,
        "ms_neighbors": st.us_neighbors / n / 1e3, "ms_setup_pairs": st.us_setup_pairs / n / 1e3,
        "ms_short": st.us_short / n / 1e3, "ms_spme": st.us_spme / n / 1e3,
        "n_list_pairs": int(st.n_list_pairs), "n_interacting": int(st.n_interacting),
    }
