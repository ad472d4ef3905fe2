"""Shared test setup: the `gpu` marker, in-tree builds when missing, and
ctypes access to the oracle's host replay library."""

import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

LIB = os.path.join(ROOT, "dynamics_amd", "libdyn_nb.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
CPUREF_LIB = os.path.join(ROOT, "oracle", "_build", "libcpuref.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    # A fresh checkout has no built libraries (they are git-ignored); build them
    # here so the CPU suite can check the ABI and the host replays.
    if not os.path.exists(LIB):
        from dynamics_amd import _build
        _build.build(verbose=False)
    if not (os.path.exists(ORACLE_LIB) and os.path.exists(CPUREF_LIB)):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle", "cpu")], check=True)


def has_gpu() -> bool:
    from dynamics_amd import api
    return api.device_count() > 0


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu():
        pytest.fail("no GPU visible: gpu-marked tests need an MI355X")
    return 0


_orc = None


def oracle_lib():
    global _orc
    if _orc is None:
        L = C.CDLL(ORACLE_LIB)
        f = C.c_float
        for name in ["orc_rsqrtf", "orc_exp_neg", "orc_erfcx", "orc_erfcf", "orc_inv_cbrt"]:
            getattr(L, name).restype = f
            getattr(L, name).argtypes = [f]
        L.orc_fx.restype = C.c_int64
        L.orc_fx.argtypes = [f]
        for name in ["orc_coul_erfc", "orc_coul_erfc_tab", "orc_coul_erf"]:
            getattr(L, name).argtypes = [f, f, f, C.POINTER(f), C.POINTER(f)]
        L.orc_lj.argtypes = [f, f, f, C.POINTER(f), C.POINTER(f)]
        L.orc_soft_core.argtypes = [f, f, f, f, C.POINTER(f), C.POINTER(f), C.POINTER(f)]
        L.orc_bspline4.argtypes = [f, C.c_void_p, C.c_void_p]
        L.orc_bspline5.argtypes = [f, C.c_void_p, C.c_void_p]
        L.orc_min_image_ref.restype = f
        L.orc_min_image_ref.argtypes = [f, f]
        L.orc_norm2_ref.restype = f
        L.orc_norm2_ref.argtypes = [f, f, f]
        _orc = L
    return _orc


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def replay_ref_pairs(system, pairs, cutoff, alpha, lj_cutoff=np.inf, lambda_alch=0.0,
                     coulomb_disabled=False, lj_disabled=False, energy_all=False):
    """Host replay of the reference pair-list kernel: (fx (n_sites,3) int64,
    energy, virial, dhdl), sites class-major."""
    L = oracle_lib()
    nd, nw = system.n_dyn, system.n_water
    pos = np.ascontiguousarray(np.concatenate(
        [system.pos_dyn, system.pos_o, system.pos_m, system.pos_h0, system.pos_h1]).astype(np.float32))
    fx = np.zeros((nd + 4 * nw, 3), np.int64)
    e, w, d = C.c_double(), C.c_double(), C.c_double()
    arr = {k: np.ascontiguousarray(v) for k, v in pairs.items() if v is not None}
    n = len(arr["tgt"])
    L.orc_ref_pairs.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32] + [C.c_void_p] * 15 + [
        C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_double, C.c_int,
        C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    ext = np.ascontiguousarray(system.box, np.float32)
    L.orc_ref_pairs(ptr(pos), nd, nw, n, *[ptr(arr.get(k)) for k in [
        "tgt", "src", "sigma", "eps", "q_tgt", "q_src", "atom_type_tgt", "water_site_tgt",
        "atom_type_src", "water_site_src", "scale14", "calc_lj", "calc_coul", "symmetric",
        "alch"]], ptr(ext), cutoff, alpha, lj_cutoff, int(coulomb_disabled), int(lj_disabled),
        float(lambda_alch), int(energy_all), ptr(fx), C.byref(e), C.byref(w), C.byref(d))
    return fx, e.value, w.value, d.value


def engine_lj_table(system):
    L = oracle_lib()
    nt = system.lj_types.shape[0]
    tab = np.zeros(((nt + 2) * (nt + 2), 2), np.float32)
    L.orc_engine_lj_table.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    se = np.ascontiguousarray(system.lj_types, np.float32)
    L.orc_engine_lj_table(ptr(se) if nt else None, nt, ptr(tab))
    return tab


def engine_lj_se(system):
    """(σ, ε) per type pair as the engine combines them for soft-core:
    Lorentz–Berthelot in f32 over the solute types, water O and "none"."""
    lt = np.asarray(system.lj_types, np.float32).reshape(-1, 2)
    f = np.float32
    o_sigma = f(1.777167268) * (f(2.0) / f(1.122462048309373))  # as the engine evaluates it
    sig = np.concatenate([lt[:, 0], [o_sigma], [0.0]]).astype(np.float32)
    eps = np.concatenate([lt[:, 1], [np.float32(0.2128008130)], [0.0]]).astype(np.float32)
    s_ab = (np.float32(0.5) * (sig[:, None] + sig[None, :])).astype(np.float32)
    e_ab = np.sqrt((eps[:, None] * eps[None, :]).astype(np.float32)).astype(np.float32)
    return np.ascontiguousarray(np.stack([s_ab, e_ab], -1).reshape(-1, 2), np.float32)


def replay_engine_tiles(dbg, system, cutoff, lj_cutoff, beta, alch=False, lam=0.0,
                        energies=False, per_item=False):
    """Host replay of the pair tiles over an engine list: (3, n_slots) int64
    fixed-point slot forces, and with energies=True also the six sums
    (e_lj water, e_lj solute, e_coul water, e_coul solute, virial, dH/dλ),
    every pair counted from both of its rows (per_item: (n_items, 6) too)."""
    L = oracle_lib()
    tab = engine_lj_table(system)
    tse = engine_lj_se(system)
    nt = system.lj_types.shape[0] + 2
    o = system.lj_types.shape[0]
    c6_oo, c12_oo = tab[o * nt + o]
    ns = len(dbg["slot_site"])
    fx = np.zeros((3, ns), np.int64)
    L.orc_engine_tiles.argtypes = [C.c_void_p] * 8 + [C.c_uint64, C.c_uint32, C.c_void_p,
                                                      C.c_void_p, C.c_int, C.c_void_p, C.c_float,
                                                      C.c_float, C.c_float, C.c_float, C.c_float,
                                                      C.c_int, C.c_float, C.c_void_p, C.c_void_p,
                                                      C.c_void_p]
    e = np.zeros(6, np.float64)
    ie = np.zeros((len(dbg["items"]), 6), np.float64)
    box = np.ascontiguousarray(system.box, np.float32)
    cm = np.ascontiguousarray(dbg["cmask"].astype(np.uint64))
    items = np.ascontiguousarray(dbg["items"], np.uint32)
    L.orc_engine_tiles(ptr(dbg["xyzq"]), ptr(dbg["slot_type"]), ptr(dbg["slot_gid"]),
                       ptr(dbg["cl_class"]), ptr(dbg["row_ptr"]),
                       ptr(dbg["cj"]) if len(dbg["cj"]) else None,
                       ptr(cm) if len(cm) else None, ptr(items) if len(items) else None,
                       len(items), ns, ptr(tab), ptr(tse), nt, ptr(box), cutoff, lj_cutoff, beta,
                       float(c6_oo), float(c12_oo), 1 if alch else 0, float(lam), ptr(fx), ptr(e),
                       ptr(ie) if per_item and len(ie) else None)
    if per_item:
        return fx, e, ie
    return (fx, e) if energies else fx


def replay_pme_spread(xyzq, K, box, order):
    L = oracle_lib()
    grid = np.zeros(tuple(K), np.int64)
    L.orc_pme_spread.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    xyzq = np.ascontiguousarray(xyzq, np.float32)
    Ka = np.ascontiguousarray(K, np.int32)
    b = np.ascontiguousarray(box, np.float32)
    L.orc_pme_spread(ptr(xyzq), xyzq.shape[0], ptr(Ka), ptr(b), order, ptr(grid))
    return grid


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def exclusion_csr(system):
    """Per solute site, the sorted partners of its excluded and 1-4 pairs
    (both directions), as the engine's list build uses them."""
    nd = system.n_dyn
    sets = [set() for _ in range(nd)]
    for arr in (system.excl_12_13, system.pairs_14):
        for i, j in np.asarray(arr, np.int64).reshape(-1, 2):
            sets[i].add(int(j))
            sets[j].add(int(i))
    ptr = np.zeros(nd + 1, np.uint32)
    idx = []
    for i in range(nd):
        v = sorted(sets[i])
        idx.extend(v)
        ptr[i + 1] = ptr[i] + len(v)
    return ptr, np.asarray(idx if idx else [0], np.uint32)


NULL_ENTRY = (1 << 30) - 1  # j-slot field of a padding entry in the host replay


def replay_engine_list(system, cutoff, list_buffer=0.0, cap_cl=None, cap_groups=None):
    """Host restatement of the engine's pair list: returns (n_clusters,
    slot_site, cl_class, row_ptr (groups), cj (8 per group: slot | class << 30,
    padding = NULL_ENTRY | class << 30), cmask (per group))."""
    L = oracle_lib()
    pos, _, _, _, _, _ = system.site_arrays()
    ns = pos.shape[0]
    pos4 = np.zeros((ns, 4), np.float32)
    pos4[:, :3] = pos
    box = np.ascontiguousarray(system.box, np.float32)
    ptr_, idx = exclusion_csr(system)
    cap_cl = cap_cl or (ns // 8 + 3 * 64 * 64 + 1)
    cap_groups = cap_groups or cap_cl * 256
    slot_site = np.zeros(8 * cap_cl, np.int32)
    cl_class = np.zeros(cap_cl, np.uint8)
    row_ptr = np.zeros(cap_cl + 1, np.uint32)
    cj = np.zeros(8 * cap_groups, np.uint32)
    cm = np.zeros(cap_groups, np.uint64)
    ncl = C.c_uint32()
    ne = C.c_uint64()
    L.orc_engine_list.restype = C.c_int
    L.orc_engine_list.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_float,
                                  C.c_float, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]
    rc = L.orc_engine_list(system.n_dyn, system.n_water, ptr(pos4), ptr(box), cutoff, list_buffer,
                           ptr(ptr_), ptr(idx), cap_cl, cap_groups, C.byref(ncl), ptr(slot_site),
                           ptr(cl_class), ptr(row_ptr), ptr(cj), ptr(cm), C.byref(ne))
    assert rc == 0, "replay capacity exceeded"
    n, m = ncl.value, ne.value
    return n, slot_site[:8 * n], cl_class[:n], row_ptr[:n + 1], cj[:8 * m], cm[:m]


def normalize_null_entries(cj, n_slots):
    """GPU lists name their own null slot index; map it to NULL_ENTRY."""
    cj = np.asarray(cj, np.uint32).copy()
    slot = cj & np.uint32(NULL_ENTRY)
    cj[slot >= n_slots] = (cj[slot >= n_slots] & ~np.uint32(NULL_ENTRY)) | np.uint32(NULL_ENTRY)
    return cj


def host_engine_debug(system, cutoff, alch=None, chunk=16):
    """The engine's list and slot arrays as the host restatement builds them,
    in the layout of Engine.debug_export (slot positions wrapped with
    np.mod, which the minimum image makes equivalent)."""
    n, slot_site, cl_class, row_ptr, cj, cm = replay_engine_list(system, cutoff)
    pos, q, _, _, mol, _ = system.site_arrays()
    box = np.asarray(system.box, np.float32)
    w = np.mod(pos.astype(np.float32), box).astype(np.float32)
    nt = system.lj_types.shape[0]
    tw = np.empty(len(q), np.int32)
    tw[:system.n_dyn] = system.type_dyn
    if alch is not None:
        tw[:system.n_dyn] |= np.where(alch, 0x10000, 0).astype(np.int32)
    tw[system.n_dyn:] = np.tile(np.array([nt, nt + 1, nt + 1, nt + 1], np.int32), system.n_water)
    gid = np.where(mol >= 0, mol, -2 - np.arange(len(q))).astype(np.int32)
    ns = 8 * n
    xyzq = np.zeros((ns, 4), np.float32)
    st = np.full(ns, nt + 1, np.int32)
    sg = np.full(ns, -1, np.int32)
    ok = slot_site >= 0
    xyzq[ok, :3] = w[slot_site[ok]]
    xyzq[ok, 3] = q[slot_site[ok]]
    first = slot_site[(np.arange(ns) // 8) * 8]
    xyzq[~ok, :3] = w[first[~ok]]
    st[ok] = tw[slot_site[ok]]
    sg[ok] = gid[slot_site[ok]]
    items = [(I, g0) for I in range(n) for g0 in range(row_ptr[I], row_ptr[I + 1], chunk)]
    items = np.asarray(items, np.uint32).reshape(-1, 2)
    return {"slot_site": slot_site, "xyzq": xyzq, "slot_type": st, "slot_gid": sg,
            "cl_class": cl_class, "row_ptr": row_ptr, "cj": cj, "cmask": cm, "items": items}


def slot_forces_to_sites(fx, dbg, n_sites):
    """(3, n_slots) fixed point → (n_sites, 3) float64 forces."""
    F = np.zeros((n_sites, 3))
    ok = dbg["slot_site"] >= 0
    F[dbg["slot_site"][ok]] = fx[:, ok].T * 2.0 ** -32
    return F
