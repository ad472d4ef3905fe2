"""ctypes binding of libdyn_nb.so (include/dyn_nb.h) with thin Python classes
named after the reference objects they stand in for.  Arrays in and out are
numpy; everything numeric happens in the library.  There is no CPU fallback:
if the library is missing, importing this module raises.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdyn_nb.so")

DYN_OK = 0
DYN_ERR_ARG = 1
DYN_ERR_HIP = 2
DYN_ERR_STATE = 3
DYN_ERR_CAPACITY = 4
DYN_ERR_NOMEM = 5
DYN_ERR_LIMIT = 6

MODE_REFERENCE = 0
MODE_STANDARD = 1

COMPUTE_SHORT = 1
COMPUTE_RECIP = 2
COMPUTE_CORR = 4
COMPUTE_ALL = 7
COMPUTE_REBUILD = 8
COMPUTE_ASYNC = 16
COMPUTE_NO_ENERGY = 32
COMPUTE_RECIP_REUSE = 64
PREC_PAIR_F64 = 1
PREC_PME_F64 = 2


class DynError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                      "(hipcc for gfx950); there is no CPU fallback")

lib = C.CDLL(LIB_PATH)

_p = C.c_void_p
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class NbOpts(C.Structure):
    _fields_ = [("lj_cutoff", C.c_float), ("energy_all_pairs", C.c_uint32)]


class MdParams(C.Structure):
    _fields_ = [("dt", C.c_float), ("temperature", C.c_float), ("gamma", C.c_float),
                ("shake_tol", C.c_float), ("seed", C.c_uint64), ("spme_ratio", C.c_uint32),
                ("nstpcouple", C.c_uint32), ("pressure_target", C.c_float), ("tau_p", C.c_float),
                ("compressibility", C.c_double), ("direct_launch", C.c_uint32),
                ("staged_builds", C.c_uint32)]


class MdEnergies(C.Structure):
    _fields_ = [("kinetic", C.c_double), ("kinetic_rigid", C.c_double), ("bond", C.c_double),
                ("angle", C.c_double), ("dihedral", C.c_double), ("temperature", C.c_double),
                ("dof", C.c_double), ("steps", C.c_uint64), ("pressure", C.c_double),
                ("virial", C.c_double), ("kinetic_com", C.c_double), ("volume", C.c_double),
                ("box", C.c_double * 3), ("pcouple_steps", C.c_uint64),
                ("potential_nonbonded", C.c_double), ("potential_bonded", C.c_double),
                ("potential", C.c_double), ("dh_dl", C.c_double), ("density", C.c_double),
                ("repeats", C.c_uint64), ("graph_steps", C.c_uint64)]


class EngineParams(C.Structure):
    _fields_ = [("cutoff", C.c_float), ("lj_cutoff", C.c_float), ("beta", C.c_float),
                ("grid", C.c_uint32 * 3), ("order", C.c_uint32), ("mode", C.c_uint32),
                ("list_buffer", C.c_float), ("precision", C.c_uint32),
                ("mesh_spacing", C.c_float), ("item_groups", C.c_uint32),
                ("single_stream", C.c_uint32)]


class Energies(C.Structure):
    _fields_ = [("lj", C.c_double), ("coul_sr", C.c_double), ("recip", C.c_double),
                ("self", C.c_double), ("excl_corr", C.c_double), ("total", C.c_double),
                ("reference", C.c_double), ("virial_sr", C.c_double),
                ("virial_recip", C.c_double), ("dh_dl", C.c_double),
                ("pair_count", C.c_uint64), ("list_entries", C.c_uint64),
                ("list_builds", C.c_uint64), ("pair_count_kind", C.c_uint64 * 3)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["pair_count_kind"] = tuple(int(v) for v in self.pair_count_kind)
        return d


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("dyn_version", C.c_int, [])
_sig("dyn_device_count", C.c_int, [])
_sig("dyn_global_last_error", C.c_char_p, [])
_sig("dyn_nb_create", C.c_int, [C.POINTER(_p), C.c_int, C.c_uint32, C.c_uint32, C.c_float,
                                C.c_float, C.POINTER(NbOpts)])
_sig("dyn_nb_destroy", None, [_p])
_sig("dyn_nb_last_error", C.c_char_p, [_p])
_sig("dyn_nb_set_pairs", C.c_int, [_p, C.c_uint32, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                   _p, _p, _p, _p])
_sig("dyn_nb_compute", C.c_int, [_p, _p, _p, _p, _p, _p, _f32p, C.c_uint8, C.c_uint8,
                                 C.c_double, _p, _p, _p, _p, _p])
_sig("dyn_nb_get_fixed_forces", C.c_int, [_p, _i64p, C.c_uint64])
_sig("dyn_nb_correct_14", C.c_int, [_p, C.c_uint32, _p, _p, _p, _p, _f32p, _f64p, _p])
_sig("dyn_nb_correct_14_noalloc", C.c_int, [_p, C.c_uint32, _p, _p, _p, _p, _f32p, _f64p, _p])
_sig("dyn_nb_build_neighbors", C.c_int, [_p, _p, C.c_uint32, _p, C.c_uint32, _p, _f32p,
                                         C.c_uint8, C.c_float, _u32p, _p, C.c_uint64,
                                         C.POINTER(C.c_uint64)])
_sig("dyn_pme_create", C.c_int, [C.POINTER(_p), C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_float, C.c_float, C.c_float, C.c_float, C.c_uint32])
_sig("dyn_pme_destroy", None, [_p])
_sig("dyn_pme_last_error", C.c_char_p, [_p])
_sig("dyn_pme_forces_and_virial", C.c_int, [_p, C.c_uint32, _p, _p, _p, _p, _p])
_sig("dyn_pme_get_fixed_grid", C.c_int, [_p, _i64p, C.c_uint64])
_sig("dyn_pme_grid_n", C.c_int, [C.c_float, C.c_float, C.c_float, C.c_float,
                                 C.POINTER(C.c_uint32)])
_sig("dyn_engine_create", C.c_int, [C.POINTER(_p), C.c_int, C.POINTER(EngineParams), C.c_uint32,
                                    C.c_uint32, _p, _p, C.c_uint32, _p, C.c_uint32, _p,
                                    C.c_uint32, _p])
_sig("dyn_engine_destroy", None, [_p])
_sig("dyn_engine_last_error", C.c_char_p, [_p])
_sig("dyn_engine_set_box", C.c_int, [_p, _f32p])
_sig("dyn_engine_set_positions", C.c_int, [_p, _p, _p, _p, _p, _p])
_sig("dyn_engine_set_positions_device", C.c_int, [_p, C.c_void_p])
_sig("dyn_engine_compute", C.c_int, [_p, C.c_uint32])
_sig("dyn_engine_set_alchemical", C.c_int, [_p, C.c_uint32, C.c_uint32, C.c_double])
_sig("dyn_engine_set_molecules", C.c_int, [_p, C.c_uint32, _p])
_sig("dyn_engine_get_pme_grid", C.c_int, [_p, C.POINTER(C.c_uint32)])
_sig("dyn_engine_get_mol_energies", C.c_int, [_p, _f64p, C.c_uint32])
_sig("dyn_md_create", C.c_int, [C.POINTER(_p), _p, C.POINTER(MdParams), _p,
                                C.c_uint32, _p, _p, C.c_uint32, _p, _p, C.c_uint32, _p, _p,
                                C.c_uint32, _p, _p])
_sig("dyn_md_destroy", None, [_p])
_sig("dyn_md_last_error", C.c_char_p, [_p])
_sig("dyn_md_set_state", C.c_int, [_p, _p, _p, _p, _p])
_sig("dyn_md_get_state", C.c_int, [_p, _p, _p, _p, _p])
_sig("dyn_md_step", C.c_int, [_p, C.c_uint32])
_sig("dyn_md_get_energies", C.c_int, [_p, C.POINTER(MdEnergies)])
_sig("dyn_engine_sync", C.c_int, [_p])
_sig("dyn_engine_get_forces", C.c_int, [_p, _p, _p])
_sig("dyn_engine_get_fixed_forces", C.c_int, [_p, _i64p, C.c_uint64])
_sig("dyn_engine_forces_device", C.c_void_p, [_p])
_sig("dyn_engine_get_energies", C.c_int, [_p, C.POINTER(Energies)])
_sig("dyn_engine_debug_item_partials", C.c_int, [_p, C.c_void_p, C.c_uint64])
_sig("dyn_debug_arg_uploads", C.c_uint64, [])
_sig("dyn_debug_layout_generation", C.c_uint64, [])
_sig("dyn_debug_capture_refusal", C.c_char_p, [])
_sig("dyn_engine_debug_inject_status", C.c_int, [_p, C.c_uint32, C.c_uint32, C.c_uint32])
_sig("dyn_engine_debug_export", C.c_int, [_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                          C.POINTER(C.c_uint64), _p, _p, _p, _p, _p, _p, _p, _p,
                                          _p, C.POINTER(C.c_uint64), _p])
_sig("dyn_batch_create", C.c_int, [C.POINTER(_p), C.c_uint32, C.POINTER(_p)])
_sig("dyn_batch_destroy", None, [_p])
_sig("dyn_batch_last_error", C.c_char_p, [_p])
_sig("dyn_batch_step", C.c_int, [_p, C.c_uint32])
_sig("dyn_batch_profile", C.c_int, [_p, C.c_int])
_sig("dyn_batch_profile_read", C.c_int, [_p, C.c_uint32, C.POINTER(C.c_char_p),
                                         C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint32)])
_sig("dyn_engine_profile", C.c_int, [_p, C.c_int])
_sig("dyn_engine_profile_read", C.c_int, [_p, C.c_uint32, C.POINTER(C.c_char_p),
                                          C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                          C.POINTER(C.c_uint32)])

# Every symbol include/dyn_nb.h declares (checked by the CPU test suite).
EXPORTED = [
    "dyn_version", "dyn_device_count", "dyn_global_last_error",
    "dyn_nb_create", "dyn_nb_destroy", "dyn_nb_last_error", "dyn_nb_set_pairs",
    "dyn_nb_compute", "dyn_nb_get_fixed_forces", "dyn_nb_correct_14", "dyn_nb_correct_14_noalloc",
    "dyn_nb_build_neighbors",
    "dyn_pme_create", "dyn_pme_destroy", "dyn_pme_last_error", "dyn_pme_forces_and_virial",
    "dyn_pme_get_fixed_grid", "dyn_pme_grid_n",
    "dyn_engine_create", "dyn_engine_destroy", "dyn_engine_last_error", "dyn_engine_set_box",
    "dyn_engine_set_positions", "dyn_engine_set_positions_device", "dyn_engine_compute",
    "dyn_engine_set_alchemical", "dyn_engine_set_molecules", "dyn_engine_get_mol_energies",
    "dyn_engine_get_pme_grid",
    "dyn_md_create", "dyn_md_destroy", "dyn_md_last_error", "dyn_md_set_state",
    "dyn_md_get_state", "dyn_md_step", "dyn_md_get_energies",
    "dyn_engine_sync", "dyn_engine_get_forces", "dyn_engine_get_fixed_forces",
    "dyn_engine_forces_device", "dyn_engine_get_energies", "dyn_engine_debug_export",
    "dyn_engine_debug_item_partials", "dyn_debug_arg_uploads", "dyn_debug_layout_generation", "dyn_debug_capture_refusal",
    "dyn_engine_debug_inject_status",
    "dyn_engine_profile",
    "dyn_engine_profile_read",
    "dyn_batch_create", "dyn_batch_destroy", "dyn_batch_last_error", "dyn_batch_step",
    "dyn_batch_profile", "dyn_batch_profile_read",
]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def arg_uploads() -> int:
    """Kernel argument-pack uploads so far in this process."""
    return int(lib.dyn_debug_arg_uploads())


def layout_generation() -> int:
    """The library's device-layout generation (step graphs are recorded and
    replayed only while it holds still)."""
    return int(lib.dyn_debug_layout_generation())


def capture_refusal() -> str:
    """Why the last step-graph capture was refused ("" if none was)."""
    return lib.dyn_debug_capture_refusal().decode(errors="replace")


def device_count() -> int:
    return int(lib.dyn_device_count())


def _check(rc, msg_fn):
    if rc != DYN_OK:
        raise DynError(rc, msg_fn().decode(errors="replace") if msg_fn() else "")


def pme_grid_n(l, spacing):
    out = (C.c_uint32 * 3)()
    rc = lib.dyn_pme_grid_n(float(l[0]), float(l[1]), float(l[2]), float(spacing), out)
    _check(rc, lib.dyn_global_last_error)
    return tuple(int(v) for v in out)


class ForcesPositsGpu:
    """The reference's per-system device state plus its force entry
    (ForcesPositsGpu::new, PerNeighborGpu::new, force_nonbonded_gpu in
    src/gpu_interface.rs)."""

    def __init__(self, n_dyn: int, n_water: int, cutoff_ewald: float, alpha_ewald: float,
                 device: int = 0, lj_cutoff: float = float("inf"), energy_all_pairs: bool = False):
        self.h = C.c_void_p()
        opts = NbOpts(lj_cutoff, 1 if energy_all_pairs else 0)
        rc = lib.dyn_nb_create(C.byref(self.h), device, n_dyn, n_water, cutoff_ewald,
                               alpha_ewald, C.byref(opts))
        _check(rc, lib.dyn_global_last_error)
        self.n_dyn, self.n_water = n_dyn, n_water

    def close(self):
        if self.h:
            lib.dyn_nb_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def _err(self):
        return lib.dyn_nb_last_error(self.h)

    def set_pairs(self, p: dict):
        """p: the PerNeighborGpu arrays (tgt, src, sigma, eps, q_tgt, q_src,
        atom_type_tgt, water_site_tgt, atom_type_src, water_site_src, scale14,
        calc_lj, calc_coul, symmetric, alch)."""
        n = int(len(p["tgt"]))
        keep = {k: _c(p[k], dt) for k, dt in [
            ("tgt", np.uint32), ("src", np.uint32), ("sigma", np.float32), ("eps", np.float32),
            ("q_tgt", np.float32), ("q_src", np.float32), ("atom_type_tgt", np.uint8),
            ("water_site_tgt", np.uint8), ("atom_type_src", np.uint8),
            ("water_site_src", np.uint8), ("scale14", np.uint8), ("calc_lj", np.uint8),
            ("calc_coul", np.uint8), ("symmetric", np.uint8)]}
        alch = _c(p["alch"], np.uint8) if p.get("alch") is not None else None
        rc = lib.dyn_nb_set_pairs(self.h, n, *[_ptr(keep[k]) for k in [
            "tgt", "src", "sigma", "eps", "q_tgt", "q_src", "atom_type_tgt", "water_site_tgt",
            "atom_type_src", "water_site_src", "scale14", "calc_lj", "calc_coul", "symmetric"]],
            _ptr(alch))
        _check(rc, self._err)

    def compute(self, pos_dyn, pos_o, pos_m, pos_h0, pos_h1, extent, coulomb_disabled=False,
                lj_disabled=False, lambda_alch=0.0):
        """force_nonbonded_gpu: returns (f_dyn (n,3) f64, f_water (nw,4,3) f64
        in (o, h0, h1, m) order, virial, energy, dh_dl)."""
        arrs = [_c(a, np.float32) for a in (pos_dyn, pos_o, pos_m, pos_h0, pos_h1)]
        f_dyn = np.zeros((self.n_dyn, 3), np.float64)
        f_w = np.zeros((self.n_water, 4, 3), np.float64)
        vir, en, dh = C.c_double(), C.c_double(), C.c_double()
        rc = lib.dyn_nb_compute(self.h, *[_ptr(a) for a in arrs], _c(extent, np.float32),
                                int(coulomb_disabled), int(lj_disabled), float(lambda_alch),
                                _ptr(f_dyn), _ptr(f_w), C.byref(vir), C.byref(en), C.byref(dh))
        _check(rc, self._err)
        return f_dyn, f_w, vir.value, en.value, dh.value

    def fixed_forces(self):
        n = (self.n_dyn + 4 * self.n_water) * 3
        out = np.zeros(n, np.int64)
        _check(lib.dyn_nb_get_fixed_forces(self.h, out, n), self._err)
        return out.reshape(-1, 3)

    def correct_14(self, i14, j14, pos_dyn, q_dyn, extent, f_dyn, virial=0.0, noalloc=False):
        """The 1-4 correction (src/non_bonded.rs:763-790) added into f_dyn;
        returns virial + its virial.  noalloc: the allocation-free entry."""
        i14 = _c(i14, np.uint32)
        j14 = _c(j14, np.uint32)
        pd = _c(pos_dyn, np.float32)
        qd = _c(q_dyn, np.float32)
        w = C.c_double(virial)
        fn = lib.dyn_nb_correct_14_noalloc if noalloc else lib.dyn_nb_correct_14
        rc = fn(self.h, len(i14), _ptr(i14), _ptr(j14), _ptr(pd), _ptr(qd),
                _c(extent, np.float32), f_dyn, C.byref(w))
        _check(rc, self._err)
        return w.value

    def build_neighbors(self, posits_outer, posits_inner, is_static, extent, symmetric,
                        skin_sq_w_cutoff):
        """build_neighbors (src/neighbors.rs:180-281) → CSR (row_ptr, col)."""
        po = _c(posits_outer, np.float32).reshape(-1, 3)
        pi = _c(posits_inner, np.float32).reshape(-1, 3)
        st = None if is_static is None else _c(is_static, np.uint8)
        n_outer = po.shape[0]
        row_ptr = np.zeros(n_outer + 1, np.uint32)
        cap = max(1024, n_outer * 64)
        total = C.c_uint64()
        for _ in range(2):
            col = np.zeros(cap, np.uint32)
            rc = lib.dyn_nb_build_neighbors(self.h, _ptr(po), n_outer, _ptr(pi), pi.shape[0],
                                            _ptr(st), _c(extent, np.float32), int(symmetric),
                                            float(skin_sq_w_cutoff), row_ptr, _ptr(col), cap,
                                            C.byref(total))
            if rc == DYN_ERR_CAPACITY and total.value > cap:
                cap = int(total.value)
                continue
            _check(rc, self._err)
            return row_ptr, col[: int(total.value)]
        raise DynError(DYN_ERR_CAPACITY, "capacity retry failed")


class PmeRecip:
    """Replacement of ewald::PmeRecip (reciprocal-space SPME)."""

    def __init__(self, n, l, alpha, order=4, device=0):
        self.h = C.c_void_p()
        rc = lib.dyn_pme_create(C.byref(self.h), device, int(n[0]), int(n[1]), int(n[2]),
                                float(l[0]), float(l[1]), float(l[2]), float(alpha), int(order))
        _check(rc, lib.dyn_global_last_error)
        self.n = tuple(int(v) for v in n)

    def close(self):
        if self.h:
            lib.dyn_pme_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def _err(self):
        return lib.dyn_pme_last_error(self.h)

    def forces_and_virial(self, pos, q):
        pos = _c(pos, np.float32).reshape(-1, 3)
        q = _c(q, np.float32)
        f = np.zeros_like(pos)
        e, w = C.c_double(), C.c_double()
        rc = lib.dyn_pme_forces_and_virial(self.h, pos.shape[0], _ptr(pos), _ptr(q), _ptr(f),
                                           C.byref(e), C.byref(w))
        _check(rc, self._err)
        return f, e.value, w.value

    def forces(self, pos, q):
        f, e, _ = self.forces_and_virial(pos, q)
        return f, e

    def fixed_grid(self):
        n = self.n[0] * self.n[1] * self.n[2]
        out = np.zeros(n, np.int64)
        _check(lib.dyn_pme_get_fixed_grid(self.h, out, n), self._err)
        return out.reshape(self.n)


class Engine:
    """Native nonbonded engine over positions + topology."""

    def __init__(self, system, cutoff=10.0, lj_cutoff=None, beta=0.3123, grid=(0, 0, 0),
                 order=4, mode=MODE_REFERENCE, list_buffer=0.0, device=0, precision=0,
                 mesh_spacing=0.0, item_groups=0, single_stream=False):
        """precision: PREC_PAIR_F64 | PREC_PME_F64 bits (0: FP32).  mesh_spacing
        > 0 (with grid 0): the SPME grid follows the box (regen_pme).
        item_groups: j groups per work item of the pair kernel (0: 16).
        single_stream: every kernel on the engine's stream (no second stream)."""
        s = system
        prm = EngineParams(cutoff, cutoff if lj_cutoff is None else lj_cutoff, beta,
                           (C.c_uint32 * 3)(*[int(g) for g in grid]), order, mode, list_buffer,
                           precision, mesh_spacing, item_groups, int(bool(single_stream)))
        self.h = C.c_void_p()
        q = _c(s.q_dyn, np.float32)
        t = _c(s.type_dyn, np.int32)
        lj = _c(s.lj_types, np.float32).reshape(-1)
        ex = _c(s.excl_12_13, np.uint32).reshape(-1)
        p14 = _c(s.pairs_14, np.uint32).reshape(-1)
        self._keep = (q, t, lj, ex, p14)
        rc = lib.dyn_engine_create(C.byref(self.h), device, C.byref(prm), s.n_dyn, s.n_water,
                                   _ptr(q), _ptr(t), s.lj_types.shape[0], _ptr(lj),
                                   s.excl_12_13.shape[0], _ptr(ex), s.pairs_14.shape[0],
                                   _ptr(p14))
        _check(rc, lib.dyn_global_last_error)
        self.n_dyn, self.n_water = s.n_dyn, s.n_water
        self.n_sites = s.n_dyn + 4 * s.n_water
        self.set_box(s.box)
        self.set_positions(s)

    def close(self):
        if self.h:
            lib.dyn_engine_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def _err(self):
        return lib.dyn_engine_last_error(self.h)

    def set_box(self, box):
        _check(lib.dyn_engine_set_box(self.h, _c(box, np.float32)), self._err)

    def set_positions(self, s):
        arrs = [_c(a, np.float32) for a in (s.pos_dyn, s.pos_o, s.pos_m, s.pos_h0, s.pos_h1)]
        _check(lib.dyn_engine_set_positions(self.h, *[_ptr(a) for a in arrs]), self._err)

    def set_positions_device(self, ptr: int):
        _check(lib.dyn_engine_set_positions_device(self.h, C.c_void_p(ptr)), self._err)

    def compute(self, flags=COMPUTE_ALL):
        _check(lib.dyn_engine_compute(self.h, flags), self._err)

    def set_alchemical(self, first, count, lam):
        """Decouple solute atoms [first, first + count) at λ = lam (count 0: off)."""
        _check(lib.dyn_engine_set_alchemical(self.h, int(first), int(count), float(lam)), self._err)

    def pme_grid(self):
        out = (C.c_uint32 * 3)()
        _check(lib.dyn_engine_get_pme_grid(self.h, out), self._err)
        return tuple(int(v) for v in out)

    def set_molecules(self, mol_start):
        """Molecule-pair energy matrix over the solute cut at mol_start (first
        atom of each molecule; empty: off)."""
        ms = _c(mol_start, np.uint32)
        self.n_mol = int(ms.size)
        _check(lib.dyn_engine_set_molecules(self.h, self.n_mol, _ptr(ms) if ms.size else None),
               self._err)

    def mol_energies(self):
        """(n_mol, n_mol) solute molecule-pair energies of the last compute, kcal/mol."""
        n = getattr(self, "n_mol", 0)
        out = np.zeros((max(n, 1), max(n, 1)), np.float64)
        _check(lib.dyn_engine_get_mol_energies(self.h, out, n), self._err)
        return out

    def sync(self):
        _check(lib.dyn_engine_sync(self.h), self._err)

    def forces(self):
        f_dyn = np.zeros((self.n_dyn, 3), np.float64)
        f_w = np.zeros((self.n_water, 4, 3), np.float64)
        _check(lib.dyn_engine_get_forces(self.h, _ptr(f_dyn), _ptr(f_w)), self._err)
        return f_dyn, f_w

    def fixed_forces(self):
        """(3, n_sites) int64, sites solute then per water O, M, H0, H1."""
        n = 3 * self.n_sites
        out = np.zeros(n, np.int64)
        _check(lib.dyn_engine_get_fixed_forces(self.h, out, n), self._err)
        return out.reshape(3, self.n_sites)

    def site_forces(self):
        return self.fixed_forces().T.astype(np.float64) * (1.0 / 4294967296.0)

    def energies(self) -> dict:
        e = Energies()
        _check(lib.dyn_engine_get_energies(self.h, C.byref(e)), self._err)
        return e.as_dict()

    def debug_export(self):
        """Cluster layout, list (row_ptr and cmask per group of 8 j entries, cj
        per entry: slot | class << 30), work items and per-slot fixed-point
        pair forces."""
        ns, nc, ne, ni = C.c_uint32(), C.c_uint32(), C.c_uint64(), C.c_uint64()
        _check(lib.dyn_engine_debug_export(self.h, C.byref(ns), C.byref(nc), C.byref(ne),
                                           *[None] * 9, C.byref(ni), None), self._err)
        ns, nc, ne, ni = ns.value, nc.value, ne.value, ni.value
        d = {"slot_site": np.zeros(ns, np.int32), "xyzq": np.zeros((ns, 4), np.float32),
             "slot_type": np.zeros(ns, np.int32), "slot_gid": np.zeros(ns, np.int32),
             "cl_class": np.zeros(nc, np.uint8), "row_ptr": np.zeros(nc + 1, np.uint32),
             "cj": np.zeros(max(8 * ne, 1), np.uint32), "cmask": np.zeros(max(ne, 1), np.uint64),
             "items": np.zeros((max(ni, 1), 2), np.uint32), "fx_slot": np.zeros((3, ns), np.int64)}
        _check(lib.dyn_engine_debug_export(
            self.h, None, None, None, *[_ptr(d[k]) for k in [
                "slot_site", "xyzq", "slot_type", "slot_gid", "cl_class", "row_ptr", "cj",
                "cmask", "items"]], None, _ptr(d["fx_slot"])), self._err)
        d["cj"] = d["cj"][:8 * ne]
        d["cmask"] = d["cmask"][:ne]
        d["items"] = d["items"][:ni]
        return d

    def debug_inject_status(self, bits, skip=0, count=1):
        """Test hook: after `skip` evaluations, the next `count` get `bits`
        ORed into the device status word (8 = the list's work items outgrew
        their buffer) before the pair kernel, as the device's own checks
        would set them."""
        _check(lib.dyn_engine_debug_inject_status(self.h, int(bits), int(skip), int(count)),
               self._err)

    def debug_item_partials(self, n_items):
        """(n_items, 9) per work item sums of the last compute (see the C header)."""
        out = np.zeros((9, n_items), np.float64)
        _check(lib.dyn_engine_debug_item_partials(self.h, _ptr(out), int(n_items)), self._err)
        return out.T.copy()

    def profile(self, enable=True):
        """True/1: time every stage; 2: the pair kernel only; False/0: off."""
        _check(lib.dyn_engine_profile(self.h, int(enable)), self._err)

    def profile_read(self):
        mx = 64
        names = (C.c_char_p * mx)()
        ms = (C.c_double * mx)()
        cnt = (C.c_uint64 * mx)()
        n = C.c_uint32()
        _check(lib.dyn_engine_profile_read(self.h, mx, names, ms, cnt, C.byref(n)), self._err)
        return {names[i].decode(): (ms[i], int(cnt[i])) for i in range(min(n.value, mx))}


class Md:
    """Device-resident MD around an Engine (dyn_md_*): positions, velocities
    and forces stay on the GPU between steps."""

    def __init__(self, engine, ff, dt=0.002, temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=0,
                 spme_ratio=1, nstpcouple=0, pressure_target=1.0, tau_p=5.0,
                 compressibility=4.5e-5, direct_launch=False, staged_builds=False):
        """nstpcouple > 0 couples the pressure (C-rescale) every nstpcouple
        steps to pressure_target bar with time constant tau_p ps.
        direct_launch: launch every kernel directly (no step graphs).
        staged_builds: list builds run beside the pair kernel and serve from
        the next step on (dyn_md_params.staged_builds)."""
        self.engine = engine
        self.h = C.c_void_p()
        p = MdParams(dt, temperature, gamma, shake_tol, seed, spme_ratio, nstpcouple,
                     pressure_target, tau_p, compressibility, 1 if direct_launch else 0,
                     1 if staged_builds else 0)
        arrs = [_c(ff.mass, np.float32),
                _c(ff.bonds, np.uint32), _c(ff.bond_k_r0, np.float32),
                _c(ff.angles, np.uint32), _c(ff.angle_k_t0, np.float32),
                _c(ff.dihedrals, np.uint32), _c(ff.dihe_k_n_phase, np.float32),
                _c(ff.constraints, np.uint32), _c(ff.constraint_r0, np.float32)]
        self._keep = arrs
        rc = lib.dyn_md_create(C.byref(self.h), engine.h, C.byref(p), _ptr(arrs[0]),
                               len(ff.bonds), _ptr(arrs[1]), _ptr(arrs[2]),
                               len(ff.angles), _ptr(arrs[3]), _ptr(arrs[4]),
                               len(ff.dihedrals), _ptr(arrs[5]), _ptr(arrs[6]),
                               len(ff.constraints), _ptr(arrs[7]), _ptr(arrs[8]))
        _check(rc, lib.dyn_global_last_error)
        self.n_dyn, self.n_water = engine.n_dyn, engine.n_water

    def close(self):
        if self.h:
            lib.dyn_md_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def _err(self):
        return lib.dyn_md_last_error(self.h)

    def set_state(self, pos_dyn, vel_dyn, pos_water, vel_water):
        """pos_water (n_water, 4, 3) in O, M, H0, H1 order; vel_water
        (n_water, 3, 3) in O, H0, H1 order."""
        a = [_c(pos_dyn, np.float32), _c(vel_dyn, np.float32), _c(pos_water, np.float32),
             _c(vel_water, np.float32)]
        _check(lib.dyn_md_set_state(self.h, *[_ptr(x) for x in a]), self._err)

    def state(self):
        pd = np.zeros((self.n_dyn, 3), np.float32)
        vd = np.zeros((self.n_dyn, 3), np.float32)
        pw = np.zeros((self.n_water, 4, 3), np.float32)
        vw = np.zeros((self.n_water, 3, 3), np.float32)
        _check(lib.dyn_md_get_state(self.h, _ptr(pd), _ptr(vd), _ptr(pw), _ptr(vw)), self._err)
        return pd, vd, pw, vw

    def step(self, n=1):
        _check(lib.dyn_md_step(self.h, int(n)), self._err)

    def energies(self):
        e = MdEnergies()
        _check(lib.dyn_md_get_energies(self.h, C.byref(e)), self._err)
        d = {k: getattr(e, k) for k, _ in MdEnergies._fields_}
        d["box"] = tuple(float(v) for v in e.box)
        return d


class Batch:
    """Independent systems stepped together (dyn_batch_*): every kernel of a
    step is one launch for all members.  Members keep their own state;
    Md.step refuses them while the batch lives."""

    def __init__(self, mds):
        self.mds = list(mds)
        self.h = C.c_void_p()
        arr = (_p * len(self.mds))(*[m.h for m in self.mds])
        _check(lib.dyn_batch_create(C.byref(self.h), len(self.mds), arr), lib.dyn_global_last_error)

    def close(self):
        if self.h:
            lib.dyn_batch_destroy(self.h)
            self.h = C.c_void_p()

    __del__ = close

    def _err(self):
        return lib.dyn_batch_last_error(self.h)

    def step(self, n=1):
        _check(lib.dyn_batch_step(self.h, int(n)), self._err)

    def profile(self, enable=True):
        _check(lib.dyn_batch_profile(self.h, int(enable)), self._err)

    def profile_read(self):
        mx = 64
        names = (C.c_char_p * mx)()
        ms = (C.c_double * mx)()
        cnt = (C.c_uint64 * mx)()
        n = C.c_uint32()
        _check(lib.dyn_batch_profile_read(self.h, mx, names, ms, cnt, C.byref(n)), self._err)
        return {names[i].decode(): (ms[i], int(cnt[i])) for i in range(min(n.value, mx))}


def relax(engine, ff, pos_dyn, pos_water, steps=300, dt=0.002):
    """Removes close contacts of a built structure by an overdamped quench:
    MD at 0 K with friction 10⁴/ps (velocities are reset every step, so each
    step moves atoms along their clamped accelerations, as steepest descent).
    Returns relaxed (pos_dyn, pos_water)."""
    m = Md(engine, ff, dt=dt, temperature=0.0, gamma=1.0e4, seed=0)
    nd, nw = engine.n_dyn, engine.n_water
    m.set_state(pos_dyn, np.zeros((nd, 3), np.float32), pos_water, np.zeros((nw, 3, 3), np.float32))
    m.step(steps)
    pd, _, pw, _ = m.state()
    m.close()
    return pd, pw
