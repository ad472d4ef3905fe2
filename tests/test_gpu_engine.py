"""Native engine on the GPU: cluster-pair list + 8×8 wave tiles + corrections
+ SPME.  Fixed-point pair sums bit-exact against the host tile replay over
the GPU-built list; forces within 1e-5 relative RMS and energies against the
FP64 oracle of the same Hamiltonian, in reference and standard modes."""

import copy
import os
import re

import numpy as np
import pytest

from conftest import normalize_null_entries, rel_rms, replay_engine_list, replay_engine_tiles
from dynamics_amd import api, systems
from oracle import fp64, reference_path as rp

pytestmark = pytest.mark.gpu

RC = 10.0
# j groups per work item (DYN_CHUNK of the engine's header)
CHUNK = int(re.search(r"#define DYN_CHUNK (\d+)", open(os.path.join(
    os.path.dirname(__file__), "..", "dynamics_amd", "csrc", "engine.h")).read()).group(1))
BETA = float(np.float32(0.3123))  # the engine takes β as f32


@pytest.fixture(scope="module")
def cases():
    return {
        "water512": systems.small_water(),
        "solvated": systems.small_solvated(),
    }


def _engine(s, **kw):
    kw.setdefault("cutoff", RC)
    kw.setdefault("beta", BETA)
    return api.Engine(s, **kw)


@pytest.mark.parametrize("name", ["water512", "solvated"])
def test_short_range_vs_fp64_and_replay(gpu, cases, name):
    s = cases[name]
    eng = _engine(s)
    eng.compute(api.COMPUTE_SHORT | api.COMPUTE_CORR)
    F = eng.site_forces()
    E = eng.energies()
    ref = fp64.pair_sum_fp64(s, RC, RC, BETA)
    err = rel_rms(F, ref["forces"])
    assert err <= 1e-5, f"force rel RMS {err:.2e}"
    # the tiles count their pairs, the correction kernel the 1-4 pairs; pairs
    # within f32 rounding of the cutoff may fall either side
    assert abs(E["pair_count"] - ref["pair_count"]) <= 2
    # Energy components are sums of many cancelling pair terms; the 1e-6 gate
    # applies to the total (test_full_evaluation_vs_fp64).  Components are held
    # to 5e-6 of their own magnitude (f32 pair arithmetic, ~1e-7 per term).
    assert abs(E["lj"] - ref["e_lj"]) <= 5e-6 * abs(ref["e_lj"]) + 1e-6, (E["lj"], ref["e_lj"])
    assert abs(E["coul_sr"] - ref["e_coul"]) <= 5e-6 * abs(ref["e_coul"]), (E["coul_sr"], ref["e_coul"])
    assert abs(E["virial_sr"] - ref["virial"]) <= 5e-6 * abs(ref["virial"]), (E["virial_sr"], ref["virial"])
    # fixed-point pair sums: bit-exact against the host replay of the tiles
    dbg = eng.debug_export()
    fx, _, ie = replay_engine_tiles(dbg, s, RC, RC, BETA, per_item=True)
    assert np.array_equal(fx, dbg["fx_slot"]), "slot fixed-point sums differ from replay"
    _check_item_sums(eng, dbg, ie)


def _check_item_sums(eng, dbg, ie):
    """Per work item energies, virial and dH/dλ against the host replay: a
    sum the replay leaves at zero (e.g. a solute energy of a water-only item)
    is zero on the device, the rest agree to f32 rounding."""
    g = eng.debug_item_partials(len(dbg["items"]))[:, [0, 1, 2, 3, 4, 8]]
    assert np.array_equal(g == 0.0, ie == 0.0), "energy classification differs from replay"
    scale = np.abs(ie).sum(0) + 1e-12
    assert (np.abs(g - ie).sum(0) <= 1e-5 * scale).all(), (np.abs(g - ie).sum(0), scale)


@pytest.mark.parametrize("name,buffer", [("water512", 0.0), ("solvated", 0.0), ("solvated", 1.5),
                                         ("chain2600_opc8600", 0.0)])
def test_pair_list_bit_exact(gpu, cases, name, buffer):
    """Clusters, slots, j entries (with padding) and exclusion masks equal
    the host restatement of the list build exactly."""
    s = cases[name] if name in cases else systems.config1()
    print(f"list parity: {name} sites={s.n_sites} buffer={buffer}", flush=True)
    eng = _engine(s, list_buffer=buffer)
    eng.compute(api.COMPUTE_SHORT)
    dbg = eng.debug_export()
    n, slot_site, cl_class, row_ptr, cj, cm = replay_engine_list(s, RC, buffer)
    assert len(dbg["cl_class"]) == n
    assert np.array_equal(dbg["slot_site"], slot_site)
    assert np.array_equal(dbg["cl_class"], cl_class)
    assert np.array_equal(dbg["row_ptr"], row_ptr)
    assert np.array_equal(normalize_null_entries(dbg["cj"], 8 * n), cj)
    grp_row = np.repeat(np.arange(n), np.diff(row_ptr.astype(np.int64)))
    ss = (cl_class[grp_row] == 0) & ((cj[::8] >> 30) == 0)
    assert np.array_equal(dbg["cmask"][ss], cm[ss]), "solute–solute exclusion masks differ"
    # work items: every row cut into chunks of at most CHUNK groups, in order
    items = dbg["items"]
    ch = CHUNK
    nch = (np.diff(row_ptr.astype(np.int64)) + ch - 1) // ch
    want_I = np.repeat(np.arange(n), nch)
    want_g = np.concatenate([row_ptr[i] + ch * np.arange(k) for i, k in enumerate(nch)])
    assert np.array_equal(items[:, 0], want_I) and np.array_equal(items[:, 1], want_g)


def test_pair_count_matches_oracle(gpu, cases):
    s = cases["water512"]
    eng = _engine(s)
    eng.compute(api.COMPUTE_SHORT)
    ref = fp64.pair_sum_fp64(s, RC, RC, BETA)
    assert abs(eng.energies()["pair_count"] - ref["pair_count"]) <= 2


@pytest.mark.parametrize("mode", [api.MODE_REFERENCE, api.MODE_STANDARD])
@pytest.mark.parametrize("name", ["water512", "solvated"])
def test_full_evaluation_vs_fp64(gpu, cases, name, mode):
    s = cases[name]
    K = api.pme_grid_n(s.box, 1.0)
    eng = _engine(s, grid=K, order=4, mode=mode)
    eng.compute(api.COMPUTE_ALL)
    F = eng.site_forces()
    E = eng.energies()
    ref = fp64.reference_energy_and_forces(s, RC, RC, BETA, K, 4,
                                           mode="reference" if mode == api.MODE_REFERENCE else "standard")
    err = rel_rms(F, ref["forces"])
    assert err <= 1e-5, f"force rel RMS {err:.2e}"
    assert abs(E["recip"] - ref["e_recip"]) <= 1e-6 * abs(ref["e_recip"]) + 1e-6
    assert abs(E["self"] - ref["e_self"]) <= 1e-9 * abs(ref["e_self"])
    tot_scale = abs(ref["e_lj"]) + abs(ref["e_coul"]) + abs(ref["e_recip"]) + abs(ref["e_self"]) + abs(ref["e_corr"])
    assert abs(E["total"] - ref["e_total"]) <= 1e-6 * tot_scale, (E["total"], ref["e_total"])
    if mode == api.MODE_REFERENCE:
        assert abs(E["reference"] - ref["e_reference"]) <= 1e-6 * tot_scale
    # the reciprocal virial is the k-space sum plus the correction terms,
    # which largely cancel in standard mode: hold it to the scale of the parts
    w_scale = abs(ref["virial_kspace"]) + abs(ref["virial_corr"])
    assert abs(E["virial_recip"] - ref["virial_recip"]) <= 1e-5 * w_scale, (E["virial_recip"], ref["virial_recip"], w_scale)


def test_deterministic(gpu, cases):
    s = cases["solvated"]
    eng = _engine(s, grid=(27, 27, 27), order=4)
    eng.compute(api.COMPUTE_ALL)
    a = eng.fixed_forces().copy()
    ea = eng.energies()
    eng.compute(api.COMPUTE_ALL | api.COMPUTE_REBUILD)
    assert np.array_equal(a, eng.fixed_forces())
    eb = eng.energies()
    assert ea["total"] == eb["total"] and ea["virial_sr"] == eb["virial_sr"]


@pytest.mark.parametrize("alch", [False, True])
def test_no_energy_forces_identical(gpu, cases, alch):
    """DYN_COMPUTE_NO_ENERGY gives the same fixed-point forces, bit for bit,
    and get_energies refuses until a compute with energies."""
    s = cases["solvated"]
    eng = _engine(s, grid=(27, 27, 27), order=4)
    if alch:
        eng.set_alchemical(0, 40, 0.4)
    eng.compute(api.COMPUTE_ALL)
    a = eng.fixed_forces().copy()
    ea = eng.energies()
    eng.compute(api.COMPUTE_ALL | api.COMPUTE_NO_ENERGY)
    assert np.array_equal(a, eng.fixed_forces())
    with pytest.raises(api.DynError, match="skipped energies"):
        eng.energies()
    eng.compute(api.COMPUTE_ALL)
    eb = eng.energies()
    for k in ("list_builds",):  # every compute without a list buffer rebuilds
        ea.pop(k), eb.pop(k)
    assert eb == ea


def test_recip_reuse(gpu, cases):
    """DYN_COMPUTE_RECIP_REUSE (the reference's SPME cadence): the kept SPME
    forces of the last SPME evaluation are added to fresh short-range and
    1-4 terms, without the 1-4 vacuum correction — bit for bit the sum of
    the two separate evaluations; energy and virial of SPME are kept too."""
    s = cases["solvated"]
    eng = _engine(s, grid=(27, 27, 27), order=4, mode=api.MODE_REFERENCE)
    eng.compute(api.COMPUTE_RECIP)
    R = eng.fixed_forces().copy()
    er = eng.energies()
    eng.compute(api.COMPUTE_SHORT | api.COMPUTE_CORR | api.COMPUTE_RECIP_REUSE)
    C = eng.fixed_forces().copy()
    ec = eng.energies()
    eng.compute(api.COMPUTE_SHORT | api.COMPUTE_CORR)
    B = eng.fixed_forces().copy()
    assert np.array_equal(C, B + R)
    assert ec["recip"] == er["recip"] and ec["virial_recip"] == er["virial_recip"]
    # nothing kept (the last compute ran no SPME): the flag runs SPME, with
    # the 1-4 vacuum correction, as COMPUTE_ALL does
    eng.compute(api.COMPUTE_SHORT | api.COMPUTE_CORR | api.COMPUTE_RECIP_REUSE)
    D = eng.fixed_forces().copy()
    eng.compute(api.COMPUTE_ALL)
    assert np.array_equal(D, eng.fixed_forces())
    # a new box drops what was kept
    eng.set_box(s.box)
    eng.compute(api.COMPUTE_SHORT | api.COMPUTE_CORR | api.COMPUTE_RECIP_REUSE)
    assert np.array_equal(eng.fixed_forces(), D)
    eng.close()


def test_list_reuse_with_buffer(gpu, cases):
    """With a list buffer the list is reused while sites move less than half
    the buffer; the result equals a fresh build on the moved positions."""
    s = cases["water512"]
    eng = _engine(s, list_buffer=1.0)
    eng.compute(api.COMPUTE_SHORT)
    assert eng.energies()["list_builds"] == 1
    rng = np.random.default_rng(0)
    moved = systems.water_box(512, 24.8, 8, seed=1)
    shift = (rng.uniform(-1, 1, (512, 3)) * 0.15).astype(np.float32)
    for name in ["pos_o", "pos_m", "pos_h0", "pos_h1"]:
        setattr(moved, name, (getattr(moved, name) + shift).astype(np.float32))
    eng.set_positions(moved)
    eng.compute(api.COMPUTE_SHORT)
    assert eng.energies()["list_builds"] == 1
    F = eng.site_forces()
    ref = fp64.pair_sum_fp64(moved, RC, RC, BETA)
    assert rel_rms(F, ref["forces"]) <= 1e-5
    # a move past half the buffer triggers a rebuild
    for name in ["pos_o", "pos_m", "pos_h0", "pos_h1"]:
        getattr(moved, name)[0] += np.float32(0.8)
    eng.set_positions(moved)
    eng.compute(api.COMPUTE_SHORT)
    assert eng.energies()["list_builds"] == 2


def test_engine_matches_reference_pair_list_path(gpu, cases):
    """At a fresh Verlet list the reference's pair list (CPU semantics, LJ cut
    at 10 Å) and the engine's cluster-pair list evaluate the same pair set."""
    s = cases["solvated"]
    nb = rp.build_all_neighbors(s, skin=4.0, cutoff=RC)
    pairs = rp.setup_pairs(s, nb)
    ctx = api.ForcesPositsGpu(s.n_dyn, s.n_water, RC, BETA, lj_cutoff=RC)
    ctx.set_pairs(pairs)
    f_dyn, f_w, vir, en, _ = ctx.compute(s.pos_dyn, s.pos_o, s.pos_m, s.pos_h0, s.pos_h1, s.box)
    eng = _engine(s)
    eng.compute(api.COMPUTE_SHORT | api.COMPUTE_CORR)
    ed, ew = eng.forces()
    # the engine's CORR kernel also adds the 1-4 vacuum correction only when
    # reciprocal space is on; with grid 0 it adds the scaled 1-4 real-space terms
    a = np.concatenate([f_dyn, f_w.reshape(-1, 3)])
    b = np.concatenate([ed, ew.reshape(-1, 3)])
    assert rel_rms(b, a) <= 1e-5


def test_config0_full(gpu):
    s = systems.config0()
    eng = _engine(s, grid=(40, 40, 40), order=4)
    eng.compute(api.COMPUTE_ALL)
    F = eng.site_forces()
    ref = fp64.reference_energy_and_forces(s, RC, RC, BETA, (40, 40, 40), 4)
    err = rel_rms(F, ref["forces"])
    assert err <= 1e-5, f"force rel RMS {err:.2e}"
    E = eng.energies()
    assert abs(E["pair_count"] - ref["pair_count"]) <= 4


def _full_size_check(s, rc, K, mode=api.MODE_REFERENCE):
    eng = _engine(s, cutoff=rc, lj_cutoff=rc, grid=K, order=4, mode=mode)
    eng.compute(api.COMPUTE_ALL | api.COMPUTE_REBUILD)
    F = eng.site_forces()
    E = eng.energies()
    eng.close()
    ref = fp64.reference_energy_and_forces(
        s, rc, rc, BETA, K, 4, mode="reference" if mode == api.MODE_REFERENCE else "standard")
    err = rel_rms(F, ref["forces"])
    print(f"force rel RMS vs FP64 {err:.3e}", flush=True)
    assert err <= 1e-5, f"force rel RMS {err:.2e}"
    tot_scale = abs(ref["e_lj"]) + abs(ref["e_coul"]) + abs(ref["e_recip"]) + abs(ref["e_self"]) + abs(ref["e_corr"])
    assert abs(E["total"] - ref["e_total"]) <= 1e-6 * tot_scale, (E["total"], ref["e_total"])
    assert abs(E["reference"] - ref["e_reference"]) <= 1e-6 * tot_scale
    assert abs(E["recip"] - ref["e_recip"]) <= 1e-6 * abs(ref["e_recip"])
    # pairs within f32 rounding of the cutoff may fall either side
    assert abs(E["pair_count"] - ref["pair_count"]) <= max(4, 1e-6 * ref["pair_count"])
    # net force: the pair and correction sums are antisymmetric; SPME's net
    # force is a discretisation residue, small against the RMS force
    net = np.abs(F.sum(axis=0)).max() / np.sqrt((F ** 2).sum(axis=1).mean())
    assert net < 1e-2 * np.sqrt(len(F)), net


_RELAXED = {}


def relaxed_config1():
    """The 37,000-site solvated chain after the benchmark's quench of its
    built close contacts (the configuration the benchmark starts from)."""
    if "c1" not in _RELAXED:
        s = systems.config1()
        ff = systems.chain_force_field(s)
        eng = api.Engine(s, cutoff=10.0, lj_cutoff=10.0, beta=BETA, grid=(64, 64, 64), order=4,
                         list_buffer=1.0)
        pd, pw = api.relax(eng, ff, s.pos_dyn, systems.water_positions(s), steps=400)
        eng.close()
        r = copy.copy(s)
        r.pos_dyn = pd
        r.pos_o, r.pos_m, r.pos_h0, r.pos_h1 = (np.ascontiguousarray(pw[:, k]) for k in range(4))
        _RELAXED["c1"] = r
    return _RELAXED["c1"]


def test_bench_workload_vs_fp64(gpu):
    """The 37,000-site solvated chain at the benchmark's settings (rc 10 Å,
    SPME 64³ order 4, reference mode, list rebuilt), from the benchmark's
    relaxed start."""
    s = relaxed_config1()
    print(f"bench workload parity: sites={s.n_sites}", flush=True)
    _full_size_check(s, 10.0, (64, 64, 64))


def test_water_128A_vs_fp64(gpu):
    """69,900 OPC waters in a 128 Å box, rc 9 Å, SPME 128³ order 4: the
    largest single-system workload, at full size against FP64."""
    s = systems.config2()
    print(f"128 A water parity: sites={s.n_sites}", flush=True)
    _full_size_check(s, 9.0, (128, 128, 128))


def _alch_check(s, alch0, n_alch, lam, K, rc=RC, replay=False):
    eng = _engine(s, cutoff=rc, lj_cutoff=rc, grid=K, order=4, mode=api.MODE_REFERENCE)
    eng.set_alchemical(alch0, n_alch, lam)
    eng.compute(api.COMPUTE_ALL)
    F = eng.site_forces()
    E = eng.energies()
    alch = np.zeros(s.n_dyn, bool)
    alch[alch0:alch0 + n_alch] = True
    ref = fp64.reference_energy_and_forces(s, rc, rc, BETA, K, 4, alch=alch, lam=lam)
    err = rel_rms(F, ref["forces"])
    print(f"alchemical lambda={lam}: force rel RMS {err:.3e}, dH/dl {E['dh_dl']:.6f} vs "
          f"{ref['dh_dl']:.6f}", flush=True)
    assert err <= 1e-5, f"force rel RMS {err:.2e}"
    tot_scale = abs(ref["e_lj"]) + abs(ref["e_coul"]) + abs(ref["e_recip"]) + abs(ref["e_self"])
    assert abs(E["total"] - ref["e_total"]) <= 1e-6 * tot_scale, (E["total"], ref["e_total"])
    assert abs(E["reference"] - ref["e_reference"]) <= 1e-6 * tot_scale
    # dH/dλ is a sum over the molecule's cross pairs and the SPME cross term
    assert abs(E["dh_dl"] - ref["dh_dl"]) <= 1e-5 * max(abs(ref["dh_dl"]), 1.0), (E["dh_dl"], ref["dh_dl"])
    if replay:
        eng.compute(api.COMPUTE_SHORT)
        dbg = eng.debug_export()
        fx, _, ie = replay_engine_tiles(dbg, s, rc, rc, BETA, alch=True, lam=lam, per_item=True)
        assert np.array_equal(fx, dbg["fx_slot"]), "alchemical slot sums differ from replay"
        _check_item_sums(eng, dbg, ie)
    eng.close()


@pytest.mark.parametrize("lam", [0.0, 0.2, 0.5, 0.8, 1.0])
def test_alchemical_vs_fp64(gpu, cases, lam):
    """nonbonded_force_alchemical_kernel semantics (src/cuda/cuda.cu:367-566)
    with the SPME cross-term scaling (src/non_bonded.rs:709-742), on a
    12-atom stretch of the chain (1-4 pairs cross its boundary)."""
    _alch_check(cases["solvated"], 4, 12, lam, (27, 27, 27), replay=True)


def test_alchemical_bench_system(gpu):
    """40 decoupled atoms of the 37,000-site solvated chain at λ = 0.6."""
    s = relaxed_config1()
    _alch_check(s, 0, 40, 0.6, (64, 64, 64))


def test_alchemical_range_change_rebuilds(gpu, cases):
    s = cases["solvated"]
    eng = _engine(s, grid=(27, 27, 27), order=4)
    eng.compute(api.COMPUTE_ALL)
    plain = eng.fixed_forces().copy()
    eng.set_alchemical(4, 12, 0.3)
    eng.compute(api.COMPUTE_ALL)
    assert not np.array_equal(plain, eng.fixed_forces())
    eng.set_alchemical(0, 0, 0.3)   # off again: back to the plain result exactly
    eng.compute(api.COMPUTE_ALL)
    assert np.array_equal(plain, eng.fixed_forces())
    with pytest.raises(api.DynError):
        eng.set_alchemical(s.n_dyn - 2, 5, 0.5)   # beyond the solute
    std = _engine(s, grid=(27, 27, 27), order=4, mode=api.MODE_STANDARD)
    with pytest.raises(api.DynError):
        std.set_alchemical(0, 4, 0.5)             # reference mode only


def test_engine_errors(gpu, cases):
    s = cases["water512"]
    with pytest.raises(api.DynError):
        _engine(s, lj_cutoff=12.0)           # LJ cutoff beyond the list cutoff
    with pytest.raises(api.DynError):
        _engine(s, beta=0.5)                 # β·rc beyond the erfc range
    with pytest.raises(api.DynError):
        _engine(s, cutoff=13.0)              # beyond half the 24.8 Å box
    eng = _engine(s)
    bad = systems.small_water()
    bad.pos_o[0, 0] = np.nan
    with pytest.raises(api.DynError):
        eng.set_positions(bad)


def test_moved_site_rebuilds_on_device(gpu, cases):
    """An asynchronous compute decides on the device whether to rebuild: a
    site that moved half the list buffer since the last build makes the
    evaluation rebuild before its pair kernel, so the result equals a fresh
    build on the moved positions bit for bit, with no error."""
    s = cases["water512"]
    eng = _engine(s, list_buffer=1.0)
    eng.compute(api.COMPUTE_ALL)
    assert eng.energies()["list_builds"] == 1
    moved = copy.deepcopy(s)
    for name in ["pos_o", "pos_m", "pos_h0", "pos_h1"]:
        getattr(moved, name)[3] += np.float32(0.6)
    eng.set_positions(moved)
    eng.compute(api.COMPUTE_ALL | api.COMPUTE_ASYNC)
    eng.sync()
    assert eng.energies()["list_builds"] == 2
    fresh = _engine(moved, list_buffer=1.0)
    fresh.compute(api.COMPUTE_ALL)
    assert np.array_equal(eng.fixed_forces(), fresh.fixed_forces())
    eng.close()
    fresh.close()
