"""Native engine on the GPU: cluster-pair list + 8×8 wave tiles + corrections
+ SPME.  Fixed-point pair sums bit-exact against the host tile replay over
the GPU-built list; forces within 1e-5 relative RMS and energies against the
FP64 oracle of the same Hamiltonian, in reference and standard modes."""

import numpy as np
import pytest

from conftest import rel_rms, replay_engine_tiles
from dynamics_amd import api, systems
from oracle import fp64, reference_path as rp

pytestmark = pytest.mark.gpu

RC = 10.0
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
    # 1-4 pairs are evaluated by the correction kernel, not counted by the tiles;
    # pairs within f32 rounding of the cutoff may fall either side
    n64 = fp64.pair_sum_fp64(s, RC, RC, BETA, include_14=False)["pair_count"]
    assert abs(E["pair_count"] - n64) <= 2
    # Energy components are sums of many cancelling pair terms; the 1e-6 gate
    # applies to the total (test_full_evaluation_vs_fp64).  Components are held
    # to 5e-6 of their own magnitude (f32 pair arithmetic, ~1e-7 per term).
    assert abs(E["lj"] - ref["e_lj"]) <= 5e-6 * abs(ref["e_lj"]) + 1e-6, (E["lj"], ref["e_lj"])
    assert abs(E["coul_sr"] - ref["e_coul"]) <= 5e-6 * abs(ref["e_coul"]), (E["coul_sr"], ref["e_coul"])
    assert abs(E["virial_sr"] - ref["virial"]) <= 5e-6 * abs(ref["virial"]), (E["virial_sr"], ref["virial"])
    # fixed-point pair sums: bit-exact against the host replay of the tiles
    dbg = eng.debug_export()
    fx = replay_engine_tiles(dbg, s, RC, RC, BETA)
    assert np.array_equal(fx, dbg["fx_slot"]), "slot fixed-point sums differ from replay"


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
