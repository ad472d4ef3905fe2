"""The reference device boundary on the GPU (force_nonbonded_gpu over the
flattened pair list, src/gpu_interface.rs:463-632; 1-4 correction,
src/non_bonded.rs:763-790): bit-exact fixed-point sums against the host
replay, forces within 1e-5 relative RMS of the FP64 oracle."""

import numpy as np
import pytest

from conftest import rel_rms, replay_ref_pairs
from dynamics_amd import api, systems
from oracle import fp64, reference_path as rp

pytestmark = pytest.mark.gpu

RC, ALPHA = 10.0, 0.3123


def _flat_fx(system, fx):
    return fx.reshape(-1, 3)


@pytest.fixture(scope="module")
def solv():
    s = systems.small_solvated()
    nb = rp.build_all_neighbors(s, skin=4.0, cutoff=RC)
    pairs = rp.setup_pairs(s, nb)
    return s, pairs


def _ctx(s, **kw):
    return api.ForcesPositsGpu(s.n_dyn, s.n_water, RC, ALPHA, **kw)


def _compute(ctx, s, **kw):
    return ctx.compute(s.pos_dyn, s.pos_o, s.pos_m, s.pos_h0, s.pos_h1, s.box, **kw)


def _fp64_as_sites(f_dyn, f_w):
    return np.concatenate([f_dyn, f_w.reshape(-1, 3)])


@pytest.mark.parametrize("lj_cutoff", [np.inf, 10.0])
def test_compat_bit_exact_and_fp64(gpu, solv, lj_cutoff):
    s, pairs = solv
    ctx = _ctx(s, lj_cutoff=lj_cutoff)
    ctx.set_pairs(pairs)
    f_dyn, f_w, vir, en, dh = _compute(ctx, s)
    fx = ctx.fixed_forces()
    fx_r, e_r, w_r, _ = replay_ref_pairs(s, pairs, RC, ALPHA, lj_cutoff=lj_cutoff)
    assert np.array_equal(fx, fx_r), "fixed-point sums differ from the host replay"
    assert abs(en - e_r) <= 1e-9 * abs(e_r) + 1e-9
    assert abs(vir - w_r) <= 1e-9 * abs(w_r)
    d64, w64, vir64, e64, _ = fp64.ref_pairs_fp64(s, pairs, RC, ALPHA, lj_cutoff=lj_cutoff)
    err = rel_rms(_fp64_as_sites(f_dyn, f_w), _fp64_as_sites(d64, w64))
    assert err <= 1e-5, f"force rel RMS {err:.2e}"
    assert abs(en - e64) <= 1e-6 * abs(e64), (en, e64)
    assert abs(vir - vir64) <= 1e-6 * abs(vir64), (vir, vir64)
    ctx.close()


def test_compat_deterministic(gpu, solv):
    s, pairs = solv
    ctx = _ctx(s)
    ctx.set_pairs(pairs)
    a = _compute(ctx, s)
    fa = ctx.fixed_forces().copy()
    b = _compute(ctx, s)
    assert np.array_equal(fa, ctx.fixed_forces())
    assert a[2] == b[2] and a[3] == b[3]


@pytest.mark.parametrize("lam", [0.0, 0.2, 0.5, 0.7, 1.0])
def test_compat_alchemical(gpu, solv, lam):
    """nonbonded_force_alchemical_kernel (src/cuda/cuda.cu:367-566): one
    molecule's cross pairs soft-core LJ, Coulomb scaled by the staged schedule."""
    s = solv[0]
    nb = rp.build_all_neighbors(s, skin=4.0, cutoff=RC)
    alch = np.zeros(s.n_dyn, bool)
    alch[:12] = True  # the decoupled "molecule"
    pairs = rp.setup_pairs(s, nb, alch_atoms=alch)
    assert pairs["alch"].any()
    ctx = _ctx(s)
    ctx.set_pairs(pairs)
    f_dyn, f_w, vir, en, dh = _compute(ctx, s, lambda_alch=lam)
    fx_r, e_r, w_r, dh_r = replay_ref_pairs(s, pairs, RC, ALPHA, lambda_alch=lam)
    assert np.array_equal(ctx.fixed_forces(), fx_r)
    d64, w64, vir64, e64, dh64 = fp64.ref_pairs_fp64(s, pairs, RC, ALPHA, lambda_alch=lam)
    assert rel_rms(_fp64_as_sites(f_dyn, f_w), _fp64_as_sites(d64, w64)) <= 1e-5
    assert abs(dh - dh64) <= 1e-5 * max(1.0, abs(dh64)), (dh, dh64)
    assert abs(en - e64) <= 1e-6 * abs(e64)


def test_compat_overrides(gpu, solv):
    s, pairs = solv
    ctx = _ctx(s)
    ctx.set_pairs(pairs)
    for cd, ld in [(True, False), (False, True), (True, True)]:
        f_dyn, f_w, vir, en, _ = _compute(ctx, s, coulomb_disabled=cd, lj_disabled=ld)
        d64, w64, vir64, e64, _ = fp64.ref_pairs_fp64(s, pairs, RC, ALPHA, coulomb_disabled=cd,
                                                      lj_disabled=ld)
        if cd and ld:
            assert not f_dyn.any() and not f_w.any() and en == 0.0
        else:
            assert rel_rms(_fp64_as_sites(f_dyn, f_w), _fp64_as_sites(d64, w64)) <= 1e-5


def test_compat_config0_water(gpu):
    """2,048 OPC waters, 39.4 Å box: the reference's full Verlet list
    (rc + 4 Å skin, ~3.9 M entries)."""
    s = systems.config0()
    nb = rp.build_all_neighbors(s, skin=4.0, cutoff=RC)
    pairs = rp.setup_pairs(s, nb)
    assert len(pairs["tgt"]) > 3_000_000
    ctx = _ctx(s)
    ctx.set_pairs(pairs)
    f_dyn, f_w, vir, en, _ = _compute(ctx, s)
    fx_r, e_r, w_r, _ = replay_ref_pairs(s, pairs, RC, ALPHA)
    assert np.array_equal(ctx.fixed_forces(), fx_r)
    d64, w64, vir64, e64, _ = fp64.ref_pairs_fp64(s, pairs, RC, ALPHA)
    err = rel_rms(f_w.reshape(-1, 3), w64.reshape(-1, 3))
    assert err <= 1e-5, f"force rel RMS {err:.2e}"
    assert en == 0.0  # all pairs are water–water: the reference reports no energy
    assert abs(vir - vir64) <= 1e-6 * abs(vir64)


def test_correct_14(gpu, solv):
    s = solv[0]
    ctx = _ctx(s)
    f = np.zeros((s.n_dyn, 3))
    w = ctx.correct_14(s.pairs_14[:, 0], s.pairs_14[:, 1], s.pos_dyn, s.q_dyn, s.box, f)
    f64, w64 = fp64.corr_14_vacuum(s.pos_dyn, s.q_dyn, s.pairs_14.astype(np.int64), s.box)
    assert rel_rms(f, f64) <= 1e-6
    assert abs(w - w64) <= 1e-6 * abs(w64)


def test_compat_errors(gpu, solv):
    s, pairs = solv
    ctx = _ctx(s)
    with pytest.raises(api.DynError) as e:
        _compute(ctx, s)
    assert e.value.code == api.DYN_ERR_STATE
    bad = dict(pairs)
    bad["tgt"] = pairs["tgt"].copy()
    bad["tgt"][0] = 10 ** 6
    with pytest.raises(api.DynError) as e:
        ctx.set_pairs(bad)
    assert e.value.code == api.DYN_ERR_ARG
    ctx.set_pairs(pairs)
    p = s.pos_dyn.copy()
    p[3, 1] = np.nan
    with pytest.raises(api.DynError) as e:
        ctx.compute(p, s.pos_o, s.pos_m, s.pos_h0, s.pos_h1, s.box)
    assert e.value.code == api.DYN_ERR_ARG
    # the handle stays usable after an error
    f_dyn, f_w, vir, en, _ = _compute(ctx, s)
    assert np.isfinite(f_dyn).all()
    # an empty pair list gives zero forces
    empty = {k: v[:0] for k, v in pairs.items()}
    ctx.set_pairs(empty)
    f_dyn, f_w, vir, en, _ = _compute(ctx, s)
    assert not f_dyn.any() and not f_w.any() and vir == 0.0
