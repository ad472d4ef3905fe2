"""Neighbor-list build on the GPU (build_neighbors, src/neighbors.rs:180-281):
identical CSR lists to the exact f32 brute-force restatement."""

import numpy as np
import pytest

from dynamics_amd import api, systems
from oracle import reference_path as rp

pytestmark = pytest.mark.gpu


def _check(ctx, outer, inner, static, ext, symmetric, thr):
    rp_g, col_g = ctx.build_neighbors(outer, inner, static, ext, symmetric, thr)
    rp_o, col_o = rp.build_neighbors(outer, inner, static, ext, symmetric, thr)
    assert np.array_equal(rp_g, rp_o)
    assert np.array_equal(col_g, col_o)
    return rp_g, col_g


@pytest.mark.parametrize("seed", [2, 5])
def test_build_all_neighbors_matches_bruteforce(gpu, seed):
    s = systems.solvated_chain(27.0, 600, 30, 40, seed=seed, lattice_n=10)
    ctx = api.ForcesPositsGpu(s.n_dyn, s.n_water, 10.0, 0.3123)
    ext = s.box.astype(np.float32)
    thr = np.float32(14.0 * 14.0)
    wpos = rp.wrap_ref(s.pos_o, np.zeros(3, np.float32), ext)
    static = np.zeros(s.n_dyn, np.uint8)
    static[::7] = 1
    _check(ctx, s.pos_dyn, s.pos_dyn, static, ext, True, thr)
    _check(ctx, s.pos_dyn, s.pos_dyn, None, ext, True, thr)
    _check(ctx, s.pos_dyn, wpos, None, ext, False, thr)
    _check(ctx, wpos, wpos, None, ext, True, thr)


def test_unwrapped_solute_and_threshold_ties(gpu):
    """Solute coordinates outside the box and points placed exactly on the
    threshold distance and on half-box separations."""
    rng = np.random.default_rng(7)
    ext = np.array([30.0, 30.0, 30.0], np.float32)
    p = (rng.uniform(-45, 75, (400, 3))).astype(np.float32)
    p[1] = p[0] + np.array([12.0, 0, 0], np.float32)   # |d| = 12 exactly: strict '<' excludes
    p[3] = p[2] + np.array([15.0, 0, 0], np.float32)   # half box
    p[5] = p[4] + np.array([-15.0, 0, 0], np.float32)
    ctx = api.ForcesPositsGpu(400, 0, 10.0, 0.3)
    _check(ctx, p, p, None, ext, True, np.float32(144.0))


def test_config0_water_water(gpu):
    s = systems.config0()
    ext = s.box.astype(np.float32)
    wpos = rp.wrap_ref(s.pos_o, np.zeros(3, np.float32), ext)
    ctx = api.ForcesPositsGpu(0, s.n_water, 10.0, 0.3123)
    rp_g, col_g = _check(ctx, wpos, wpos, None, ext, True, np.float32(196.0))
    # symmetric: j ∈ row i ⇔ i ∈ row j
    n = len(rp_g) - 1
    rows = np.repeat(np.arange(n), np.diff(rp_g.astype(np.int64)))
    a = set(zip(rows.tolist(), col_g.tolist()))
    assert all((j, i) in a for (i, j) in list(a)[:20000])


def test_neighbor_errors_and_empty(gpu):
    ctx = api.ForcesPositsGpu(4, 0, 10.0, 0.3)
    p = np.zeros((4, 3), np.float32)
    ext = np.full(3, 20.0, np.float32)
    with pytest.raises(api.DynError):
        ctx.build_neighbors(p, p, np.zeros(4, np.uint8), ext, False, 100.0)
    with pytest.raises(api.DynError):
        ctx.build_neighbors(p, p[:3], None, ext, True, 100.0)
    rp_g, col_g = ctx.build_neighbors(p, np.zeros((0, 3), np.float32), None, ext, False, 100.0)
    assert not rp_g.any() and col_g.size == 0
