"""The allocation-free 1-4 correction (dyn_nb_correct_14_noalloc) against the
allocating entry point it mirrors (dyn_nb_correct_14, src/non_bonded.rs:
763-790), through the C ABI: bit-identical forces and virial on two calls in
a row, the second smaller (fewer pairs, then a context with fewer atoms),
both accumulating into non-zero caller arrays; sizes beyond the fixed
scratch area are refused with DYN_ERR_LIMIT."""

import numpy as np
import pytest

from dynamics_amd import api, systems
from oracle import fp64

pytestmark = pytest.mark.gpu

RC, ALPHA = 10.0, 0.3123


def _both(ctx, i, j, pos, q, box, f0, w0):
    fa = f0.copy()
    wa = ctx.correct_14(i, j, pos, q, box, fa, virial=w0)
    fb = f0.copy()
    wb = ctx.correct_14(i, j, pos, q, box, fb, virial=w0, noalloc=True)
    return fa, wa, fb, wb


def test_noalloc_matches_allocating_entry(gpu):
    s = systems.small_solvated()
    rng = np.random.default_rng(5)
    p14 = np.asarray(s.pairs_14, np.int64)
    print(f"no-alloc 1-4: {s.n_dyn} atoms, {len(p14)} pairs", flush=True)
    ctx = api.ForcesPositsGpu(s.n_dyn, s.n_water, RC, ALPHA)
    f0 = rng.normal(size=(s.n_dyn, 3))
    # call 1: every 1-4 pair
    fa, wa, fb, wb = _both(ctx, p14[:, 0], p14[:, 1], s.pos_dyn, s.q_dyn, s.box, f0, 1.25)
    assert np.array_equal(fa, fb) and wa == wb
    f64, w64 = fp64.corr_14_vacuum(s.pos_dyn, s.q_dyn, p14, s.box)
    assert np.linalg.norm((fb - f0) - f64) <= 1e-6 * np.linalg.norm(f64)
    assert abs((wb - 1.25) - w64) <= 1e-6 * abs(w64)
    # call 2: a smaller input on the same device (a third of the pairs)
    sub = p14[: len(p14) // 3]
    fa, wa, fb, wb = _both(ctx, sub[:, 0], sub[:, 1], s.pos_dyn, s.q_dyn, s.box, f0, 0.0)
    assert np.array_equal(fa, fb) and wa == wb
    ctx.close()
    # call 3: a context with fewer atoms (the scratch keeps call 2's contents)
    n = s.n_dyn // 2
    keep = (p14[:, 0] < n) & (p14[:, 1] < n)
    small = p14[keep]
    ctx2 = api.ForcesPositsGpu(n, 0, RC, ALPHA)
    f0s = rng.normal(size=(n, 3))
    fa, wa, fb, wb = _both(ctx2, small[:, 0], small[:, 1], s.pos_dyn[:n], s.q_dyn[:n], s.box,
                           f0s, -3.0)
    assert np.array_equal(fa, fb) and wa == wb
    ctx2.close()


def test_noalloc_limits(gpu):
    n = (1 << 16) + 1
    ctx = api.ForcesPositsGpu(n, 0, RC, ALPHA)
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = np.arange(n, dtype=np.float32) * 1e-3
    q = np.ones(n, np.float32)
    f = np.zeros((n, 3))
    with pytest.raises(api.DynError) as e:
        ctx.correct_14([0], [1], pos, q, np.full(3, 100.0), f, noalloc=True)
    assert e.value.code == api.DYN_ERR_LIMIT
    # the allocating entry takes the same call
    ctx.correct_14([0], [1], pos, q, np.full(3, 100.0), f)
    ctx.close()
    ctx = api.ForcesPositsGpu(4, 0, RC, ALPHA)
    many = np.zeros((1 << 18) + 1, np.uint32)
    with pytest.raises(api.DynError) as e:
        ctx.correct_14(many, many + 1, np.eye(4, 3, dtype=np.float32) * 2.0,
                       np.ones(4, np.float32), np.full(3, 20.0), np.zeros((4, 3)), noalloc=True)
    assert e.value.code == api.DYN_ERR_LIMIT
    ctx.close()
