"""The precision options (DYN_PREC_PAIR_F64, DYN_PREC_PME_F64) against the
FP64 oracle: each part in FP64 arithmetic matches the FP64 result to the
fixed-point resolution of the force sums (2^-32 kcal/(mol·Å) per term) and
the spread grid's (2^-32 per charge unit), orders of magnitude inside the
FP32 path's error, which is measured alongside."""

import numpy as np
import pytest

from conftest import rel_rms
from dynamics_amd import api, systems
from oracle import fp64

pytestmark = pytest.mark.gpu

BETA = float(np.float32(0.3123))
RC = 10.0


@pytest.fixture(scope="module")
def solvated():
    return systems.small_solvated()


def _engine(s, prec, K):
    return api.Engine(s, cutoff=RC, lj_cutoff=RC, beta=BETA, grid=K, order=4,
                      mode=api.MODE_REFERENCE, precision=prec)


def test_pair_f64(gpu, solvated):
    s = solvated
    ref = fp64.pair_sum_fp64(s, RC, RC, BETA, include_14=False)
    err = {}
    for prec in (0, api.PREC_PAIR_F64):
        eng = _engine(s, prec, (27, 27, 27))
        eng.compute(api.COMPUTE_SHORT)
        err[prec] = rel_rms(eng.site_forces(), ref["forces"])
        E = eng.energies()
        e_err = abs(E["lj"] + E["coul_sr"] - ref["e_lj"] - ref["e_coul"])
        print(f"pair precision {prec}: force rel RMS {err[prec]:.2e}, energy err {e_err:.2e}",
              flush=True)
        if prec:
            assert e_err <= 1e-9 * (abs(ref["e_lj"]) + abs(ref["e_coul"]))
        eng.close()
    assert err[api.PREC_PAIR_F64] <= 1e-9
    assert err[0] <= 1e-5 and err[api.PREC_PAIR_F64] < 1e-2 * err[0]


def test_pme_f64(gpu, solvated):
    s = solvated
    K = (27, 27, 27)
    pos, q, _, _, _, _ = s.site_arrays()
    f_ref, e_ref, w_ref = fp64.spme_fp64(pos.astype(np.float64), q.astype(np.float64), s.box, K, 4,
                                         BETA)
    err = {}
    for prec in (0, api.PREC_PME_F64):
        eng = _engine(s, prec, K)
        eng.compute(api.COMPUTE_RECIP)
        F = eng.site_forces()
        err[prec] = rel_rms(F, f_ref)
        E = eng.energies()
        print(f"PME precision {prec}: force rel RMS {err[prec]:.2e}, energy "
              f"{E['recip']:.10f} vs {e_ref:.10f}", flush=True)
        if prec:
            assert abs(E["recip"] - e_ref) <= 1e-9 * abs(e_ref)
        eng.close()
    assert err[api.PREC_PME_F64] <= 1e-8
    assert err[0] <= 1e-5 and err[api.PREC_PME_F64] < 1e-2 * err[0]


def test_both_f64_full(gpu, solvated):
    s = solvated
    K = (27, 27, 27)
    ref = fp64.reference_energy_and_forces(s, RC, RC, BETA, K, 4)
    eng = _engine(s, api.PREC_PAIR_F64 | api.PREC_PME_F64, K)
    eng.compute(api.COMPUTE_ALL)
    err = rel_rms(eng.site_forces(), ref["forces"])
    E = eng.energies()
    print(f"FP64 pairs + FP64 SPME: force rel RMS {err:.2e}, total {E['total']:.8f} vs "
          f"{ref['e_total']:.8f}", flush=True)
    # the 1-4 corrections stay in f32 arithmetic: a few hundred pairs
    assert err <= 1e-7
    eng.close()


def test_precision_errors(gpu, solvated):
    with pytest.raises(api.DynError):
        _engine(solvated, 4, (27, 27, 27))           # unknown bit
    eng = _engine(solvated, api.PREC_PAIR_F64, (27, 27, 27))
    with pytest.raises(api.DynError):
        eng.set_alchemical(4, 12, 0.5)               # alchemical runs in f32 only
    eng.close()
