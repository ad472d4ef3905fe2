"""FP64 evaluation of the nonbonded Hamiltonian (test infrastructure only).

  pair_sum_fp64      all non-excluded site pairs with min-image r < rc:
                     LJ 12-6 (src/forces.rs:10-21) within lj_cutoff, erfc
                     Coulomb (src/cuda/cuda.cu:147-180), 1-4 pairs scaled by
                     0.5 / (1/1.2) (src/non_bonded.rs:55-56), energies split
                     into solute-involving and water–water pairs
                     (src/non_bonded.rs:337-342).  Pinned by the reference's
                     closed-form tests (tests/forces.rs:222-507): LJ minimum
                     −ε at 2^(1/6)σ, F = −dE/dr by central difference,
                     Newton-3, zero at r ≥ cutoff, erfc from scipy.
  ref_pairs_fp64     FP64 evaluation of an explicit flattened pair list with
                     the reference kernel's per-pair semantics
                     (src/cuda/cuda.cu:185-362 and :367-566, alchemical).
  bspline            cardinal B-splines M_n and their derivatives (Essmann
                     1995 eq. 4.1); pinned against the closed-form cubic.
  spme_fp64          smooth PME (Essmann 1995) of the same grid, order and β
                     as the engine.  Pinned against plain Ewald sums
                     (ewald_recip_fp64) on small systems and against the
                     reference's SPME tests (tests/spme.rs: neutral pair →
                     vacuum Coulomb within 1 %, F = −dE/dx).
  ewald_recip_fp64   direct reciprocal-space Ewald sum (independent of any
                     grid), with forces and the k-space virial.
  mol_energies       (pair_sum_fp64 with mol_start) the molecule-pair energy
                     matrix of solute–solute pairs (src/non_bonded.rs:
                     350-361).  Pinned by separability: for molecules A, B
                     the diagonal M_AA is the pair energy of A alone and
                     M_AB = E(A ∪ B) − E(A) − E(B), each E from the sum over
                     a sub-system holding only those atoms
                     (tests/test_oracle_mol_energies.py).
  corr_14_vacuum     the reference's force-only 1-4 correction
                     (src/non_bonded.rs:763-790).
  erf_pair_fp64      qq·erf(βr)/r pair term for exclusion corrections.
"""

from __future__ import annotations

import math

import numpy as np
from scipy.special import erfc, erf

F64 = np.float64


def min_image(d, box):
    return d - box * np.round(d / box)


def site_view(system):
    pos, q, sig, eps, mol, code = system.site_arrays()
    return (pos.astype(F64), q.astype(F64), sig.astype(np.float32), eps.astype(np.float32),
            mol, code)


def _pairs_within(pos, box, rc):
    from scipy.spatial import cKDTree
    w = np.mod(pos, box)
    w = np.where(w >= box, 0.0, w)
    tree = cKDTree(w, boxsize=box)
    pr = tree.query_pairs(rc * (1 + 1e-9) + 1e-9, output_type="ndarray")
    if pr.size == 0:
        return np.zeros((0, 2), np.int64)
    return pr.astype(np.int64)


def lj_fp64(r, sigma, eps):
    """E, |F| (along +r̂ on the target) of the 12-6 potential."""
    sr6 = (sigma / r) ** 6
    e = 4 * eps * (sr6 * sr6 - sr6)
    f = 24 * eps * (2 * sr6 * sr6 - sr6) / r
    return e, f


def coul_erfc_fp64(r, qq, beta):
    e = qq * erfc(beta * r) / r
    f = qq * (erfc(beta * r) / r ** 2 + 2 * beta / math.sqrt(math.pi) * np.exp(-(beta * r) ** 2) / r)
    return e, f


def erf_pair_fp64(r, qq, beta):
    """qq·erf(βr)/r and its |F| (= −dE/dr)."""
    e = qq * erf(beta * r) / r
    f = qq * (erf(beta * r) / r ** 2 - 2 * beta / math.sqrt(math.pi) * np.exp(-(beta * r) ** 2) / r)
    return e, f


def schedule(lam):
    """Staged decoupling (src/alchemical.rs:82-100): Coulomb is switched off
    over λ ∈ [0, ½], then LJ over [½, 1].  Returns (coulomb_scale,
    d coulomb_scale/dλ, LJ λ, d LJ λ/dλ)."""
    lam = min(1.0, max(0.0, float(lam)))
    if lam < 0.5:
        return 1.0 - 2.0 * lam, -2.0, 0.0, 0.0
    return 0.0, 0.0, 2.0 * lam - 1.0, 2.0


def pair_sum_fp64(system, rc=10.0, lj_rc=None, beta=0.3123, include_14=True, alch=None, lam=0.0,
                  mol_start=None):
    """Forces (n_sites, 3) in site order (solute, then per water O, M, H0, H1)
    and energies of all pairs inside the cutoff.  alch (bool per solute atom)
    marks a decoupled molecule: pairs with exactly one atom in it take
    soft-core LJ and scaled Coulomb at λ = lam (src/non_bonded.rs:300-345,
    :883-941), and "dh_dl" is returned.  mol_start (first solute atom of
    each molecule) adds "mol_energies": M[m_i, m_j] sums the energy of every
    solute–solute pair, both orders for m_i ≠ m_j, once on the diagonal
    (src/non_bonded.rs:350-361)."""
    lj_rc = rc if lj_rc is None else lj_rc
    pos, q, sig, eps, mol, code = site_view(system)
    box = system.box.astype(F64)
    nd = system.n_dyn
    pr = _pairs_within(pos, box, rc)
    i, j = pr[:, 0], pr[:, 1]
    swap = i > j
    i, j = np.where(swap, j, i), np.where(swap, i, j)
    d = min_image(pos[i] - pos[j], box)
    r = np.sqrt(np.einsum("ij,ij->i", d, d))
    keep = r < rc
    # exclusions: same water molecule; solute 1-2/1-3; 1-4 handled separately
    keep &= ~((mol[i] >= 0) & (mol[i] == mol[j]))
    both = (i < nd) & (j < nd)
    key = i * (nd + 1) + j
    if nd:
        ex = system.excl_12_13.astype(np.int64)
        p14 = system.pairs_14.astype(np.int64)
        exk = np.sort(ex[:, 0] * (nd + 1) + ex[:, 1]) if ex.size else np.zeros(0, np.int64)
        p4k = np.sort(p14[:, 0] * (nd + 1) + p14[:, 1]) if p14.size else np.zeros(0, np.int64)
        in_ex = both & np.isin(key, exk)
        in_14 = both & np.isin(key, p4k)
    else:
        in_ex = np.zeros_like(keep)
        in_14 = np.zeros_like(keep)
    keep &= ~in_ex
    i, j, d, r, in_14 = i[keep], j[keep], d[keep], r[keep], in_14[keep]
    if not include_14:
        m = ~in_14
        i, j, d, r, in_14 = i[m], j[m], d[m], r[m], in_14[m]
    s_ab = (np.float32(0.5) * (sig[i] + sig[j])).astype(F64)
    e_ab = np.sqrt((eps[i] * eps[j]).astype(np.float32)).astype(F64)
    qq = q[i] * q[j]
    lj_on = (e_ab > 0) & (r <= lj_rc)
    e_lj, f_lj = lj_fp64(r, s_ab, e_ab)
    e_lj = np.where(lj_on, e_lj, 0.0)
    f_lj = np.where(lj_on, f_lj, 0.0)
    c_on = qq != 0
    e_c, f_c = coul_erfc_fp64(r, qq, beta)
    e_c = np.where(c_on, e_c, 0.0)
    f_c = np.where(c_on, f_c, 0.0)
    dh = np.zeros_like(r)
    if alch is not None and np.any(alch):
        a_site = np.zeros(pos.shape[0], bool)
        a_site[:nd] = np.asarray(alch, bool)
        cross = a_site[i] ^ a_site[j]
        cs, cds, ll, lld = schedule(lam)
        e_sc, f_sc, dh_sc = soft_core_fp64(r, s_ab, e_ab, ll)
        e_lj = np.where(cross & lj_on, e_sc, e_lj)
        f_lj = np.where(cross & lj_on, f_sc, f_lj)
        dh = np.where(cross & lj_on, dh_sc * lld, 0.0)
    s_lj = np.where(in_14, 0.5, 1.0)
    s_c = np.where(in_14, 1.0 / 1.2, 1.0)
    e_lj *= s_lj
    f_lj *= s_lj
    dh *= s_lj
    e_c *= s_c
    f_c *= s_c
    if alch is not None and np.any(alch):
        dh = dh + np.where(cross, e_c * cds, 0.0)
        e_c = np.where(cross, e_c * cs, e_c)
        f_c = np.where(cross, f_c * cs, f_c)
    fmag = f_lj + f_c
    fvec = d * (fmag / r)[:, None]
    n = pos.shape[0]
    F = np.zeros((n, 3))
    np.add.at(F, i, fvec)
    np.add.at(F, j, -fvec)
    solute = (i < nd) | (j < nd)
    interacting = lj_on | c_on
    extra = {}
    if mol_start is not None:
        starts = np.asarray(mol_start, np.int64)
        nm = len(starts)
        M = np.zeros((nm, nm))
        ss = (i < nd) & (j < nd)
        mi = np.searchsorted(starts, i[ss], side="right") - 1
        mj = np.searchsorted(starts, j[ss], side="right") - 1
        e = (e_lj + e_c)[ss]
        np.add.at(M, (mi, mj), e)
        off = mi != mj
        np.add.at(M, (mj[off], mi[off]), e[off])
        extra["mol_energies"] = M
    return {
        **extra,
        "forces": F,
        "e_lj": float(e_lj.sum()), "e_coul": float(e_c.sum()),
        "e_lj_solute": float(e_lj[solute].sum()), "e_coul_solute": float(e_c[solute].sum()),
        "virial": float((fmag * r).sum()),
        "pair_count": int(interacting.sum()),
        "pair_count_lj": int(lj_on.sum()), "pair_count_coul": int(c_on.sum()),
        "dh_dl": float(dh.sum()),
    }


def ref_pairs_fp64(system, pairs, cutoff, alpha, lj_cutoff=np.inf, lambda_alch=0.0,
                   coulomb_disabled=False, lj_disabled=False, energy_all=False):
    """FP64 evaluation of a flattened pair list with the reference kernel's
    rules: skip r² < 1e-16, LJ only where calc_lj (no cutoff unless given),
    Coulomb zero at r ≥ cutoff, 1-4 scaling, alchemical soft-core with the
    staged schedule, energy of pairs whose target is a solute atom.
    Forces returned in the per-class layout (f_dyn (nd,3), f_water (nw,4,3)
    in o, h0, h1, m order)."""
    nd, nw = system.n_dyn, system.n_water
    box = system.box.astype(F64)
    blocks = [system.pos_dyn, system.pos_o, system.pos_m, system.pos_h0, system.pos_h1]
    allpos = np.concatenate([b.astype(F64) for b in blocks]) if nd + nw else np.zeros((0, 3))
    off = lambda at, ws: np.where(at == 0, 0, nd + (ws.astype(np.int64) - 1) * nw)
    ti = off(pairs["atom_type_tgt"], pairs["water_site_tgt"]) + pairs["tgt"].astype(np.int64)
    si = off(pairs["atom_type_src"], pairs["water_site_src"]) + pairs["src"].astype(np.int64)
    d = min_image(allpos[ti] - allpos[si], box)
    r2 = np.einsum("ij,ij->i", d, d)
    ok = r2 >= 1e-16
    r = np.sqrt(np.where(ok, r2, 1.0))
    sig = pairs["sigma"].astype(F64)
    eps = pairs["eps"].astype(F64)
    qq = pairs["q_tgt"].astype(F64) * pairs["q_src"].astype(F64)
    alch = pairs["alch"].astype(bool) if pairs.get("alch") is not None else np.zeros(len(r), bool)
    lam = min(1.0, max(0.0, lambda_alch))
    if lam < 0.5:
        c_scale, c_ds, lj_lam, lj_dl = 1 - 2 * lam, -2.0, 0.0, 0.0
    else:
        c_scale, c_ds, lj_lam, lj_dl = 0.0, 0.0, 2 * lam - 1, 2.0
    lj_on = pairs["calc_lj"].astype(bool) & (not lj_disabled) & (r <= lj_cutoff) & ok
    e_lj, f_lj = lj_fp64(r, sig, eps)
    dh = np.zeros_like(r)
    if alch.any():
        e_sc, f_sc, dh_sc = soft_core_fp64(r, sig, eps, lj_lam)
        e_lj = np.where(alch, e_sc, e_lj)
        f_lj = np.where(alch, f_sc, f_lj)
        dh = np.where(alch, dh_sc * lj_dl, 0.0)
    e_lj = np.where(lj_on, e_lj, 0.0)
    f_lj = np.where(lj_on, f_lj, 0.0)
    dh = np.where(lj_on, dh, 0.0)
    c_on = pairs["calc_coul"].astype(bool) & (not coulomb_disabled) & (r < cutoff) & ok
    e_c, f_c = coul_erfc_fp64(r, qq, alpha)
    e_c = np.where(c_on, e_c, 0.0)
    f_c = np.where(c_on, f_c, 0.0)
    s14 = pairs["scale14"].astype(bool)
    e_lj = np.where(s14, 0.5 * e_lj, e_lj)
    f_lj = np.where(s14, 0.5 * f_lj, f_lj)
    dh = np.where(s14, 0.5 * dh, dh)
    e_c = np.where(s14, e_c / 1.2, e_c)
    f_c = np.where(s14, f_c / 1.2, f_c)
    cs = np.where(alch, c_scale, 1.0)
    fmag = f_lj + f_c * cs
    e = e_lj + e_c * cs
    dh = dh + np.where(alch, e_c * c_ds, 0.0)
    fvec = d * (fmag / r)[:, None]
    fvec[~ok] = 0
    n = allpos.shape[0]
    F = np.zeros((n, 3))
    np.add.at(F, ti, fvec)
    sym = pairs["symmetric"].astype(bool)
    np.add.at(F, si[sym], -fvec[sym])
    tgt_solute = pairs["atom_type_tgt"] == 0
    energy = float(e[ok & (tgt_solute | energy_all)].sum())
    virial = float(np.einsum("ij,ij->i", d, fvec)[ok].sum())
    f_dyn = F[:nd]
    blocksF = [F[nd + k * nw: nd + (k + 1) * nw] for k in range(4)]  # O, M, H0, H1
    f_water = np.stack([blocksF[0], blocksF[2], blocksF[3], blocksF[1]], axis=1)
    return f_dyn, f_water, virial, energy, float(dh[alch & ok].sum()) if alch.any() else 0.0


def soft_core_fp64(r, sigma, eps, lam, alpha=0.5, sigma_min=3.0):
    """Beutler soft-core (p = 1): E = (1−λ) V_LJ(r_sc), r_sc⁶ = r⁶ + α σ_sc⁶ λ.
    Returns (E, |F|, dE/dλ)."""
    ss6 = np.maximum(sigma, sigma_min) ** 6
    rsc6 = r ** 6 + alpha * ss6 * lam
    rsc = rsc6 ** (1 / 6)
    sr6 = (sigma / rsc) ** 6
    hard_e = 4 * eps * (sr6 * sr6 - sr6)
    hard_f = 24 * eps * (2 * sr6 * sr6 - sr6) / rsc
    scale = 1 - lam
    e = scale * hard_e
    f = scale * hard_f * (r / rsc) ** 5
    drsc = alpha * ss6 / (6 * rsc ** 5)
    dh = -hard_e - scale * hard_f * drsc
    z = eps == 0
    return np.where(z, 0, e), np.where(z, 0, f), np.where(z, 0, dh)


# --- B-splines and SPME ----------------------------------------------------------
def bspline(x, n):
    """Cardinal B-spline M_n(x) (support (0, n)) and dM_n/dx, vectorised."""
    x = np.asarray(x, F64)

    def M(xx, k):
        if k == 2:
            return np.where((xx >= 0) & (xx <= 2), 1 - np.abs(xx - 1), 0.0)
        return (xx * M(xx, k - 1) + (k - xx) * M(xx - 1, k - 1)) / (k - 1)

    return M(x, n), M(x, n - 1) - M(x - 1, n - 1)


def bspline_moduli(K, n):
    """|b(m)|² of Essmann eq. 4.4 with the usual fix at zeros of the denominator."""
    mk = np.array([bspline(np.array([k + 1.0]), n)[0][0] for k in range(n - 1)])
    m = np.arange(K)
    s = (mk[None, :] * np.exp(2j * np.pi * np.outer(m, np.arange(n - 1)) / K)).sum(axis=1)
    mod = np.abs(s) ** 2
    for k in range(K):
        if mod[k] < 1e-7:
            mod[k] = 0.5 * (mod[(k - 1) % K] + mod[(k + 1) % K])
    return 1.0 / mod


def influence(K, box, beta, n):
    V = float(np.prod(box))
    f = [np.fft.fftfreq(K[a], d=1.0 / K[a]) / box[a] for a in range(3)]  # m'/L
    fx, fy, fz = np.meshgrid(*f, indexing="ij")
    msq = fx ** 2 + fy ** 2 + fz ** 2
    b = [bspline_moduli(K[a], n) for a in range(3)]
    B = b[0][:, None, None] * b[1][None, :, None] * b[2][None, None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        G = np.exp(-np.pi ** 2 * msq / beta ** 2) / (np.pi * V * msq) * B
    G[0, 0, 0] = 0.0
    return G, msq


def spme_fp64(pos, q, box, K, order, beta):
    """Returns (forces (n,3), E_rec, W_rec) without the self energy."""
    pos = np.asarray(pos, F64).reshape(-1, 3)
    q = np.asarray(q, F64)
    box = np.asarray(box, F64)
    K = tuple(int(k) for k in K)
    n = order
    u = np.mod(pos, box) / box * np.array(K)
    k0 = np.floor(u).astype(np.int64)
    w = u - k0
    j = np.arange(n)
    th, dth, idx = [], [], []
    for a in range(3):
        x = w[:, a:a + 1] + (n - 1 - j)[None, :]  # weight for grid k0 - n + 1 + j is M(w + n-1-j)
        m_, dm_ = bspline(x, n)
        th.append(m_)
        dth.append(dm_)
        idx.append(np.mod(k0[:, a:a + 1] - n + 1 + j[None, :], K[a]))
    Q = np.zeros(K)
    W = (q[:, None, None, None] * th[0][:, :, None, None] * th[1][:, None, :, None]
         * th[2][:, None, None, :])
    ix = np.broadcast_to(idx[0][:, :, None, None], W.shape)
    iy = np.broadcast_to(idx[1][:, None, :, None], W.shape)
    iz = np.broadcast_to(idx[2][:, None, None, :], W.shape)
    np.add.at(Q, (ix.ravel(), iy.ravel(), iz.ravel()), W.ravel())
    G, msq = influence(K, box, beta, n)
    FQ = np.fft.fftn(Q)
    S2 = np.abs(FQ) ** 2
    E = 0.5 * float((G * S2).sum())
    Wv = 0.5 * float((G * S2 * (1 - 2 * np.pi ** 2 * msq / beta ** 2)).sum())
    phi = np.real(np.fft.ifftn(G * FQ)) * np.prod(K)
    P = phi[ix, iy, iz]
    sc = np.array(K) / box
    fx = -(q * sc[0] * (P * dth[0][:, :, None, None] * th[1][:, None, :, None] * th[2][:, None, None, :]).sum(axis=(1, 2, 3)))
    fy = -(q * sc[1] * (P * th[0][:, :, None, None] * dth[1][:, None, :, None] * th[2][:, None, None, :]).sum(axis=(1, 2, 3)))
    fz = -(q * sc[2] * (P * th[0][:, :, None, None] * th[1][:, None, :, None] * dth[2][:, None, None, :]).sum(axis=(1, 2, 3)))
    return np.stack([fx, fy, fz], axis=1), E, Wv


def spme_grid_fp64(pos, q, box, K, order):
    """The spread charge grid alone (for fixed-point grid checks)."""
    pos = np.asarray(pos, F64).reshape(-1, 3)
    q = np.asarray(q, F64)
    box = np.asarray(box, F64)
    n = order
    u = np.mod(pos, box) / box * np.array(K)
    k0 = np.floor(u).astype(np.int64)
    w = u - k0
    j = np.arange(n)
    Q = np.zeros(K)
    th, idx = [], []
    for a in range(3):
        m_, _ = bspline(w[:, a:a + 1] + (n - 1 - j)[None, :], n)
        th.append(m_)
        idx.append(np.mod(k0[:, a:a + 1] - n + 1 + j[None, :], K[a]))
    W = (q[:, None, None, None] * th[0][:, :, None, None] * th[1][:, None, :, None]
         * th[2][:, None, None, :])
    ix = np.broadcast_to(idx[0][:, :, None, None], W.shape)
    iy = np.broadcast_to(idx[1][:, None, :, None], W.shape)
    iz = np.broadcast_to(idx[2][:, None, None, :], W.shape)
    np.add.at(Q, (ix.ravel(), iy.ravel(), iz.ravel()), W.ravel())
    return Q


def ewald_recip_fp64(pos, q, box, beta, mmax):
    """Direct reciprocal Ewald sum over |m_a| ≤ mmax: (forces, E_rec, W_rec)."""
    pos = np.asarray(pos, F64).reshape(-1, 3)
    q = np.asarray(q, F64)
    box = np.asarray(box, F64)
    V = float(np.prod(box))
    rng = np.arange(-mmax, mmax + 1)
    mx, my, mz = np.meshgrid(rng, rng, rng, indexing="ij")
    m = np.stack([mx.ravel() / box[0], my.ravel() / box[1], mz.ravel() / box[2]], axis=1)
    msq = (m ** 2).sum(axis=1)
    nz = msq > 0
    m, msq = m[nz], msq[nz]
    f = np.exp(-np.pi ** 2 * msq / beta ** 2) / msq
    th = 2 * np.pi * pos @ m.T  # (n, M)
    ph = np.exp(1j * th)
    S = (q[:, None] * ph).sum(axis=0)
    S2 = np.abs(S) ** 2
    E = float((f * S2).sum()) / (2 * np.pi * V)
    Wv = float((f * S2 * (1 - 2 * np.pi ** 2 * msq / beta ** 2)).sum()) / (2 * np.pi * V)
    # F_j = (2 q_j / V) Σ_m f(m) m Im[S* e^{iθ_j}]
    im = np.imag(np.conj(S)[None, :] * ph)  # (n, M)
    F = (2 * q / V)[:, None] * ((f[None, :] * im) @ m)
    return F, E, Wv


def self_energy(q, beta):
    return -beta / math.sqrt(math.pi) * float(np.sum(np.asarray(q, F64) ** 2))


def corr_14_vacuum(pos_dyn, q_dyn, pairs_14, box):
    """Force-only 1-4 correction (1/1.2 − 1)·qi·qj/r² along r̂ (reference
    mode) and its virial r·ΔF."""
    pos = np.asarray(pos_dyn, F64)
    q = np.asarray(q_dyn, F64)
    box = np.asarray(box, F64)
    F = np.zeros_like(pos)
    if len(pairs_14) == 0:
        return F, 0.0
    i, j = pairs_14[:, 0], pairs_14[:, 1]
    d = min_image(pos[i] - pos[j], box)
    r = np.sqrt((d ** 2).sum(axis=1))
    df = d * ((1 / 1.2 - 1) * q[i] * q[j] / r ** 3)[:, None]
    np.add.at(F, i, df)
    np.add.at(F, j, -df)
    return F, float((d * df).sum())


def site_charges(system):
    _, q, _, _, _, _ = site_view(system)
    return q


def reference_energy_and_forces(system, rc, lj_rc, beta, K, order, mode="reference", alch=None,
                                lam=0.0):
    """Whole nonbonded evaluation in FP64: short range + corrections + SPME,
    forces in site order, the engine's energy breakdown.  With a decoupled
    molecule (alch, bool per solute atom) the SPME sum is evaluated three
    times, as the reference does (src/non_bonded.rs:709-742): all charges,
    the environment only, the molecule only; the cross term is scaled by the
    schedule's Coulomb factor and contributes its λ-derivative to dh_dl."""
    sr = pair_sum_fp64(system, rc, lj_rc, beta, alch=alch, lam=lam)
    pos, q, _, _, mol, _ = site_view(system)
    box = system.box.astype(F64)
    F = sr["forces"].copy()
    out = {"e_lj": sr["e_lj"], "e_coul": sr["e_coul"], "virial_sr": sr["virial"],
           "pair_count": sr["pair_count"], "dh_dl": sr["dh_dl"]}
    e_rec = w_rec = 0.0
    e_self = 0.0
    if K is not None:
        qi = np.nonzero(q != 0)[0]
        if alch is not None and np.any(alch):
            a_site = np.zeros(pos.shape[0], bool)
            a_site[:system.n_dyn] = np.asarray(alch, bool)
            q_env = np.where(a_site, 0.0, q)
            q_mol = np.where(a_site, q, 0.0)
            f_all, e_all, w_all = spme_fp64(pos[qi], q[qi], box, K, order, beta)
            f_env, e_env, w_env = spme_fp64(pos[qi], q_env[qi], box, K, order, beta)
            f_mol, e_mol, w_mol = spme_fp64(pos[qi], q_mol[qi], box, K, order, beta)
            cs, cds, _, _ = schedule(lam)
            f_r = f_env + f_mol + cs * (f_all - f_env - f_mol)
            e_cross = e_all - e_env - e_mol
            e_rec = e_env + e_mol + cs * e_cross
            w_rec = w_env + w_mol + cs * (w_all - w_env - w_mol)
            out["dh_dl"] += cds * e_cross
        else:
            f_r, e_rec, w_rec = spme_fp64(pos[qi], q[qi], box, K, order, beta)
        F[qi] += f_r
        e_self = self_energy(q, beta)
    w_corr = 0.0
    e_corr = 0.0
    nd = system.n_dyn
    if mode == "reference":
        if K is not None and nd and len(system.pairs_14):
            f14, w14 = corr_14_vacuum(system.pos_dyn, system.q_dyn, system.pairs_14.astype(np.int64), box)
            F[:nd] += f14
            w_corr += w14
    else:
        pl = []
        if nd and len(system.excl_12_13):
            pl.append((system.excl_12_13.astype(np.int64), 1.0))
        if nd and len(system.pairs_14):
            pl.append((system.pairs_14.astype(np.int64), 1.0 - 1.0 / 1.2))
        nw = system.n_water
        if nw:
            b = nd + 4 * np.arange(nw)
            wp = np.concatenate([np.stack([b + 1, b + 2], 1), np.stack([b + 1, b + 3], 1),
                                 np.stack([b + 2, b + 3], 1)])
            pl.append((wp, 1.0))
        for pr, fac in pl:
            i, j = pr[:, 0], pr[:, 1]
            d = min_image(pos[i] - pos[j], box)
            r = np.sqrt((d ** 2).sum(axis=1))
            e, f = erf_pair_fp64(r, q[i] * q[j], beta)
            fv = -fac * d * (f / r)[:, None]
            np.add.at(F, i, fv)
            np.add.at(F, j, -fv)
            e_corr += -fac * float(e.sum())
            w_corr += float((d * fv).sum())
        if K is not None:
            V = float(np.prod(box))
            e_self += -np.pi * float(q.sum()) ** 2 / (2 * V * beta ** 2)
    out.update(forces=F, e_recip=e_rec, e_self=e_self, e_corr=e_corr, virial_recip=w_rec + w_corr,
               virial_kspace=w_rec, virial_corr=w_corr)
    out["e_total"] = out["e_lj"] + out["e_coul"] + e_rec + e_self + e_corr
    out["e_reference"] = sr["e_lj_solute"] + sr["e_coul_solute"] + e_rec + e_self
    return out
