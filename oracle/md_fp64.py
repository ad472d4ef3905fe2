"""FP64 restatement of the reference's MD step (test infrastructure only):
the checker for dynamics_amd's device-resident step, never a product path.

Follows:
  Langevin-middle step        src/integrate.rs:100-190
  velocity Verlet (γ = 0)     src/integrate.rs:191-260
  kick / drift                src/integrate.rs:478-590 (MAX_ACCEL :26)
  rigid-water drift           src/solvent/opc_settle.rs:107-198
  M-site force projection     src/solvent/mod.rs:397-407, C_H/C_O :89-90
  Langevin velocity update    src/thermostat.rs:308-362
  SHAKE / RATTLE              src/bonded.rs:141-304
  bonded terms                src/bonded_forces.rs:11-200 (harmonic bond
                              and angle; periodic dihedral written with the
                              IUPAC torsion angle)
  pressure, C-rescale         src/barostat.rs:324-351 (scale factor),
                              :419-431 (P = (2K + W)/3V), measured where
                              src/integrate.rs:172-192 / :241-310 do; with
                              pressure groups (each water, each constraint
                              cluster, each other atom): K of the groups'
                              centres of mass, W = W_atomic − Σ (r − R_g)·F
The Langevin noise uses the same counter-based deviates as the device
(SplitMix64 finalizer over seed, step and index; Box–Muller), evaluated in
float64.  Pinned by: the SplitMix64 known answer (tests/test_oracle_md.py),
−∇U finite differences for the bonded forces, rigid-body invariants of the
water drift (geometry, linear and angular momentum), constraint residuals,
and energy conservation of the γ = 0 step.  The group virial is pinned by
W = −dU/dλ, U's derivative when every group's centre of mass and the box
scale by λ (tests/test_oracle_pressure.py).
"""

from __future__ import annotations

import copy
import dataclasses

import numpy as np

from . import fp64

KCAL_TO_NATIVE = 418.4
BAR_PER_KCAL_MOL_A3 = 69476.95457055373   # src/barostat.rs:22
KB_BAR_A3 = 138.0649                      # src/barostat.rs:26
BAROSTAT_INDEX = 0xB0B0000000000000       # the barostat's deviate in the step's stream
KB = 0.83144626
M_O, M_H = 16.0, 1.008
M_W = M_O + 2 * M_H
O_EP_R = 0.15939833
RA = 0.5395199719801114
C_H = (O_EP_R / RA) / 2.0
C_O = 1.0 - 2.0 * C_H
MAX_ACCEL = 1e5
EPS_CONSTR = 1e-8
SHAKE_MAX_IT = 100
U64 = np.uint64
MASK = (1 << 64) - 1


def mix64(z):
    """SplitMix64 output for state z (the state advanced by the golden gamma
    first), on Python ints or uint64 arrays."""
    with np.errstate(over="ignore"):
        z = (np.asarray(z, U64) + U64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> U64(30))) * U64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> U64(27))) * U64(0x94D049BB133111EB)
        return z ^ (z >> U64(31))


def normal(seed, step, idx):
    """Standard normal deviates for indices idx at (seed, step)."""
    with np.errstate(over="ignore"):
        h = mix64(mix64(mix64(U64(seed)) + U64(step)) + np.asarray(idx, U64))
    u1 = ((h >> U64(40)).astype(np.float64) + 0.5) / 16777216.0
    u2 = (((h >> U64(16)) & U64(0xFFFFFF)).astype(np.float64) + 0.5) / 16777216.0
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def min_image(d, box):
    """src/barostat.rs:165-168: dv − round(dv/ext)·ext, round half away."""
    q = d / box
    return d - np.sign(q) * np.floor(np.abs(q) + 0.5) * box


def bonded_fp64(pos, ff, box, virial=False):
    """Forces (n_dyn, 3) and energies (bond, angle, dihedral) of the bonded
    terms; with virial=True also W = Σ_a r_a·F_a per term, positions taken
    relative to one of its atoms."""
    pos = np.asarray(pos, np.float64)
    box = np.asarray(box, np.float64)
    F = np.zeros_like(pos)
    e = [0.0, 0.0, 0.0]
    w = 0.0
    if len(ff.bonds):
        i, j = ff.bonds[:, 0].astype(np.int64), ff.bonds[:, 1].astype(np.int64)
        k, r0 = ff.bond_k_r0[:, 0].astype(np.float64), ff.bond_k_r0[:, 1].astype(np.float64)
        d = min_image(pos[j] - pos[i], box)
        r = np.linalg.norm(d, axis=1)
        fi = d * (2 * k * (r - r0) / np.maximum(r, 1e-8))[:, None]
        np.add.at(F, i, fi)
        np.add.at(F, j, -fi)
        e[0] = float((k * (r - r0) ** 2).sum())
        w += -float((d * fi).sum())
    if len(ff.angles):
        i, j, l = (ff.angles[:, c].astype(np.int64) for c in range(3))
        k, t0 = ff.angle_k_t0[:, 0].astype(np.float64), ff.angle_k_t0[:, 1].astype(np.float64)
        b0 = min_image(pos[i] - pos[j], box)
        b2 = min_image(pos[l] - pos[j], box)
        l0 = np.linalg.norm(b0, axis=1)
        l2 = np.linalg.norm(b2, axis=1)
        u0, u2 = b0 / l0[:, None], b2 / l2[:, None]
        ct = np.clip((u0 * u2).sum(1), -1, 1)
        th = np.arccos(ct)
        st = np.maximum(np.sqrt(1 - ct * ct), 1e-8)
        dedt = 2 * k * (th - t0)
        g0 = (u0 * ct[:, None] - u2) / (l0 * st)[:, None]
        g2 = (u2 * ct[:, None] - u0) / (l2 * st)[:, None]
        f0, f2 = -g0 * dedt[:, None], -g2 * dedt[:, None]
        np.add.at(F, i, f0)
        np.add.at(F, l, f2)
        np.add.at(F, j, -(f0 + f2))
        e[1] = float((k * (th - t0) ** 2).sum())
        w += float((b0 * f0).sum() + (b2 * f2).sum())
    if len(ff.dihedrals):
        i, j, k, l = (ff.dihedrals[:, c].astype(np.int64) for c in range(4))
        kk = ff.dihe_k_n_phase[:, 0].astype(np.float64)
        n = ff.dihe_k_n_phase[:, 1].astype(np.float64)
        g = ff.dihe_k_n_phase[:, 2].astype(np.float64)
        b1 = min_image(pos[j] - pos[i], box)
        b2 = min_image(pos[k] - pos[j], box)
        b3 = min_image(pos[l] - pos[k], box)
        n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
        b2l = np.linalg.norm(b2, axis=1)
        phi = np.arctan2(b2l * (b1 * n2).sum(1), (n1 * n2).sum(1))
        arg = n * phi - g
        dedphi = -kk * n * np.sin(arg)
        n1s, n2s, b2s = (n1 * n1).sum(1), (n2 * n2).sum(1), b2l ** 2
        gi = n1 * (-b2l / n1s)[:, None]
        gl = n2 * (b2l / n2s)[:, None]
        ta = ((b1 * b2).sum(1) / b2s)[:, None]
        tb = ((b3 * b2).sum(1) / b2s)[:, None]
        gj = -gi * (1 + ta) + gl * tb
        gk = gi * ta - gl * (1 + tb)
        for idx, gv in ((i, gi), (j, gj), (k, gk), (l, gl)):
            np.add.at(F, idx, -gv * dedphi[:, None])
        e[2] = float((kk * (1 + np.cos(arg))).sum())
        fi_, fk_, fl_ = (-gv * dedphi[:, None] for gv in (gi, gk, gl))
        w += float((-b1 * fi_).sum() + (b2 * fk_).sum() + ((b2 + b3) * fl_).sum())
    if virial:
        return F, e, w
    return F, e


@dataclasses.dataclass
class MdState:
    x_dyn: np.ndarray   # (n_dyn, 3)
    v_dyn: np.ndarray
    x_w: np.ndarray     # (n_water, 4, 3) O, M, H0, H1
    v_w: np.ndarray     # (n_water, 3, 3) O, H0, H1
    step: int = 0


def _system_at(system, st):
    s = copy.copy(system)
    s.pos_dyn = st.x_dyn
    s.pos_o, s.pos_m, s.pos_h0, s.pos_h1 = (st.x_w[:, k] for k in range(4))
    return s


class Model:
    """The forces and parameters of one MD setup (FP64)."""

    def __init__(self, system, ff, rc, beta, K, order=4, mode="reference", dt=0.002,
                 temperature=300.0, gamma=1.0, shake_tol=1e-5, seed=0, nstpcouple=0,
                 pressure_target=1.0, tau_p=5.0, compressibility=4.5e-5):
        self.system, self.ff = system, ff
        self.rc, self.beta, self.K, self.order, self.mode = rc, beta, K, order, mode
        self.dt, self.T, self.gamma, self.tol, self.seed = dt, temperature, gamma, shake_tol, seed
        self.box = system.box.astype(np.float64)
        self.mass = ff.mass.astype(np.float64)
        self.nd, self.nw = system.n_dyn, system.n_water
        # constraint clusters (connected components), constraints in input order
        parent = list(range(self.nd))

        def find(a):
            while parent[a] != a:
                parent[a] = parent[parent[a]]
                a = parent[a]
            return a

        for a, b in ff.constraints:
            parent[find(int(a))] = find(int(b))
        groups = {}
        for k, (a, b) in enumerate(ff.constraints):
            groups.setdefault(find(int(a)), []).append(k)
        self.clusters = list(groups.values())
        # pressure groups of the solute: atoms of each constraint cluster, then
        # every atom in no constraint
        in_cons = np.zeros(self.nd, bool)
        if len(ff.constraints):
            in_cons[np.asarray(ff.constraints, np.int64).reshape(-1)] = True
        comp = {}
        for a in range(self.nd):
            if in_cons[a]:
                comp.setdefault(find(a), []).append(a)
        order_roots = []
        for k in range(len(ff.constraints)):
            r = find(int(ff.constraints[k][0]))
            if r not in order_roots:
                order_roots.append(r)
        self.pgroups = [comp[r] for r in order_roots] + [[a] for a in range(self.nd) if not in_cons[a]]
        self.nst, self.p0, self.tau_p, self.kappa = nstpcouple, pressure_target, tau_p, compressibility
        self.f = None
        self.energy = None
        self.pressure_log = []

    def forces(self, st):
        s = _system_at(self.system, st)
        s.box = self.box.copy()
        r = fp64.reference_energy_and_forces(s, self.rc, self.rc, self.beta, self.K, self.order,
                                             mode=self.mode)
        F = r["forces"].copy()
        fb, eb, wb = bonded_fp64(st.x_dyn, self.ff, self.box, virial=True)
        F[:self.nd] += fb
        self.f = F
        self.w_atomic = r["virial_sr"] + r["virial_recip"] + wb
        self.energy = {"nonbonded": r["e_total"], "bond": eb[0], "angle": eb[1], "dihedral": eb[2]}
        return F

    def _groups(self, st):
        """(positions (n,3), velocities, masses, forces) per pressure group;
        water: O, H0, H1 with the M-site force projected as the kick does."""
        nd = self.nd
        out = []
        for g in self.pgroups:
            out.append((st.x_dyn[g], st.v_dyn[g], self.mass[g], self.f[g]))
        if self.nw:
            fw = self.f[nd:].reshape(self.nw, 4, 3)
            fo = fw[:, 0] + C_O * fw[:, 1]
            f0 = fw[:, 2] + C_H * fw[:, 1]
            f1 = fw[:, 3] + C_H * fw[:, 1]
            xw = np.stack([st.x_w[:, 0], st.x_w[:, 2], st.x_w[:, 3]], axis=1)
            fws = np.stack([fo, f0, f1], axis=1)
            out.append(("water", xw, st.v_w, np.array([M_O, M_H, M_H]), fws))
        return out

    def pressure(self, st):
        """(P bar, W group virial kcal/mol, K_com kcal/mol) from the last
        forces and the current velocities."""
        kin = 0.0
        cvir = 0.0
        for item in self._groups(st):
            if isinstance(item[0], str):
                _, xw, vw, mw, fws = item
                d = min_image(xw - xw[:, :1], self.box)
                M = mw.sum()
                R = (d * mw[None, :, None]).sum(1) / M
                V = (vw * mw[None, :, None]).sum(1) / M
                kin += 0.5 * M * float((V * V).sum())
                cvir += float(((d - R[:, None]) * fws).sum())
                continue
            x, v, m, f = item
            d = min_image(x - x[:1], self.box)
            M = m.sum()
            R = (d * m[:, None]).sum(0) / M
            V = (v * m[:, None]).sum(0) / M
            kin += 0.5 * M * float(V @ V)
            cvir += float(((d - R) * f).sum())
        kin /= KCAL_TO_NATIVE
        W = self.w_atomic - cvir
        P = (2 * kin + W) / (3 * float(np.prod(self.box))) * BAR_PER_KCAL_MOL_A3
        return P, W, kin

    def c_rescale_mu(self, P, step):
        """Barostat::scale_factor (src/barostat.rs:324-351) over nstpcouple·dt."""
        dt = self.dt * self.nst
        V = float(np.prod(self.box))
        det = (self.kappa / self.tau_p) * (P - self.p0) * dt
        sig = np.sqrt(2 * self.kappa * KB_BAR_A3 * self.T * dt / (self.tau_p * V))
        xi = float(normal(self.seed, step, BAROSTAT_INDEX))
        return float(np.exp(np.clip(det + sig * xi, -0.1, 0.1) / 3))

    def rescale(self, st, mu):
        """Every pressure group's centre of mass to μR (about the origin),
        velocities by 1/μ, the box by μ."""
        for g in self.pgroups:
            x = st.x_dyn[g]
            d = min_image(x - x[:1], self.box)
            m = self.mass[g]
            R = x[0] + (d * m[:, None]).sum(0) / m.sum()
            st.x_dyn[g] = x + (mu - 1) * R
        st.v_dyn = st.v_dyn / mu
        if self.nw:
            mw = np.array([M_O, M_H, M_H])
            xw = np.stack([st.x_w[:, 0], st.x_w[:, 2], st.x_w[:, 3]], axis=1)
            d = min_image(xw - xw[:, :1], self.box)
            R = xw[:, 0] + (d * mw[None, :, None]).sum(1) / mw.sum()
            st.x_w = st.x_w + (mu - 1) * R[:, None, :]
            st.v_w = st.v_w / mu
        self.box = self.box * mu

    def barostat(self, st):
        P, W, kin = self.pressure(st)
        mu = self.c_rescale_mu(P, st.step)
        self.pressure_log.append((st.step, P, W, kin, mu))
        if abs(mu - 1) >= 1e-12:
            self.rescale(st, mu)

    def kick(self, st, h):
        F = self.f
        nd, nw = self.nd, self.nw
        if nd:
            a = F[:nd] * (KCAL_TO_NATIVE / self.mass)[:, None]
            an = np.linalg.norm(a, axis=1)
            big = an > MAX_ACCEL
            a[big] *= (MAX_ACCEL / an[big])[:, None]
            st.v_dyn = st.v_dyn + a * h
        if nw:
            fw = F[nd:].reshape(nw, 4, 3)
            fm = fw[:, 1]
            fo = fw[:, 0] + C_O * fm
            f0 = fw[:, 2] + C_H * fm
            f1 = fw[:, 3] + C_H * fm
            st.v_w = st.v_w + np.stack([fo * (KCAL_TO_NATIVE / M_O), f0 * (KCAL_TO_NATIVE / M_H),
                                        f1 * (KCAL_TO_NATIVE / M_H)], axis=1) * h

    def drift(self, st, h):
        if self.nd:
            st.x_dyn = st.x_dyn + st.v_dyn * h
        if self.nw:
            st.x_w, st.v_w = rigid_drift(st.x_w, st.v_w, h, self.box)

    def shake(self, st):
        x = st.x_dyn
        m = self.mass
        for cl in self.clusters:
            for _ in range(SHAKE_MAX_IT):
                mc = 0.0
                for k in cl:
                    i, j = (int(v) for v in self.ff.constraints[k])
                    r0 = float(self.ff.constraint_r0[k])
                    d = x[j] - x[i]
                    d2 = float(d @ d)
                    lam = (d2 - r0 * r0) / (2 * (1 / m[i] + 1 / m[j]) * max(d2, EPS_CONSTR))
                    corr = d * lam
                    x[i] = x[i] + corr / m[i]
                    x[j] = x[j] - corr / m[j]
                    mc = max(mc, float(np.linalg.norm(corr)))
                if mc < self.tol:
                    break

    def rattle(self, st):
        x, v, m = st.x_dyn, st.v_dyn, self.mass
        for cl in self.clusters:
            for k in cl:
                i, j = (int(q) for q in self.ff.constraints[k])
                r = x[j] - x[i]
                dv = v[j] - v[i]
                lam = float(dv @ r) / ((1 / m[i] + 1 / m[j]) * max(float(r @ r), EPS_CONSTR))
                v[i] = v[i] + r * lam / m[i]
                v[j] = v[j] - r * lam / m[j]

    def langevin(self, st):
        c = np.exp(-self.gamma * self.dt)
        num = KB * self.T * max(0.0, 1 - c * c)
        nd, nw = self.nd, self.nw
        if nd:
            idx = 3 * np.arange(nd)[:, None] + np.arange(3)[None, :]
            z = normal(self.seed, st.step, idx)
            st.v_dyn = c * st.v_dyn + np.sqrt(num / self.mass)[:, None] * z
        if nw:
            base = nd + 4 * np.arange(nw)
            sites = np.stack([base, base + 2, base + 3], axis=1)
            z = normal(self.seed, st.step, 3 * sites[:, :, None] + np.arange(3))
            sig = np.array([np.sqrt(num / M_O), np.sqrt(num / M_H), np.sqrt(num / M_H)])
            st.v_w = c * st.v_w + sig[None, :, None] * z

    def step(self, st):
        """One step; the state is updated in place and returned."""
        st.x_dyn = np.array(st.x_dyn, np.float64)
        st.v_dyn = np.array(st.v_dyn, np.float64)
        st.x_w = np.array(st.x_w, np.float64)
        st.v_w = np.array(st.v_w, np.float64)
        if self.f is None:
            self.forces(st)
        h = 0.5 * self.dt
        if self.gamma > 0:
            self.kick(st, h)
            self.drift(st, h)
            self.shake(st)
            self.rattle(st)
            self.langevin(st)
            self.rattle(st)
            self.drift(st, h)
            self.shake(st)
        else:
            self.kick(st, h)
            self.drift(st, self.dt)
            self.shake(st)
            self.rattle(st)
        couple = self.nst > 0 and (st.step + 1) % self.nst == 0
        self.forces(st)
        if couple and self.gamma > 0:
            self.barostat(st)
        self.kick(st, h)
        self.rattle(st)
        if couple and self.gamma <= 0:
            self.barostat(st)
        st.step += 1
        return st


def kinetic(st, mass):
    ke = 0.5 * (mass[:, None] * st.v_dyn ** 2).sum() if len(mass) else 0.0
    mw = np.array([M_O, M_H, M_H])
    ke += 0.5 * (mw[None, :, None] * st.v_w ** 2).sum()
    return ke / KCAL_TO_NATIVE


def rigid_drift(xw, vw, dt, box):
    """Rigid drift of every water (vectorized over molecules)."""
    o = xw[:, 0]
    h0 = o + min_image(xw[:, 2] - o, box)
    h1 = o + min_image(xw[:, 3] - o, box)
    vo, vh0, vh1 = vw[:, 0], vw[:, 1], vw[:, 2]
    rc = (o * M_O + h0 * M_H + h1 * M_H) / M_W
    vc = (vo * M_O + vh0 * M_H + vh1 * M_H) / M_W
    rs = [o - rc, h0 - rc, h1 - rc]
    us = [vo - vc, vh0 - vc, vh1 - vc]
    ms = [M_O, M_H, M_H]
    L = sum(np.cross(r, u) * m for r, u, m in zip(rs, us, ms))
    I = np.zeros((len(o), 3, 3))
    for r, m in zip(rs, ms):
        r2 = (r * r).sum(1)
        I += m * (r2[:, None, None] * np.eye(3) - r[:, :, None] * r[:, None, :])
    w = np.linalg.solve(I, L[:, :, None])[:, :, 0]
    wd = w * dt
    th = np.linalg.norm(wd, axis=1)

    def rot(r):
        n = wd / np.where(th > 0, th, 1.0)[:, None]
        c, s = np.cos(th)[:, None], np.sin(th)[:, None]
        return r * c + np.cross(n, r) * s + n * ((n * r).sum(1))[:, None] * (1 - c)

    r2 = [rot(r) for r in rs]
    new = [rc + vc * dt + r for r in r2]
    wo = np.mod(new[0], box)
    wh0 = wo + min_image(new[1] - wo, box)
    wh1 = wo + min_image(new[2] - wo, box)
    bis = (wh0 - wo) + (wh1 - wo)
    wm = wo + bis / np.linalg.norm(bis, axis=1)[:, None] * O_EP_R
    x_out = np.stack([wo, wm, wh0, wh1], axis=1)
    v_out = np.stack([vc + np.cross(w, r) for r in r2], axis=1)
    return x_out, v_out
