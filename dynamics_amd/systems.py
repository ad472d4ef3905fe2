"""Synthetic input systems for the nonbonded path: OPC water boxes and a
solvated random-walk chain, generated from a seeded, language-independent
splitmix64 stream (no numpy RNG), so the same seed gives the same f32 inputs on
every machine.

Water geometry and parameters follow the reference's OPC model
(src/solvent/mod.rs:58-93 constants, WaterMolOpc::new :337-392), charges are
stored pre-multiplied by CHARGE_UNIT_SCALER (src/non_bonded.rs:63), LJ pairs
combine by Lorentz–Berthelot (src/non_bonded.rs:1081-1086), exclusions follow
setup_nonbonded_exclusion_scale_flags (src/prep.rs:114-143): 1-2 and 1-3
excluded, 1-4 scaled, a pair that is both 1-3 and 1-4 is treated as 1-4.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

F32 = np.float32

# --- constants as the reference evaluates them in f32 -----------------------
CHARGE_UNIT_SCALER = F32(18.2223)
O_EP_R = F32(0.15939833)
O_H_R = F32(0.87243313)
H_O_H_THETA = F32(1.808161105066)
_SIGMA_FACTOR = F32(F32(2.0) / F32(1.122462048309373))
O_RSTAR = F32(1.777167268)
O_SIGMA = F32(O_RSTAR * _SIGMA_FACTOR)
O_EPS = F32(0.2128008130)
Q_H = F32(F32(0.6791) * CHARGE_UNIT_SCALER)
Q_EP = F32(F32(-2.0) * Q_H)
SCALE_LJ_14 = F32(0.5)
SCALE_COUL_14 = F32(F32(1.0) / F32(1.2))

# water site codes as the reference encodes them (src/solvent/mod.rs:256-261)
SITE_O, SITE_M, SITE_H0, SITE_H1 = 1, 2, 3, 4


# --- splitmix64 -----------------------------------------------------------------
_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based splitmix64 (Steele, Lea, Flood 2014). Draw k values at
    once; the stream is identical to drawing them one at a time."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + idx * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        """Doubles in [0, 1) from the top 53 bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """Box–Muller on pairs of uniforms."""
        m = (n + 1) // 2
        u = self.uniform(2 * m)
        u1 = 1.0 - u[0::2]
        u2 = u[1::2]
        rad = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2)])
        return z[:n]

    def permutation(self, n: int) -> np.ndarray:
        """Fisher–Yates with uniform draws."""
        p = np.arange(n)
        u = self.uniform(n)
        for i in range(n - 1, 0, -1):
            j = int(u[i] * (i + 1))
            p[i], p[j] = p[j], p[i]
        return p


def random_quaternions(rng: SplitMix64, n: int) -> np.ndarray:
    """Uniform unit quaternions (w, x, y, z), Shoemake's method."""
    u = rng.uniform(3 * n).reshape(n, 3)
    u1, u2, u3 = u[:, 0], u[:, 1], u[:, 2]
    a = np.sqrt(1 - u1)
    b = np.sqrt(u1)
    x = a * np.sin(2 * np.pi * u2)
    y = a * np.cos(2 * np.pi * u2)
    z = b * np.sin(2 * np.pi * u3)
    w = b * np.cos(2 * np.pi * u3)
    return np.stack([w, x, y, z], axis=1)


def _rotate(q: np.ndarray, v: np.ndarray) -> np.ndarray:
    w = q[:, 0:1]
    u = q[:, 1:4]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def opc_sites(o_pos: np.ndarray, quats: np.ndarray):
    """H0, H1 and M positions of OPC waters from O positions and orientations,
    as WaterMolOpc::new places them (src/solvent/mod.rs:337-372): bisector along
    the body z axis, hydrogens in the body x-z plane, M on the bisector."""
    n = o_pos.shape[0]
    z_loc = _rotate(quats, np.tile([0.0, 0.0, 1.0], (n, 1)))
    e_loc = _rotate(quats, np.tile([1.0, 0.0, 0.0], (n, 1)))
    half = float(H_O_H_THETA) * 0.5
    h0_dir = z_loc * math.cos(half) + e_loc * math.sin(half)
    h1_dir = z_loc * math.cos(half) - e_loc * math.sin(half)
    h0_dir /= np.linalg.norm(h0_dir, axis=1, keepdims=True)
    h1_dir /= np.linalg.norm(h1_dir, axis=1, keepdims=True)
    o = o_pos.astype(np.float64)
    h0 = o + h0_dir * float(O_H_R)
    h1 = o + h1_dir * float(O_H_R)
    bis = (h0 - o) + (h1 - o)
    bis /= np.linalg.norm(bis, axis=1, keepdims=True)
    m = o + bis * float(O_EP_R)
    return h0, h1, m


@dataclasses.dataclass
class System:
    """A periodic orthorhombic system in the reference's data model: solute
    ("dynamic") atoms plus rigid 4-site OPC waters, all f32."""

    box: np.ndarray                 # (3,) f32 edge lengths; box spans [0, L)
    pos_dyn: np.ndarray             # (n_dyn, 3) f32
    q_dyn: np.ndarray               # (n_dyn,) f32, Amber-scaled
    sigma_dyn: np.ndarray           # (n_dyn,) f32
    eps_dyn: np.ndarray             # (n_dyn,) f32
    type_dyn: np.ndarray            # (n_dyn,) int32 LJ type index
    lj_types: np.ndarray            # (n_types, 2) f32 (sigma, eps)
    pos_o: np.ndarray               # (n_water, 3) f32
    pos_m: np.ndarray
    pos_h0: np.ndarray
    pos_h1: np.ndarray
    bonds: np.ndarray               # (n_bonds, 2) int32 solute indices
    excl_12_13: np.ndarray          # (k, 2) int32, i < j
    pairs_14: np.ndarray            # (k, 2) int32, i < j
    mol_start: np.ndarray           # molecule start indices over solute atoms
    name: str = ""
    n_heavy: int = 0                # chain systems: heavy atoms come first, then H

    @property
    def n_dyn(self) -> int:
        return int(self.pos_dyn.shape[0])

    @property
    def n_water(self) -> int:
        return int(self.pos_o.shape[0])

    @property
    def n_sites(self) -> int:
        return self.n_dyn + 4 * self.n_water

    def site_arrays(self):
        """All sites in one array: solute first, then per water O, M, H0, H1
        (molecule-major).  Returns (pos (n,3) f32, q (n,) f32, sigma, eps,
        water_mol (n,) int32 with -1 for solute, site code (n,) uint8)."""
        nd, nw = self.n_dyn, self.n_water
        # f32 unless a caller stored f64 positions (finite-difference checks)
        pos = np.empty((nd + 4 * nw, 3), np.result_type(F32, self.pos_dyn.dtype, self.pos_o.dtype))
        pos[:nd] = self.pos_dyn
        w = np.stack([self.pos_o, self.pos_m, self.pos_h0, self.pos_h1], axis=1)
        pos[nd:] = w.reshape(-1, 3)
        q = np.empty(nd + 4 * nw, F32)
        q[:nd] = self.q_dyn
        q[nd:] = np.tile(np.array([0.0, Q_EP, Q_H, Q_H], F32), nw)
        sig = np.empty(nd + 4 * nw, F32)
        eps = np.empty(nd + 4 * nw, F32)
        sig[:nd] = self.sigma_dyn
        eps[:nd] = self.eps_dyn
        sig[nd:] = np.tile(np.array([O_SIGMA, 0, 0, 0], F32), nw)
        eps[nd:] = np.tile(np.array([O_EPS, 0, 0, 0], F32), nw)
        mol = np.full(nd + 4 * nw, -1, np.int32)
        mol[nd:] = np.repeat(np.arange(nw, dtype=np.int32), 4)
        code = np.zeros(nd + 4 * nw, np.uint8)
        code[nd:] = np.tile(np.array([SITE_O, SITE_M, SITE_H0, SITE_H1], np.uint8), nw)
        return pos, q, sig, eps, mol, code


def _wrap(p: np.ndarray, box) -> np.ndarray:
    return np.mod(p, np.asarray(box, np.float64))


def water_box(n_water: int, box: float, lattice_n: int, seed: int = 0,
              jitter: float = 0.3, name: str = "") -> System:
    """n_water OPC waters: O at the first n_water points of a seeded
    permutation of a lattice_n³ lattice of spacing box/lattice_n, jittered
    by U(±jitter) per axis, uniformly random orientations."""
    rng = SplitMix64(seed)
    n_lat = lattice_n ** 3
    if n_water > n_lat:
        raise ValueError("lattice too small")
    perm = rng.permutation(n_lat)[:n_water]
    a = box / lattice_n
    ijk = np.stack([perm // (lattice_n * lattice_n), (perm // lattice_n) % lattice_n,
                    perm % lattice_n], axis=1).astype(np.float64)
    o = (ijk + 0.5) * a + (rng.uniform(3 * n_water).reshape(-1, 3) * 2 - 1) * jitter
    qs = random_quaternions(rng, n_water)
    h0, h1, m = opc_sites(o, qs)
    boxv = np.full(3, box, F32)
    return System(
        box=boxv, pos_dyn=np.zeros((0, 3), F32), q_dyn=np.zeros(0, F32),
        sigma_dyn=np.zeros(0, F32), eps_dyn=np.zeros(0, F32), type_dyn=np.zeros(0, np.int32),
        lj_types=np.zeros((0, 2), F32),
        pos_o=o.astype(F32), pos_m=m.astype(F32), pos_h0=h0.astype(F32), pos_h1=h1.astype(F32),
        bonds=np.zeros((0, 2), np.int32), excl_12_13=np.zeros((0, 2), np.int32),
        pairs_14=np.zeros((0, 2), np.int32), mol_start=np.zeros(0, np.int64), name=name)


def _min_image(d, box):
    return d - np.round(d / box) * box


def _chain(rng: SplitMix64, n_heavy: int, box: np.ndarray, bond: float,
           min_sep: float) -> np.ndarray:
    """Self-avoiding random walk with a fixed 111° bond angle and random
    torsions, periodic in the box."""
    pts = np.zeros((n_heavy, 3))
    pts[0] = box * 0.5
    d0 = rng.normal(3)
    d0 /= np.linalg.norm(d0)
    pts[1] = pts[0] + d0 * bond
    theta = math.radians(111.0)
    for i in range(2, n_heavy):
        prev = pts[i - 1] - pts[i - 2]
        prev /= np.linalg.norm(prev)
        ok = False
        for attempt in range(200):
            u = rng.uniform(2)
            phi = 2 * np.pi * u[0]
            # orthonormal frame around prev
            tmp = np.array([1.0, 0.0, 0.0]) if abs(prev[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
            e1 = np.cross(prev, tmp)
            e1 /= np.linalg.norm(e1)
            e2 = np.cross(prev, e1)
            th = theta if attempt < 100 else np.pi * u[1]
            dirv = (math.cos(np.pi - th) * prev + math.sin(np.pi - th) *
                    (math.cos(phi) * e1 + math.sin(phi) * e2))
            cand = pts[i - 1] + dirv * bond
            if i > 3:
                dd = _min_image(pts[: i - 3] - cand, box)
                if np.min(np.einsum("ij,ij->i", dd, dd)) < min_sep * min_sep:
                    continue
            ok = True
            break
        if not ok:
            pass  # accept the last candidate
        pts[i] = cand
    return pts


def solvated_chain(box: float, n_water: int, n_heavy: int, n_h: int, seed: int = 0,
                   lattice_n: int | None = None, clash: float = 2.4,
                   name: str = "") -> System:
    """A seeded random-walk chain of n_heavy heavy atoms (bond 1.53 Å) with
    n_h hydrogens (1.09 Å), 8 Lorentz–Berthelot LJ types (σ 1.9–3.8 Å,
    ε 0.015–0.21 kcal/mol), Gaussian charges (sd 0.35 e) shifted to zero net,
    solvated by n_water OPC waters on a jittered lattice with the sites near
    the chain removed."""
    rng = SplitMix64(seed)
    boxv = np.full(3, box, np.float64)
    heavy = _chain(rng, n_heavy, boxv, 1.53, 2.6)
    # hydrogens: each heavy gets one, n_h - n_heavy of them get a second.
    extra = n_h - n_heavy
    if extra < 0 or extra > n_heavy:
        raise ValueError("n_h must be in [n_heavy, 2 n_heavy]")
    two = np.zeros(n_heavy, bool)
    two[rng.permutation(n_heavy)[:extra]] = True
    hpos = []
    hparent = []
    from scipy.spatial import cKDTree  # local import: scipy only needed here
    heavy_tree = cKDTree(_wrap(heavy, boxv), boxsize=boxv)

    def clearance(cands, i):
        """Smallest distance from the candidate H positions to atoms they do
        not share a bond or angle with (heavy atoms beyond i ± 1, hydrogens
        of other heavy atoms)."""
        worst = np.inf
        for q in cands:
            for j in heavy_tree.query_ball_point(_wrap(q[None], boxv)[0], 3.0):
                if abs(j - i) > 1:
                    worst = min(worst, float(np.linalg.norm(_min_image(q - heavy[j], boxv))))
            if hpos:
                d = _min_image(np.asarray(hpos) - q, boxv)
                other = np.asarray(hparent) != i
                if other.any():
                    worst = min(worst, float(np.sqrt(np.einsum("ij,ij->i", d, d)[other].min())))
        return worst

    for i in range(n_heavy):
        a = heavy[i - 1] if i > 0 else heavy[i] - (heavy[i + 1] - heavy[i])
        c = heavy[i + 1] if i + 1 < n_heavy else heavy[i] - (heavy[i - 1] - heavy[i])
        b = _min_image(heavy[i] - a, boxv) + _min_image(heavy[i] - c, boxv)
        nb = np.linalg.norm(b)
        b = b / nb if nb > 1e-6 else np.array([0.0, 0.0, 1.0])
        n = np.cross(_min_image(a - heavy[i], boxv), _min_image(c - heavy[i], boxv))
        nn = np.linalg.norm(n)
        n = n / nn if nn > 1e-6 else np.array([1.0, 0.0, 0.0])
        ang = math.radians(54.75)
        m = np.cross(b, n)
        pick = two[i] or rng.uniform(1)[0] < 0.5
        # H directions at 54.75° from the bisector; rotated about it in 10°
        # steps until they clear other atoms by 2.2 Å (else the best found)
        best, best_c = None, -1.0
        for k in range(36):
            rot = math.radians(10.0 * k)
            nr = n * math.cos(rot) + m * math.sin(rot)
            d_plus = b * math.cos(ang) + nr * math.sin(ang)
            d_minus = b * math.cos(ang) - nr * math.sin(ang)
            cands = ([heavy[i] + 1.09 * d_plus, heavy[i] + 1.09 * d_minus] if two[i]
                     else [heavy[i] + 1.09 * (d_plus if pick else d_minus)])
            cl = clearance(cands, i)
            if cl > best_c:
                best, best_c = cands, cl
            if cl >= 2.2:
                break
        hpos += best
        hparent += [i] * len(best)
    hpos = np.array(hpos)
    # atom order: heavy atoms, then hydrogens
    pos_dyn = np.concatenate([heavy, hpos])
    n_dyn = pos_dyn.shape[0]
    bonds = [(i, i + 1) for i in range(n_heavy - 1)]
    bonds += [(p, n_heavy + k) for k, p in enumerate(hparent)]
    bonds = np.array(bonds, np.int32)
    # LJ types
    u = rng.uniform(16).reshape(8, 2)
    sig = 1.9 + 1.9 * u[:, 0]
    eps = 0.015 + 0.195 * u[:, 1]
    order = np.argsort(sig, kind="stable")
    lj_types = np.stack([sig[order], eps[order]], axis=1).astype(F32)
    t = np.empty(n_dyn, np.int32)
    uu = rng.uniform(n_dyn)
    t[:n_heavy] = 2 + (uu[:n_heavy] * 6).astype(np.int32)
    t[n_heavy:] = (uu[n_heavy:] * 2).astype(np.int32)
    # charges
    qe = rng.normal(n_dyn) * 0.35
    qe -= qe.mean()
    q = (qe.astype(F32) * CHARGE_UNIT_SCALER).astype(F32)
    # make the f32 net charge as close to zero as f32 allows
    q[-1] = F32(q[-1] - F32(np.sum(q.astype(np.float64))))
    excl, p14 = exclusion_sets(n_dyn, bonds)
    # waters on a jittered lattice, away from the solute
    if lattice_n is None:
        lattice_n = int(math.ceil((n_water * 1.6) ** (1 / 3)))
    a = box / lattice_n
    g = np.arange(lattice_n)
    ijk = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3).astype(np.float64)
    lat = (ijk + 0.5) * a
    keep = np.ones(lat.shape[0], bool)
    sol_w = _wrap(pos_dyn, boxv)
    from scipy.spatial import cKDTree  # local import: scipy only needed here
    tree = cKDTree(sol_w, boxsize=boxv)
    dmin, _ = tree.query(_wrap(lat, boxv), k=1)
    keep &= dmin > clash + 0.3
    cand = np.nonzero(keep)[0]
    if cand.size < n_water:
        raise ValueError(f"only {cand.size} lattice sites free for {n_water} waters")
    sel = cand[rng.permutation(cand.size)[:n_water]]
    o = lat[sel] + (rng.uniform(3 * n_water).reshape(-1, 3) * 2 - 1) * 0.3
    qs = random_quaternions(rng, n_water)
    h0, h1, m = opc_sites(o, qs)
    return System(
        box=boxv.astype(F32), pos_dyn=pos_dyn.astype(F32), q_dyn=q,
        sigma_dyn=lj_types[t, 0].copy(), eps_dyn=lj_types[t, 1].copy(), type_dyn=t,
        lj_types=lj_types,
        pos_o=o.astype(F32), pos_m=m.astype(F32), pos_h0=h0.astype(F32), pos_h1=h1.astype(F32),
        bonds=bonds, excl_12_13=excl, pairs_14=p14,
        mol_start=np.zeros(1, np.int64), name=name, n_heavy=n_heavy)


def exclusion_sets(n: int, bonds: np.ndarray):
    """1-2 and 1-3 excluded pairs and 1-4 scaled pairs (i < j) from a bond
    list, by the rules of src/prep.rs:114-143 (angles and proper dihedrals are
    enumerated from the bond graph here, as the topology's angle and dihedral
    keys would be)."""
    adj = [[] for _ in range(n)]
    for a, b in bonds:
        adj[a].append(int(b))
        adj[b].append(int(a))
    s12, s13, s14 = set(), set(), set()

    def key(i, j):
        return (i, j) if i < j else (j, i)

    for a, b in bonds:
        s12.add(key(int(a), int(b)))
    for c in range(n):
        nb = adj[c]
        for x in range(len(nb)):
            for y in range(x + 1, len(nb)):
                s13.add(key(nb[x], nb[y]))
    for b, c in bonds:
        b, c = int(b), int(c)
        for a in adj[b]:
            if a == c:
                continue
            for d in adj[c]:
                if d == b or d == a:
                    continue
                s14.add(key(a, d))
    excl = (s12 | s13) - s14
    to_arr = lambda s: (np.array(sorted(s), np.int32).reshape(-1, 2))
    return to_arr(excl), to_arr(s14)


# --- the named workloads ---------------------------------------------------------
def config0(seed: int = 0) -> System:
    """2,048 OPC waters in a 39.4 Å cube, O on a jittered 13³ lattice."""
    return water_box(2048, 39.4, 13, seed=seed, jitter=0.3, name="opc2048_box39.4")


def config1(seed: int = 0) -> System:
    """64 Å cube, 8,600 OPC waters around a 1,200-heavy-atom + 1,400-H chain."""
    return solvated_chain(64.0, 8600, 1200, 1400, seed=seed, lattice_n=24,
                          name="chain2600_opc8600_box64")


def config2(seed: int = 0) -> System:
    """69,900 OPC waters in a 128 Å cube (jittered 42³ lattice)."""
    return water_box(69900, 128.0, 42, seed=seed, jitter=0.3, name="opc69900_box128")


def config4(n: int = 128, seed: int = 4) -> list:
    """n solvated boxes with edges uniform in [24, 40] Å (≈0.0334 waters/Å³,
    ≈460–2,140 waters each) around 20–60-atom chains (half heavy atoms),
    seeded: the batched uneven-systems workload."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        L = float(rng.uniform(24.0, 40.0))
        n_atoms = int(rng.integers(20, 61))
        n_heavy = n_atoms // 2
        lat = int(np.ceil((0.0334 * L ** 3) ** (1.0 / 3.0))) + 1  # room for the chain
        n_water = int(round(0.0334 * L ** 3))
        out.append(solvated_chain(L, n_water, n_heavy, n_atoms - n_heavy, seed=1000 + k,
                                  lattice_n=lat, name=f"uneven{k}"))
    return out


def config3_lambdas(n: int = 16) -> list:
    """λ of the n alchemical windows of config 1: k/(n−1); the decoupled set is
    solute atoms [0, 40)."""
    return [k / max(n - 1, 1) for k in range(n)]


CONFIG3_ALCH = (0, 40)


def small_water(n_water: int = 512, box: float = 24.8, lattice_n: int = 8, seed: int = 1) -> System:
    return water_box(n_water, box, lattice_n, seed=seed, name=f"opc{n_water}_box{box}")


def small_solvated(seed: int = 2) -> System:
    """A 30-heavy + 40-H chain in 600 waters, 27 Å box (for parity tests)."""
    return solvated_chain(27.0, 600, 30, 40, seed=seed, lattice_n=10, name="chain70_opc600_box27")


# --- a synthetic bonded force field and initial velocities (for MD steps) -------
MASS_HEAVY = F32(12.011)
MASS_H = F32(1.008)
MASS_O = F32(16.0)
KB_NATIVE = 0.83144626  # k_B in Å²/ps²/K per amu


@dataclasses.dataclass
class ForceField:
    """Bonded terms and masses of the solute, in the reference's conventions
    (src/bonded_forces.rs: U = k(r − r0)², k(θ − θ0)², k(1 + cos(nφ − γ)))."""

    mass: np.ndarray              # (n_dyn,) amu
    bonds: np.ndarray             # (nb, 2)
    bond_k_r0: np.ndarray         # (nb, 2) kcal/mol/Å², Å
    angles: np.ndarray            # (na, 3), vertex in the middle
    angle_k_t0: np.ndarray        # (na, 2) kcal/mol/rad², rad
    dihedrals: np.ndarray         # (nd, 4)
    dihe_k_n_phase: np.ndarray    # (nd, 3) kcal/mol, periodicity, rad
    constraints: np.ndarray       # (nc, 2)
    constraint_r0: np.ndarray     # (nc,) Å


def _angle(a, b, c, box):
    u = _min_image(a - b, box)
    v = _min_image(c - b, box)
    return math.acos(max(-1.0, min(1.0, float(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v))))))


def chain_force_field(system: System, constrain_h: bool = True) -> ForceField:
    """Harmonic bonds (k 310, r0 as built) and angles (k 50, θ0 as built) and a
    threefold torsion (k 0.2) on every proper dihedral of the chain's bond
    graph; heavy atoms 12.011 amu, hydrogens 1.008 amu; bonds to hydrogen
    become constraints when constrain_h (and then carry no bond term)."""
    nd = system.n_dyn
    box = system.box.astype(np.float64)
    pos = system.pos_dyn.astype(np.float64)
    mass = np.full(nd, MASS_HEAVY, F32)
    mass[system.n_heavy:] = MASS_H
    is_h = np.zeros(nd, bool)
    is_h[system.n_heavy:] = True
    adj = [[] for _ in range(nd)]
    for a, b in system.bonds:
        adj[a].append(int(b))
        adj[b].append(int(a))
    bonds, bkr, cons, cr0 = [], [], [], []
    for a, b in system.bonds:
        a, b = int(a), int(b)
        r0 = float(np.linalg.norm(_min_image(pos[b] - pos[a], box)))
        if constrain_h and (is_h[a] or is_h[b]):
            cons.append((a, b))
            cr0.append(r0)
        else:
            bonds.append((a, b))
            bkr.append((310.0, r0))
    angles, akt = [], []
    for j in range(nd):
        nb = sorted(adj[j])
        for x in range(len(nb)):
            for y in range(x + 1, len(nb)):
                i, k = nb[x], nb[y]
                angles.append((i, j, k))
                akt.append((50.0, _angle(pos[i], pos[j], pos[k], box)))
    dih, dkp = [], []
    for j, k in system.bonds:
        j, k = int(j), int(k)
        for i in adj[j]:
            if i == k:
                continue
            for l in adj[k]:
                if l in (j, i):
                    continue
                dih.append((i, j, k, l))
                dkp.append((0.2, 3.0, 0.0))
    arr = lambda v, w, dt: np.asarray(v, dt).reshape(-1, w)
    return ForceField(mass=mass, bonds=arr(bonds, 2, np.uint32), bond_k_r0=arr(bkr, 2, F32),
                      angles=arr(angles, 3, np.uint32), angle_k_t0=arr(akt, 2, F32),
                      dihedrals=arr(dih, 4, np.uint32), dihe_k_n_phase=arr(dkp, 3, F32),
                      constraints=arr(cons, 2, np.uint32), constraint_r0=np.asarray(cr0, F32))


def water_force_field(system: System) -> ForceField:
    """No solute: an empty force field."""
    z = lambda w, dt: np.zeros((0, w), dt)
    return ForceField(mass=np.zeros(0, F32), bonds=z(2, np.uint32), bond_k_r0=z(2, F32),
                      angles=z(3, np.uint32), angle_k_t0=z(2, F32), dihedrals=z(4, np.uint32),
                      dihe_k_n_phase=z(3, F32), constraints=z(2, np.uint32),
                      constraint_r0=np.zeros(0, F32))


def initial_velocities(system: System, ff: ForceField, temperature: float = 300.0, seed: int = 7):
    """Maxwell–Boltzmann velocities: per solute atom, and per water one
    centre-of-mass velocity shared by O, H0 and H1 (no initial rotation).
    Returns (vel_dyn (n_dyn, 3), vel_water (n_water, 3, 3))."""
    rng = SplitMix64(seed)
    nd, nw = system.n_dyn, system.n_water
    v = np.zeros((nd, 3), F32)
    if nd:
        sd = np.sqrt(KB_NATIVE * temperature / ff.mass.astype(np.float64))
        v[:] = (rng.normal(3 * nd).reshape(nd, 3) * sd[:, None]).astype(F32)
    mw = float(MASS_O + 2 * MASS_H)
    vc = rng.normal(3 * nw).reshape(nw, 3) * math.sqrt(KB_NATIVE * temperature / mw)
    vw = np.repeat(vc[:, None, :], 3, axis=1).astype(F32)
    return v, vw


def water_positions(system: System) -> np.ndarray:
    """(n_water, 4, 3) positions in O, M, H0, H1 order."""
    return np.stack([system.pos_o, system.pos_m, system.pos_h0, system.pos_h1], axis=1).astype(F32)


def with_positions(system: System, pos_dyn, pos_water) -> System:
    """A copy of the system at other positions: solute (n_dyn, 3) and water
    (n_water, 4, 3) in O, M, H0, H1 order (as an MD state holds them)."""
    pw = np.asarray(pos_water, F32).reshape(-1, 4, 3)
    return dataclasses.replace(system, pos_dyn=np.ascontiguousarray(pos_dyn, F32).reshape(-1, 3),
                               pos_o=np.ascontiguousarray(pw[:, 0]),
                               pos_m=np.ascontiguousarray(pw[:, 1]),
                               pos_h0=np.ascontiguousarray(pw[:, 2]),
                               pos_h1=np.ascontiguousarray(pw[:, 3]))
