"""Pins of oracle/reference_path.py, the exact f32 restatement of the
reference's neighbor build (src/neighbors.rs:102-281) and pair flattening
(src/non_bonded.rs:541-677 with PerNeighborGpu::new's resolution,
src/gpu_interface.rs:242-411), and of the exclusion rule the systems use
(src/prep.rs:114-143).

* build_neighbors: membership equals an f64 distance test wherever the pair
  is not within rounding of the threshold; lists are symmetric, ascending,
  free of self pairs.
* setup_pairs: the counts its rules fix: solute–solute pairs j > i minus the
  excluded ones, 4 site pairs per listed (atom, water), 10 per listed water
  pair (O–O LJ + 9 charged pairs); 1-4 flagged exactly on the 1-4 set;
  Lorentz–Berthelot σ, ε; the OPC site charges.
* exclusion_sets: on a tree the 1-2/1-3 set is graph distance 1 or 2 and the
  1-4 set is distance 3; on a ring a pair that is both 1-3 and 1-4 is
  scaled, not excluded (the reference's "1-4 wins", src/prep.rs:139-142).
"""

from collections import deque

import numpy as np
import pytest

from dynamics_amd import systems
from oracle import reference_path as rp


@pytest.fixture(scope="module")
def solvated():
    s = systems.small_solvated()
    nb = rp.build_all_neighbors(s, skin=4.0, cutoff=10.0)
    return s, nb, rp.setup_pairs(s, nb)


def _lists(rpc):
    return rp.csr_rows(*rpc)


def test_build_neighbors_vs_f64(solvated):
    s, nb, _ = solvated
    thr = float(nb["thr"])
    box = s.box.astype(np.float64)
    o = nb["water_proxy"].astype(np.float64)
    d = o[:, None, :] - o[None, :, :]
    d -= box * np.round(d / box)
    r2 = (d ** 2).sum(-1)
    rows = _lists(nb["water_water"])
    member = np.zeros_like(r2, bool)
    for i, r in enumerate(rows):
        member[i, r] = True
    far = np.abs(r2 - thr) > 1e-3 * thr
    np.fill_diagonal(far, False)
    want = r2 < thr
    assert np.array_equal(member[far], want[far])
    assert not member.diagonal().any()
    assert np.array_equal(member, member.T)
    for r in rows:
        assert np.all(np.diff(r.astype(np.int64)) > 0)


def test_setup_pairs_counts(solvated):
    s, nb, P = solvated
    nd = s.n_dyn
    excl = set(map(tuple, s.excl_12_13.tolist()))
    ss_rows = _lists(nb["std_std"])
    n_ss = sum(1 for i in range(nd) for j in ss_rows[i] if j > i and (i, int(j)) not in excl)
    n_sw = len(nb["std_water"][1])
    n_ww = len(nb["water_water"][1]) // 2
    at_t, at_s = P["atom_type_tgt"], P["atom_type_src"]
    assert int(((at_t == 0) & (at_s == 0)).sum()) == n_ss
    assert int(((at_t == 0) & (at_s == 1)).sum()) == 4 * n_sw
    assert int(((at_t == 1) & (at_s == 1)).sum()) == 10 * n_ww
    assert len(P["tgt"]) == n_ss + 4 * n_sw + 10 * n_ww
    assert P["symmetric"].all()


def test_setup_pairs_exclusions_and_14(solvated):
    s, _, P = solvated
    ss = (P["atom_type_tgt"] == 0) & (P["atom_type_src"] == 0)
    t, u = P["tgt"][ss].astype(int), P["src"][ss].astype(int)
    assert np.all(u > t)
    keys = set(zip(t.tolist(), u.tolist()))
    excl = set(map(tuple, s.excl_12_13.tolist()))
    p14 = set(map(tuple, s.pairs_14.tolist()))
    assert not (keys & excl)
    flagged = {k for k, f in zip(zip(t.tolist(), u.tolist()), P["scale14"][ss]) if f}
    assert flagged == (p14 & keys)
    assert len(p14 & keys) > 0
    assert P["scale14"][~ss].sum() == 0


def test_setup_pairs_site_rules(solvated):
    s, _, P = solvated
    sw = (P["atom_type_tgt"] == 0) & (P["atom_type_src"] == 1)
    site = P["water_site_src"][sw]
    assert np.array_equal(P["calc_lj"][sw].astype(bool), site == 1)
    assert np.array_equal(P["calc_coul"][sw].astype(bool), site != 1)
    ww = (P["atom_type_tgt"] == 1) & (P["atom_type_src"] == 1)
    n_ww = int(ww.sum()) // 10
    assert int(P["calc_lj"][ww].sum()) == n_ww
    assert int(P["calc_coul"][ww].sum()) == 9 * n_ww
    assert not (P["calc_lj"][ww] & P["calc_coul"][ww]).any()
    # charges of the water sites (O has none); Lorentz–Berthelot σ, ε
    qsite = {1: 0.0, 2: systems.Q_EP, 3: systems.Q_H, 4: systems.Q_H}
    np.testing.assert_array_equal(P["q_src"][sw], np.array([qsite[int(k)] for k in site], np.float32))
    i = P["tgt"][sw].astype(int)
    sig = (np.float32(0.5) * (s.sigma_dyn[i] + np.float32(systems.O_SIGMA))).astype(np.float32)
    np.testing.assert_array_equal(P["sigma"][sw], sig)
    eps = np.sqrt((s.eps_dyn[i] * np.float32(systems.O_EPS)).astype(np.float32)).astype(np.float32)
    np.testing.assert_array_equal(P["eps"][sw], eps)


def _graph_dist(n, bonds):
    adj = [[] for _ in range(n)]
    for a, b in bonds:
        adj[a].append(int(b))
        adj[b].append(int(a))
    D = np.full((n, n), 99, np.int64)
    for s0 in range(n):
        D[s0, s0] = 0
        dq = deque([s0])
        while dq:
            x = dq.popleft()
            for y in adj[x]:
                if D[s0, y] == 99:
                    D[s0, y] = D[s0, x] + 1
                    dq.append(y)
    return D


def test_exclusion_sets_tree_is_graph_distance():
    s = systems.small_solvated()
    n = s.n_dyn
    D = _graph_dist(n, s.bonds)
    excl, p14 = systems.exclusion_sets(n, s.bonds)
    iu = np.triu_indices(n, 1)
    want_ex = {(int(a), int(b)) for a, b in zip(*iu) if D[a, b] in (1, 2)}
    want_14 = {(int(a), int(b)) for a, b in zip(*iu) if D[a, b] == 3}
    assert set(map(tuple, excl.tolist())) == want_ex
    assert set(map(tuple, p14.tolist())) == want_14


def test_exclusion_sets_ring_14_wins():
    # 5-ring 0-1-2-3-4-0: (0, 2) is 1-3 through 1 and 1-4 through 4-3
    bonds = np.array([[0, 1], [1, 2], [2, 3], [3, 4], [4, 0]])
    excl, p14 = systems.exclusion_sets(5, bonds)
    ex = set(map(tuple, excl.tolist()))
    s14 = set(map(tuple, p14.tolist()))
    assert (0, 2) in s14 and (0, 2) not in ex
    assert (0, 1) in ex and (0, 1) not in s14
    # in a 5-ring every non-bonded pair is both 1-3 and 1-4
    assert s14 == {(0, 2), (0, 3), (1, 3), (1, 4), (2, 4)}
    assert ex == {(0, 1), (1, 2), (2, 3), (3, 4), (0, 4)}
