"""The host restatement of the engine's pair list, checked against brute
force: each site sits in exactly one slot; every interacting site pair
within the cutoff is listed both ways; every listed j slot is within the
list radius of a site of the row's cluster; class segments are ordered and
padded to groups of 8; the solute–solute masks clear exactly the excluded
and 1-4 pairs."""

import numpy as np
import pytest

from conftest import NULL_ENTRY, exclusion_csr, replay_engine_list
from dynamics_amd import systems

RC = 10.0
INTERACT = np.array([[1, 1, 1], [1, 1, 0], [1, 0, 1]], bool)  # solute, O, charged


def _classes_gids(s):
    nd, nw = s.n_dyn, s.n_water
    c = np.empty(nd + 4 * nw, np.int64)
    c[:nd] = 0
    c[nd:] = np.tile(np.array([1, 2, 2, 2]), nw)
    g = np.empty(nd + 4 * nw, np.int64)
    g[:nd] = -2 - np.arange(nd)
    g[nd:] = np.repeat(np.arange(nw), 4)
    return c, g


@pytest.mark.parametrize("name", ["water512", "solvated"])
@pytest.mark.parametrize("buffer", [0.0, 1.5])
def test_list_replay_properties(name, buffer):
    s = systems.small_water() if name == "water512" else systems.small_solvated()
    n, slot_site, cl_class, row_ptr, cj, cm = replay_engine_list(s, RC, buffer)
    pos, *_ = s.site_arrays()
    ns = pos.shape[0]
    cls, gid = _classes_gids(s)
    real = slot_site[slot_site >= 0]
    assert np.array_equal(np.sort(real), np.arange(ns)), "each site in exactly one slot"
    slot_of = np.empty(ns, np.int64)
    slot_of[slot_site[slot_site >= 0]] = np.nonzero(slot_site >= 0)[0]
    cl_of = slot_of // 8
    assert np.array_equal(cl_class[cl_of], cls)
    jslot = (cj & NULL_ENTRY).astype(np.int64)
    jcls = (cj >> 30).astype(np.int64)
    grp_row = np.repeat(np.arange(n), np.diff(row_ptr.astype(np.int64)))
    ent_row = np.repeat(grp_row, 8)
    null = jslot == NULL_ENTRY
    # tags, segment order and padding
    assert np.all(jcls[~null] == cl_class[jslot[~null] // 8])
    assert np.all(jcls.reshape(-1, 8) == jcls.reshape(-1, 8)[:, :1]), "a group mixes classes"
    for I in range(n):
        seg = jcls[8 * row_ptr[I]:8 * row_ptr[I + 1]]
        assert np.all(np.diff(seg) >= 0), "class segments out of order"
        assert np.all(INTERACT[cl_class[I], seg])
    listed = set(zip(ent_row[~null].tolist(), jslot[~null].tolist()))
    assert len(listed) == int((~null).sum()), "duplicate j entries"
    box = np.asarray(s.box, np.float64)
    p = pos.astype(np.float64)
    rl = RC + buffer
    # coverage: every interacting pair within rc, both ways
    missing = 0
    for a in range(ns):
        d = p[a + 1:] - p[a]
        d -= box * np.round(d / box)
        r2 = (d * d).sum(1)
        b = np.nonzero(r2 < RC * RC)[0] + a + 1
        b = b[INTERACT[cls[a], cls[b]] & (gid[b] != gid[a])]
        for bb in b:
            if (cl_of[a], slot_of[bb]) not in listed or (cl_of[bb], slot_of[a]) not in listed:
                missing += 1
    assert missing == 0, f"{missing} interacting pairs within rc not listed"
    # minimality: each listed j slot is a site within rlist of a site of the
    # row's cluster (possibly of its own water molecule: the pair kernel
    # masks those pairs)
    slot_pos = np.zeros((len(slot_site), 3))
    slot_gid = np.full(len(slot_site), -1)
    ok = slot_site >= 0
    slot_pos[ok] = p[slot_site[ok]]
    slot_gid[ok] = gid[slot_site[ok]]
    for I, js in listed:
        pi = slot_pos[8 * I:8 * I + 8]
        gi = slot_gid[8 * I:8 * I + 8]
        d = pi - slot_pos[js]
        d -= box * np.round(d / box)
        r2 = (d * d).sum(1)
        good = (gi != -1) & (r2 < rl * rl * (1 + 1e-5))
        assert good.any(), (I, js)
    # masks of solute–solute groups
    ptr_, idx = exclusion_csr(s)
    excl = {(i, int(j)) for i in range(s.n_dyn) for j in idx[ptr_[i]:ptr_[i + 1]]}
    for gq in range(len(cm)):
        I = grp_row[gq]
        if cl_class[I] != 0 or jcls[8 * gq] != 0:
            continue
        for lane in range(64):
            ia, jb = lane % 8, lane // 8
            sa = slot_site[8 * I + ia]
            js = jslot[8 * gq + jb]
            sb = slot_site[js] if js != NULL_ENTRY else -1
            want = not (sa >= 0 and sb >= 0 and (int(sa), int(sb)) in excl)
            assert bool((int(cm[gq]) >> lane) & 1) == want, (I, gq, lane)
