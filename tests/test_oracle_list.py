"""The host restatement of the engine's cluster-pair list, checked against
brute force: every interacting site pair within the cutoff is covered by an
entry, rows are symmetric, each site sits in exactly one slot, and the
solute–solute masks clear exactly the excluded and 1-4 pairs."""

import numpy as np
import pytest

from conftest import exclusion_csr, replay_engine_list
from dynamics_amd import systems

RC = 10.0


def _classes(s):
    nd, nw = s.n_dyn, s.n_water
    c = np.empty(nd + 4 * nw, np.int64)
    c[:nd] = 0
    w = np.tile(np.array([1, 2, 2, 2]), nw)
    c[nd:] = w
    return c


@pytest.mark.parametrize("name", ["water512", "solvated"])
@pytest.mark.parametrize("buffer", [0.0, 1.5])
def test_list_replay_properties(name, buffer):
    s = systems.small_water() if name == "water512" else systems.small_solvated()
    n, slot_site, cl_class, row_ptr, cj, cm = replay_engine_list(s, RC, buffer)
    pos, *_ = s.site_arrays()
    ns = pos.shape[0]
    cls = _classes(s)
    # each site in exactly one slot; cluster class matches its sites
    real = slot_site[slot_site >= 0]
    assert np.array_equal(np.sort(real), np.arange(ns))
    slot_of = np.empty(ns, np.int64)
    slot_of[slot_site[slot_site >= 0]] = np.nonzero(slot_site >= 0)[0]
    cl_of = slot_of // 8
    assert np.array_equal(cl_class[cl_of], cls)
    # entry set and symmetry
    rows = np.repeat(np.arange(n), np.diff(row_ptr.astype(np.int64)))
    ent = set(zip(rows.tolist(), cj.tolist()))
    assert len(ent) == len(cj), "duplicate entries"
    assert all((j, i) in ent for i, j in ent), "list not symmetric"
    assert all((i, i) in ent for i in range(n)), "a cluster does not list itself"
    # coverage: all interacting site pairs within the cutoff (min image)
    box = np.asarray(s.box, np.float64)
    p = pos.astype(np.float64)
    inter = np.array([[1, 1, 1], [1, 1, 0], [1, 0, 1]], bool)  # solute, O, charged
    missing = 0
    for a in range(ns):
        d = p[a + 1:] - p[a]
        d -= box * np.round(d / box)
        r2 = (d * d).sum(1)
        b = np.nonzero(r2 < RC * RC)[0] + a + 1
        b = b[inter[cls[a], cls[b]]]
        for bb in b:
            if (cl_of[a], cl_of[bb]) not in ent:
                missing += 1
    assert missing == 0, f"{missing} interacting pairs within rc not covered"
    # masks of solute–solute entries
    ptr_, idx = exclusion_csr(s)
    excl = set()
    for i in range(s.n_dyn):
        for j in idx[ptr_[i]:ptr_[i + 1]]:
            excl.add((i, int(j)))
    for e, (I, J) in enumerate(zip(rows, cj)):
        if cl_class[I] != 0 or cl_class[J] != 0:
            continue
        for lane in range(64):
            ia, jb = lane % 8, lane // 8
            sa, sb = slot_site[8 * I + ia], slot_site[8 * J + jb]
            want = not (sa >= 0 and sb >= 0 and (int(sa), int(sb)) in excl)
            assert bool((int(cm[e]) >> lane) & 1) == want, (I, J, lane)
