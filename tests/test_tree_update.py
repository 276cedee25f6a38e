"""Tree updates on the device (afg_update_tree): the data side of
a$D_adjust_refinement (m_aX_core.f90:452-541) and the prolongation of the
boxes it adds (a$D_prolong0/1/2, m_aX_prolong.f90:39-348).

* against the reference's own output (tests/golden/update_*.npz, written by
  make_golden.py from the harness's refine2 step): kept boxes keep every
  value, added boxes get the reference's prolonged interiors, bit for bit;
* at the size of the 3D benchmark tree, through properties: kept boxes are
  unchanged, and a linear function with dyadic coefficients is prolonged
  exactly by all three orders (the stencils reproduce linear functions and
  every sum is exact in FP64);
* boxes are matched by level and index, not by id.

Device tests; also run through the CPU emulation (test_emul.py)."""
import json
import os

import numpy as np
import pytest

from afivo_amd import mg as M
from afivo_amd import problems as P
from afivo_amd import tree as T
from afivo_amd.treefile import TreeState
from helpers import assert_bitexact, random_state
from test_oracle import GOLDEN, load_golden

pytestmark = pytest.mark.gpu

UPDATES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.startswith("update_") and f.endswith(".npz"))


def load_update(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    return json.loads(str(z["meta"])), z["recs"], float(z["dr_base"]), z["added"], z["vals"]


def box_key(st, r):
    return (int(st.lvl[r]),) + tuple(int(x) for x in st.ix[r])


def downloaded_state(dt, like, recs, dr_base):
    lvl = recs[:, 1]
    st = TreeState(like.dim, like.nc, dt.n_var, int(lvl.min()), int(lvl.max()), dr_base, recs, [])
    st.data = [dt.download(iv) for iv in range(1, dt.n_var + 1)]
    return st


def check_update(old, new, prolonged, expect_added=None):
    """Kept boxes equal their old arrays in every variable; added boxes hold
    zeros outside the interiors of the prolonged variables, whose values
    expect_added(iv, r) gives (None: not checked)."""
    old_row = {box_key(old, r): r for r in range(old.n_boxes)}
    oo, on = old.offsets(), new.offsets()
    inner = (slice(1, -1),) * new.dim
    n_added = 0
    for r in range(new.n_boxes):
        ro = old_row.get(box_key(new, r))
        for iv in range(1, new.n_var + 1):
            got = new.data[iv - 1][on[r]:on[r + 1]]
            if ro is not None:
                assert_bitexact(got, old.data[iv - 1][oo[ro]:oo[ro + 1]], f"kept box {box_key(new, r)} var {iv}")
                continue
            box = new.box_array(iv, r).copy()
            if iv in prolonged:
                if expect_added is not None:
                    assert_bitexact(box[inner].ravel(), expect_added(iv, r), f"added box {box_key(new, r)} var {iv}")
                box[inner] = 0.0
            assert_bitexact(box.ravel(), np.zeros(box.size), f"added box {box_key(new, r)} var {iv}: rest")
        n_added += ro is None
    return n_added


@pytest.mark.parametrize("name", UPDATES)
def test_update_matches_reference(name):
    meta, recs, dr_base, added, vals = load_update(name)
    _, _, old, _ = load_golden(meta["base"])
    dt = M.DeviceTree.from_state(old)
    prolonged = tuple(range(1, meta["n_vars"] + 1))
    dt.update_tree(recs, prolong_order=meta["prolong_order"], vars=prolonged, dr_base=dr_base)
    assert dt.n_values == TreeState(old.dim, old.nc, 0, 0, 0, dr_base, recs, []).n_values()
    new = downloaded_state(dt, old, recs, dr_base)
    pos = {int(r): k for k, r in enumerate(added)}
    n = check_update(old, new, prolonged, lambda iv, r: vals[iv - 1, pos[r]])
    assert n == len(added) == meta["n_added"]
    dt.close()


def test_update_matches_boxes_not_ids():
    """The same tree under other ids and in another record order: every
    value stays with its box."""
    p = P.small_3d()
    st = p.state
    st.data = random_state(st, 3, seed=11)
    dt = M.DeviceTree.from_state(st)
    rng = np.random.default_rng(5)
    new_id = np.zeros(st.ids.max() + 1, dtype=np.int64)
    new_id[st.ids] = rng.permutation(st.n_boxes) + 7
    recs = st.recs.copy()
    for col in [0, 2 + st.dim] + list(range(3 + st.dim, recs.shape[1])):
        v = recs[:, col]
        recs[:, col] = np.where(v > 0, new_id[np.maximum(v, 0)], v)
    faces = rng.standard_normal(int(st.face_offsets()[-1]))
    dt.upload_fc(2, faces)
    order = rng.permutation(st.n_boxes)
    recs = recs[order]
    dt.update_tree(recs, prolong_order=2, vars=(1,))
    new = downloaded_state(dt, st, recs, st.dr_base)
    assert check_update(st, new, (1,)) == 0
    # face variables follow their boxes too
    fo = st.face_offsets()
    want = np.concatenate([faces[fo[r]:fo[r + 1]] for r in order])
    assert_bitexact(dt.download_fc(2), want, "face variable 2")
    dt.close()


def linear_values(st, coef):
    """coef . r at every cell centre of every box, ghost cells included
    (dyadic coordinates: exact in FP64)."""
    out = np.zeros(st.n_values())
    off = st.offsets()
    nc_l = st.ncell()
    for r in range(st.n_boxes):
        lvl, n = int(st.lvl[r]), int(nc_l[r])
        dr = st.dr_base * 0.5 ** (lvl - 1)
        r_min = (st.ix[r] - 1) * st.nc * dr if lvl >= 1 else np.zeros(st.dim)
        c = (np.arange(n + 2) - 0.5) * dr
        g = np.meshgrid(*([c] * st.dim), indexing="ij")   # g[d]: coordinate d, indexed (x, y, z)
        v = sum(coef[d] * (r_min[d] + g[d]) for d in range(st.dim))
        out[off[r]:off[r + 1]] = np.transpose(v).ravel()
    return out


@pytest.mark.parametrize("order", [0, 1, 2])
def test_update_full_size_properties(order):
    """The 3D benchmark tree (n_cell 8, about 2.6e6 leaf cells) refined
    around one point and coarsened elsewhere: kept boxes keep their values;
    orders 1 and 2 prolong a linear function exactly, order 0 copies the
    parent's cell."""
    p = P.c2()
    t, st = p.tree, p.state
    coef = np.array([1.0, 2.0, -4.0])
    st.data = [linear_values(st, coef), st.data[1]]
    dt = M.DeviceTree.from_state(st, 2)
    c0 = np.array([0.31, 0.62, 0.45])

    def ref(view, cur):
        out = np.array(cur, dtype=np.int64, copy=True) if np.ndim(cur) else np.full(view.ids.size, cur)
        mid = view.r_min + 0.5 * view.n_cell[:, None] * view.dr[:, None]
        d = np.sqrt(((mid - c0) ** 2).sum(axis=1))
        out[(d < 0.12) & (view.lvl < 8)] = T.REFINE
        out[(d > 0.3) & (view.lvl > 4)] = T.DEREFINE
        return out

    n_add = t.adjust_refinement(ref)
    assert n_add > 1000
    recs = t.records()
    dt.update_tree(recs, prolong_order=order, vars=(1,))
    new = downloaded_state(dt, st, recs, st.dr_base)
    exact = linear_values(new, coef)
    on = new.offsets()
    inner = (slice(1, -1),) * 3
    parent_row = {int(i): r for r, i in enumerate(recs[:, 0])}

    def expect(iv, r):
        if order > 0:
            return exact[on[r]:on[r + 1]].reshape((new.nc + 2,) * 3)[inner].ravel()
        # order 0: each child cell takes the parent cell it lies in
        pr = parent_row[int(new.parent[r])]
        pb = new.box_array(1, pr)
        h = new.nc // 2
        o = [(int(new.ix[r][d]) - 1) % 2 * h for d in range(3)]   # x, y, z offsets
        i = o[0] + (np.arange(new.nc) + 2) // 2
        j = o[1] + (np.arange(new.nc) + 2) // 2
        k = o[2] + (np.arange(new.nc) + 2) // 2
        return pb[np.ix_(k, j, i)].ravel()

    n = check_update(st, new, (1,), expect)
    assert n == n_add
    assert new.n_boxes < st.n_boxes + n_add   # some boxes were removed
    dt.close()
