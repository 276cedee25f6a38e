"""Host-side tree construction (afivo_amd.tree) against the reference's trees."""
import os

import numpy as np
import pytest

import oracle as O
from afivo_amd import problems as P
from afivo_amd import tree as T
from afivo_amd.treefile import TreeState, read_tree_file, write_tree_file
from test_oracle import CASES, load_golden


def keyed(st):
    """(lvl, ix) -> (neighbour keys or codes, parent key, child keys)."""
    idk = {int(st.recs[r, 0]): tuple(st.recs[r, 1:2 + st.dim].tolist()) for r in range(st.n_boxes)}
    out = {}
    for r in range(st.n_boxes):
        k = idk[int(st.recs[r, 0])]
        nb = tuple(idk[int(x)] if x > 0 else int(x) for x in st.neighbors[r])
        par = idk.get(int(st.parent[r]), 0)
        ch = tuple(idk[int(x)] if x > 0 else 0 for x in st.children[r])
        out[k] = (nb, par, ch)
    return out


def rhs_by_key(st):
    off = st.offsets()
    return {tuple(st.recs[r, 1:2 + st.dim].tolist()): st.data[1][off[r]:off[r + 1]]
            for r in range(st.n_boxes)}


def gauss_of(cfg):
    c = cfg["cfg"]
    dim = cfg["dim"]
    n = c["n_gauss"]
    return P.Gaussians(np.array(c["ampl"][:n], dtype=float), np.array(c["sigma"][:n], dtype=float),
                       np.array(c["r0"], dtype=float).reshape(n, dim))


@pytest.mark.parametrize("name", CASES)
def test_builder_reproduces_reference_tree(name):
    cfg, st_ref, _, _ = load_golden(name)
    c = cfg["cfg"]
    gs = gauss_of(cfg)
    p = P.make_problem(name, cfg["dim"], c["nc"], gs, c["l_uniform"], c["l_max"], c["thr"],
                       c["n_base"], c.get("n_vcycle", 0), eps_hi=c.get("eps_hi") if c.get("use_eps") else None)
    assert keyed(p.state) == keyed(st_ref)
    a, b = rhs_by_key(p.state), rhs_by_key(st_ref)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    if c.get("use_eps"):
        ea = {k: v for k, v in zip(a, [None] * len(a))}
        off_a, off_b = p.state.offsets(), st_ref.offsets()
        ka = {tuple(p.state.recs[r, 1:2 + p.state.dim].tolist()): p.state.data[3][off_a[r]:off_a[r + 1]]
              for r in range(p.state.n_boxes)}
        kb = {tuple(st_ref.recs[r, 1:2 + st_ref.dim].tolist()): st_ref.data[3][off_b[r]:off_b[r + 1]]
              for r in range(st_ref.n_boxes)}
        for k in ka:
            np.testing.assert_array_equal(ka[k], kb[k])


def test_builder_c1_size():
    p = P.c1()
    assert p.state.n_boxes == 951
    assert p.tree.n_leaf_cells() == 45568


@pytest.mark.skipif(not O.ref_available(2), reason="reference build oracle/_ref absent")
def test_builder_random_list_tree(tmp_path):
    """A tree refined at random boxes, replayed through the reference's own
    adjust_refinement (list mode): identical structure, so the builder's 2:1
    balance matches the reference's."""
    rng = np.random.default_rng(11)

    def ref(view, cur):
        h = (view.ix[:, 0] * 7919 + view.ix[:, 1] * 104729 + view.lvl * 1299709) % 1000
        return np.where((view.lvl < 3) | ((view.lvl < 6) & (h < 150)), T.REFINE, T.KEEP)

    t = T.build(2, 8, ref)
    st = t.state()
    par = st.recs[(st.children[:, 0] != 0) & (st.lvl >= 1)]
    keys = sorted((int(r[1]), int(r[3]), int(r[2])) for r in par)
    lst = tmp_path / "list.txt"
    with open(lst, "w") as f:
        f.write(f"{len(keys)}\n")
        for k in keys:
            f.write(" ".join(map(str, k)) + "\n")
    prefix, _ = O.run_ref(2, dict(nc=8, ref_mode="list", list_file=str(lst), n_gauss=1,
                                  n_vcycle=0, dump_output=False, out_file="lt"), workdir=str(tmp_path))
    st_ref = read_tree_file(prefix + ".in")
    assert keyed(st) == keyed(st_ref)
    del rng


def test_treefile_roundtrip(tmp_path):
    p = P.small_2d()
    path = tmp_path / "t.afgt"
    write_tree_file(path, p.state)
    st = read_tree_file(path)
    assert np.array_equal(st.recs, p.state.recs)
    for a, b in zip(st.data, p.state.data):
        np.testing.assert_array_equal(a, b)


def test_morton_ids_and_periodic_base():
    """Two base boxes, periodic in x: every box has a neighbour on both x sides."""
    t = T.Tree(2, 4, 0.25, coarsen_to=2)
    t.set_base([[1, 1], [2, 1]], [[2, 2, -1, -1], [1, 1, -1, -1]])
    for _ in range(3):
        t.adjust_refinement(lambda v, cur: np.where(v.lvl < 3, T.REFINE, T.KEEP))
    st = t.state()
    top = st.lvl == st.highest_lvl
    assert (st.neighbors[top][:, :2] > 0).all()
    assert (st.neighbors[top][:, 2:] != 0).all()


UPDATES = sorted(f[:-4] for f in os.listdir(os.path.join(os.path.dirname(__file__), "golden"))
                 if f.startswith("update_") and f.endswith(".npz"))


@pytest.mark.parametrize("name", UPDATES)
def test_adjust_refinement_reproduces_reference_update(name):
    """A solved golden tree refined and derefined once more (the reference's
    own a$D_adjust_refinement, fixture update_*.npz): the builder's pass with
    the same rule (refine below refine2_lmax where dr^2 max|lpl| >
    refine2_thr, else derefine, below l_uniform always refine) gives the
    same boxes and links."""
    import json
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    meta = json.loads(str(z["meta"]))
    cfg, _, st_out, _ = load_golden(meta["base"])
    c = meta["cfg"]
    gs = gauss_of(cfg)
    p = P.make_problem(name, cfg["dim"], c["nc"], gs, c["l_uniform"], c["l_max"], c["thr"],
                       c["n_base"], 0)
    assert keyed(p.state) == keyed(st_out)
    grow = P.gauss_refinement(gs, c["l_uniform"], c["refine2_lmax"], c["refine2_thr"])

    def ref(view, cur):
        out = grow(view, T.DEREFINE)
        out[view.lvl < c["l_uniform"]] = T.REFINE
        return out

    p.tree.adjust_refinement(ref)
    st_new = TreeState(cfg["dim"], c["nc"], 0, 0, 0, 1.0, z["recs"], [])
    assert keyed(p.tree.state()) == keyed(st_new)
