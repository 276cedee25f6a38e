"""Forests (batched solves), host side: packing trees into one forest and
drawing the batched configuration's trees."""
import numpy as np

from afivo_amd import forest as F


def small_forest(dim, n, seed=3, nc=8):
    """Trees of uneven depth: level 2 uniform, random refinement up to a
    per-tree maximum level."""
    l_range = (2, 5) if dim == 2 else (2, 4)
    out = []
    for t in range(n):
        st, spec = F.c3_tree(seed, t, dim=dim, l_uniform=2, p_ref=0.5, nc_choices=(nc,), l_range=l_range)
        out.append(st)
    return out


# ---------------------------------------------------------------------------
# host side (CPU)

def test_combine_and_split_roundtrip():
    sts = small_forest(2, 4)
    st, tree_of = F.combine(sts)
    assert st.n_boxes == sum(s.n_boxes for s in sts)
    assert np.array_equal(np.bincount(tree_of), [s.n_boxes for s in sts])
    assert len(np.unique(st.ids)) == st.n_boxes
    for iv in range(3):
        parts = F.split(sts, st.data[iv])
        for a, s in zip(parts, sts):
            assert np.array_equal(a, s.data[iv])
    # links stay inside a tree
    D = st.dim
    id2tree = dict(zip(st.ids.tolist(), tree_of.tolist()))
    for r in range(st.n_boxes):
        for col in [2 + D] + list(range(3 + D, 3 + D + (1 << D) + 2 * D)):
            v = int(st.recs[r, col])
            if v > 0:
                assert id2tree[v] == tree_of[r]


def test_tree_draw_is_independent_of_other_trees():
    """A tree depends only on (seed, tree index): drawing it alone or after
    others gives the same boxes and rhs."""
    a, _ = F.c3_tree(5, 7, l_uniform=2, l_range=(2, 5))
    for t in range(3):
        F.c3_tree(5, t, l_uniform=2, l_range=(2, 5))
    b, _ = F.c3_tree(5, 7, l_uniform=2, l_range=(2, 5))
    assert np.array_equal(a.recs, b.recs)
    assert np.array_equal(a.data[1], b.data[1])


def test_c3_spec_ranges():
    specs = [F.c3_spec(1, t) for t in range(200)]
    assert {s.nc for s in specs} == {8, 16}
    assert {s.l_max for s in specs} == set(range(4, 9))
    for s in specs:
        assert 1 <= s.gs.ampl.size <= 3
        assert np.all((s.gs.sigma >= 0.02) & (s.gs.sigma <= 0.08))
        assert np.all((s.gs.r0 >= 0.1) & (s.gs.r0 <= 0.9))
