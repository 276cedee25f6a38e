"""Batched solves on the device: a forest of trees in one context gives every
tree exactly the result of solving it alone (and, through that, of the
reference).  Also run through the CPU emulation (test_emul.py)."""
import numpy as np
import pytest

import oracle as O
from afivo_amd import forest as F
from afivo_amd import mg as M
from afivo_amd._lib import AfgError
from helpers import assert_bitexact
from test_forest import small_forest

pytestmark = pytest.mark.gpu


def solve_sequence(dt, mg, per_tree_norms, n_vcycle=3, partial_lvl=None):
    """FMG with residual, then V-cycles; residual norms after each step."""
    norms = []
    M.mg_fas_fmg(dt, mg, True, False)
    norms.append(per_tree_norms(dt))
    for _ in range(n_vcycle):
        M.mg_fas_vcycle(dt, mg, dt.highest_lvl if partial_lvl is None else partial_lvl, True)
        norms.append(per_tree_norms(dt))
    return norms


@pytest.mark.parametrize("dim,n_trees", [(2, 6), (3, 3)])
@pytest.mark.parametrize("bottom", [8192, 0])
def test_forest_equals_single_solves(dim, n_trees, bottom):
    sts = small_forest(dim, n_trees)
    his = [s.highest_lvl for s in sts]
    assert len(set(his)) > 1, "the forest should mix tree depths"
    mg = M.default_mg(n_base=4)
    fdt = M.DeviceTree.from_forest(sts)
    fdt.set_bottom_cells(bottom)
    assert list(fdt.forest_trees()) == his
    fnorms = solve_sequence(fdt, mg, lambda d: d.forest_max_abs(3))
    fdata = [F.split(sts, fdt.download(iv)) for iv in (1, 2, 3)]
    for t, st in enumerate(sts):
        dt = M.DeviceTree.from_state(st)
        dt.set_bottom_cells(bottom)
        norms = solve_sequence(dt, mg, lambda d: np.array([M.max_abs_cc(d, 3)]))
        for iv in (1, 2, 3):
            assert_bitexact(fdata[iv - 1][t], dt.download(iv), f"tree {t} var {iv}")
        assert np.array_equal([n[t] for n in fnorms], [n[0] for n in norms]), f"tree {t} residuals"
        dt.close()
    fdt.close()


@pytest.mark.parametrize("nc", [8, 16])
def test_forest_partial_vcycle_and_oracle(nc):
    """A V-cycle from a level below some trees' highest level runs each tree
    from min(level, its highest); every tree checked against the oracle, for
    both box sizes of the batched configuration (8^2 and 16^2)."""
    sts = small_forest(2, 4, seed=11, nc=nc)
    mg = M.default_mg(n_base=4)
    om = O.default_opts(2, 2, 4)
    fdt = M.DeviceTree.from_forest(sts)
    lvl = sorted(s.highest_lvl for s in sts)[1]
    solve_sequence(fdt, mg, lambda d: d.forest_max_abs(3), n_vcycle=2, partial_lvl=lvl)
    fdata = [F.split(sts, fdt.download(iv)) for iv in (1, 2, 3)]
    for t, st in enumerate(sts):
        ot = O.OracleTree(st)
        ot.fas_fmg(om, True, False)
        for _ in range(2):
            ot.fas_vcycle(om, min(lvl, st.highest_lvl))
        for iv in (1, 2, 3):
            assert_bitexact(fdata[iv - 1][t], ot.get(iv), f"tree {t} var {iv}")
    fdt.close()


def test_forest_rejects_links_across_trees():
    sts = small_forest(2, 2)
    st, tree_of = F.combine(sts)
    bad = tree_of.copy()
    bad[1] = 1 - bad[1]                # one box moved to the other tree
    with pytest.raises(AfgError):
        M.DeviceTree(st.dim, st.nc, 3, st.recs, st.dr_base, bad)


@pytest.mark.parametrize("dim,n_trees", [(2, 5), (3, 3)])
def test_forest_mixed_precision_equals_single_solves(dim, n_trees):
    """Mixed-precision corrections on a forest: every tree gets exactly the
    phi and residual of the same number of corrections alone, and the forest
    solve meets the tolerance on every tree."""
    sts = small_forest(dim, n_trees)
    mg = M.default_mg(n_base=4)
    fdt = M.DeviceTree.from_forest(sts)
    M.mixed_begin(fdt, mg)
    fnorms = [fdt.forest_max_abs(3)]
    for _ in range(3):
        M.mixed_correct(fdt, mg, 2)
        fnorms.append(fdt.forest_max_abs(3))
    fdata = [F.split(sts, fdt.download(iv)) for iv in (1, 3)]
    for t, st in enumerate(sts):
        dt = M.DeviceTree.from_state(st)
        M.mixed_begin(dt, mg)
        norms = [dt.forest_max_abs(3)]
        for _ in range(3):
            M.mixed_correct(dt, mg, 2)
            norms.append(dt.forest_max_abs(3))
        assert_bitexact(fdata[0][t], dt.download(1), f"tree {t} phi")
        assert_bitexact(fdata[1][t], dt.download(3), f"tree {t} residual")
        assert np.array_equal([n[t] for n in fnorms], [n[0] for n in norms]), f"tree {t} residuals"
        dt.close()
    # the whole solve: every tree below the tolerance
    it, rr = M.mixed_solve(fdt, mg, 1e-10, max_iter=30)
    rhs = fdt.forest_max_abs(2)
    res = fdt.forest_max_abs(3)
    assert rr <= 1e-10 and np.all(res <= 1e-10 * rhs) and it >= 2
    assert rr == pytest.approx(float(np.max(res / rhs)), rel=0, abs=0)
    fdt.close()


def test_mixed_correct_needs_begin():
    sts = small_forest(2, 2)
    fdt = M.DeviceTree.from_forest(sts)
    with pytest.raises(AfgError):
        M.mixed_correct(fdt, M.default_mg(n_base=4), 2)
    fdt.close()
