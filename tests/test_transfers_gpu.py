"""Ghost cells and transfers of any variable on the device: afg_gc_tree /
afg_gc_level (a$D_gc_tree with mg$D_sides_rb, a$D_gc_prolong0,
a$D_gc_interp, a$D_gc_interp_lim and the boundary table or Dirichlet /
Neumann zero), afg_restrict_tree, afg_prolong_level (a$D_prolong0/1/2),
afg_consistent_fluxes on face variables and the whole-tree arithmetic
(a$D_box_add_cc .. a$D_box_lincomb_cc).

* against the reference's own output (tests/golden/xfer_*.npz): bit for bit;
* against the oracle on the 2D and 3D benchmark trees with random values in
  every cell (ghosts included) and random boundary values: bit for bit, for
  every method and boundary condition;
* the arithmetic against numpy's IEEE operations.

Device tests; also run through the CPU emulation (test_emul.py)."""
import os

import numpy as np
import pytest

import oracle as O
from afivo_amd import mg as M
from afivo_amd import problems as P
from afivo_amd._lib import AfgError
from helpers import (BC_DIRICHLET, BC_NEUMANN, assert_bitexact, face_gradient, random_state, run_xfer,
                     xfer_expected)
from test_oracle import GOLDEN, XFERS, load_golden, load_xfer

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", XFERS)
def test_transfers_match_reference(name):
    meta, ops = load_xfer(name)
    _, _, st, _ = load_golden(meta["base"])
    dt = M.DeviceTree.from_state(st, 4)
    if meta["cfg"].get("cyl"):
        dt.set_coord(True)
    for op, (pos, val) in ops.items():
        src = meta["ops"][op]
        dt.upload(4, st.data[src - 1])
        run_xfer(dt, op, st.min_lvl, st.highest_lvl)
        assert_bitexact(dt.download(4), xfer_expected(st.data[src - 1], pos, val), f"{name} {op}")
        for iv in (1, 2, 3):
            assert_bitexact(dt.download(iv), st.data[iv - 1], f"{name} {op}: variable {iv}")
    dt.close()


@pytest.mark.parametrize("name", XFERS)
def test_second_ghost_layer_matches_reference(name):
    """afg_gc2_boxes (a$D_gc2_box with a$D_gc2_prolong1 and
    a$D_bc2_neumann_zero) of variable 4 after the gc0 step, every box: the
    reference's dump bit for bit; the same for a subset of boxes given in
    another order."""
    meta, ops = load_xfer(name)
    _, _, st, _ = load_golden(meta["base"])
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    dt = M.DeviceTree.from_state(st, 4)
    pos, val = ops["gc0"]
    dt.upload(4, xfer_expected(st.data[0], pos, val))
    assert_bitexact(dt.gc2(4), z["gc2"], name)
    n = (2 * st.dim * st.ncell() ** (st.dim - 1)).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(n)])
    pick = np.arange(st.n_boxes)[::-3]
    want = np.concatenate([z["gc2"][off[r]:off[r + 1]] for r in pick])
    assert_bitexact(dt.gc2(4, st.ids[pick]), want, f"{name} subset")
    dt.close()


def random_tree(dim, seed):
    """The benchmark tree of the dimension (2D: 951 boxes of 8^2; 3D: 5875
    boxes of 8^3) with random values in 3 variables and random Dirichlet /
    Neumann values on a third of the boundary faces."""
    p = P.c1() if dim == 2 else P.c2()
    st = p.state
    st.data = random_state(st, 3, seed)
    rng = np.random.default_rng(seed + 1)
    bcs = []
    for r in range(st.n_boxes):
        for f in range(2 * dim):
            if st.neighbors[r, f] < 0 and rng.uniform() < 0.33:
                n = int(st.ncell()[r]) ** (dim - 1)
                bcs.append((int(st.ids[r]), f + 1, int(rng.choice([BC_DIRICHLET, BC_NEUMANN])),
                            rng.standard_normal(n)))
    return st, bcs


def both(st, bcs):
    dt = M.DeviceTree.from_state(st)
    ot = O.OracleTree(st)
    for id_, nb, t, b in bcs:
        dt.set_bc(id_, nb, t, b)
        ot.set_bc(id_, nb, t, b)
    return dt, ot


@pytest.mark.parametrize("dim", [2, 3])
def test_transfers_match_oracle_full_size(dim):
    st, bcs = random_tree(dim, seed=40 + dim)
    dt, ot = both(st, bcs)
    for rb in range(4):
        for bc in (0, BC_DIRICHLET, BC_NEUMANN):
            dt.upload(1, st.data[0])
            ot.set(1, st.data[0])
            dt.gc_tree(1, rb, bc)
            ot.gc_tree(1, rb, bc)
            assert_bitexact(dt.download(1), ot.get(1), f"gc_tree rb {rb} bc {bc}")
    # one level at a time, lowest first, is the whole-tree fill (the oracle
    # holds gc_tree's result for rb 3, Neumann zero)
    dt.upload(1, st.data[0])
    for lvl in range(st.min_lvl, st.highest_lvl + 1):
        dt.gc_level(1, lvl, 3, BC_NEUMANN)
    assert_bitexact(dt.download(1), ot.get(1), "gc_level")
    # restriction into another variable, and in place
    dt.restrict_tree(2, 3)
    ot.restrict_tree(2, 3)
    assert_bitexact(dt.download(3), ot.get(3), "restrict_tree 2 -> 3")
    dt.restrict_tree(2)
    ot.restrict_tree(2)
    assert_bitexact(dt.download(2), ot.get(2), "restrict_tree 2")
    # prolongation: every order, added or not, in place and into another variable
    for order in range(3):
        for add in (False, True):
            for iv_to in (3, 1):
                for lvl in range(st.min_lvl + 1, st.highest_lvl + 1):
                    dt.prolong_level(1, iv_to, lvl, order, add)
                    ot.prolong_level(1, iv_to, lvl, order, add)
                assert_bitexact(dt.download(iv_to), ot.get(iv_to), f"prolong order {order} add {add} to {iv_to}")
    dt.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_second_ghost_layer_matches_oracle_full_size(dim):
    """afg_gc2_boxes on the benchmark trees with random values in every cell
    (ghosts included, as left by a random first layer): the oracle's
    a$D_gc2_box of every box, bit for bit."""
    st, _ = random_tree(dim, seed=60 + dim)
    dt = M.DeviceTree.from_state(st)
    ot = O.OracleTree(st)
    assert_bitexact(dt.gc2(2), ot.gc2(2), f"{dim}D gc2")
    dt.close()


def test_tree_arithmetic():
    """afg_add_cc / sub_cc / times_cc / lincomb_cc: the IEEE operations of
    a$D_box_*_cc on every value, ghosts included."""
    p = P.small_3d()
    st = p.state
    st.data = random_state(st, 3, seed=5)
    x, y, z = st.data
    dt = M.DeviceTree.from_state(st)
    dt.add_cc(1, 2)
    assert_bitexact(dt.download(2), y + x, "add")
    dt.sub_cc(1, 2)
    assert_bitexact(dt.download(2), (y + x) - x, "sub")
    dt.times_cc(-0.375, 3)
    assert_bitexact(dt.download(3), -0.375 * z, "times")
    dt.lincomb_cc(1.5, 1, -2.25, 3)
    assert_bitexact(dt.download(3), 1.5 * x + -2.25 * (-0.375 * z), "lincomb")
    assert_bitexact(dt.download(1), x, "source unchanged")
    dt.close()


def test_gc_rejects_bad_arguments():
    p = P.small_2d()
    dt = M.DeviceTree.from_state(p.state)
    with pytest.raises(AfgError):
        dt.gc_tree(1, 4, 0)
    with pytest.raises(AfgError):
        dt.gc_tree(1, 0, 7)
    with pytest.raises(AfgError):
        dt.prolong_level(1, 3, dt.min_lvl, 1)
    with pytest.raises(AfgError):
        dt.prolong_level(1, 3, dt.highest_lvl, 3)
    dt.close()


@pytest.mark.parametrize("name", XFERS)
def test_consistent_fluxes_match_reference(name):
    """a$D_consistent_fluxes on face variable 1 = the gradient of phi: the
    reference's faces bit for bit; another face variable is untouched."""
    meta, _ = load_xfer(name)
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    _, _, st, _ = load_golden(meta["base"])
    before = face_gradient(st, st.data[0])
    dt = M.DeviceTree.from_state(st)
    if meta["cfg"].get("cyl"):
        dt.set_coord(True)
    dt.upload_fc(1, before)
    dt.upload_fc(3, before)
    dt.consistent_fluxes([1])
    assert_bitexact(dt.download_fc(1), xfer_expected(before, z["flx_pos"], z["flx_val"]), name)
    assert_bitexact(dt.download_fc(3), before, "face variable 3 untouched")
    dt.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_consistent_fluxes_match_oracle_full_size(dim):
    st, _ = random_tree(dim, seed=60 + dim)
    rng = np.random.default_rng(7)
    n = int(st.face_offsets()[-1])
    f1, f2 = rng.standard_normal(n), rng.standard_normal(n)
    dt = M.DeviceTree.from_state(st)
    ot = O.OracleTree(st)
    dt.upload_fc(1, f1)
    dt.upload_fc(2, f2)
    ot.set_fc(1, f1)
    ot.set_fc(2, f2)
    dt.consistent_fluxes([1, 2])
    ot.consistent_fluxes([1, 2])
    for k in (1, 2):
        got, want = dt.download_fc(k), ot.get_fc(k)
        assert_bitexact(got, want, f"face variable {k}")
        assert np.count_nonzero(want != (f1 if k == 1 else f2)) > 0
    dt.close()


def _box_ptrs(arrays):
    from afivo_amd import _lib
    return (_lib.F64P * len(arrays))(*[a.ctypes.data_as(_lib.F64P) for a in arrays])


@pytest.mark.parametrize("dim", [2, 3])
def test_box_array_transfers(dim):
    """afg_upload_boxes / afg_download_boxes (one host array per box, as the
    Fortran drop-in passes box%cc slabs): on the benchmark tree, boxes given
    in a shuffled order and a subset, against the packed transfers, bit for
    bit; other boxes keep their values; a repeated id, a missing array or an
    unknown id is an error."""
    from afivo_amd import _lib
    p = P.c1() if dim == 2 else P.c2()
    st = p.state
    rng = np.random.default_rng(7)
    a = rng.standard_normal(st.n_values())
    b = rng.standard_normal(st.n_values())
    dt = M.DeviceTree.from_state(st, len(st.data))
    dt.upload(1, a)
    off = st.offsets()
    order = rng.permutation(st.n_boxes)[: st.n_boxes * 2 // 3]
    ids = np.ascontiguousarray(st.ids[order], dtype=np.int32)
    boxes = [np.ascontiguousarray(b[off[r]:off[r + 1]]) for r in order]
    L = _lib.load()
    dt._c(L.afg_upload_boxes(dt.h, 1, ids.size, ids.ctypes.data_as(_lib.I32P), _box_ptrs(boxes)))
    want = a.copy()
    for r in order:
        want[off[r]:off[r + 1]] = b[off[r]:off[r + 1]]
    assert_bitexact(dt.download(1), want, "upload_boxes")
    got = [np.full_like(x, np.nan) for x in boxes]
    dt._c(L.afg_download_boxes(dt.h, 1, ids.size, ids.ctypes.data_as(_lib.I32P), _box_ptrs(got)))
    for r, g in zip(order, got):
        assert_bitexact(g, b[off[r]:off[r + 1]], f"download_boxes box {st.ids[r]}")
    dup = np.ascontiguousarray(np.r_[ids[:3], ids[:1]], dtype=np.int32)
    with pytest.raises(AfgError, match="twice"):
        dt._c(L.afg_upload_boxes(dt.h, 1, 4, dup.ctypes.data_as(_lib.I32P), _box_ptrs(boxes[:3] + boxes[:1])))
    ptrs = _box_ptrs(boxes[:2])
    ptrs[1] = _lib.F64P()
    with pytest.raises(AfgError, match="no array"):
        dt._c(L.afg_download_boxes(dt.h, 1, 2, ids.ctypes.data_as(_lib.I32P), ptrs))
    bad = np.array([st.ids.max() + 5], dtype=np.int32)
    with pytest.raises(AfgError, match="no box"):
        dt._c(L.afg_upload_boxes(dt.h, 1, 1, bad.ctypes.data_as(_lib.I32P), _box_ptrs(boxes[:1])))
    assert_bitexact(dt.download(1), want, "after the rejected calls")
    dt.close()
