"""Per-box operators on the host (afg_box_*: the drop-in module's mg$D_box_*
and mg$D_sides_rb procedures) against the oracle's box procedures, bit for
bit (sign of zero included), on every box of small trees: 2D and 3D
Cartesian trees with the lpl, lpld and lpllsf families and a 2D cylindrical
tree with clpl and clpld, random values in every variable (ghosts and
corners included).  The oracle runs each procedure on the box inside its
tree (orc_box_call); the library gets the box's arrays alone.

The operators are host code: the CPU run goes through the emulation build
(tests/emul, the same source compiled with g++), the gpu-marked run through
libafivo_amd.so itself.
"""
import ctypes

import numpy as np
import pytest

import conftest
import oracle as O
from afivo_amd import _lib
from afivo_amd import mg as M
from afivo_amd import problems as P
from helpers import assert_bitexact

LPL, LPLD, LSF, CLPL, CLPLD = M.AFG_BOX_LPL, M.AFG_BOX_LPLD, M.AFG_BOX_LSF, M.AFG_BOX_CLPL, M.AFG_BOX_CLPLD
PHI, RHS, TMP, EPS, LS, BV = 1, 2, 3, 4, 5, 6


@pytest.fixture(params=["emul", pytest.param("hip", marks=pytest.mark.gpu)])
def library(request):
    previous = _lib._lib
    if request.param == "emul":
        if not __import__("os").path.exists(conftest.EMUL_LIB):
            pytest.skip("emulation library not built (make -C tests/emul)")
        conftest.load_emulation()
    else:
        _lib._lib = None
        _lib.load()
    yield request.param
    _lib._lib = previous


def tree_case(case, seed=3):
    """(state, cyl, r_base): the tree with six random variables; eps in (0.5,
    3); the level set of random sign so that every interface branch is taken."""
    if case == "2d":
        p = P.small_2d(nc=8, l_uniform=2, l_max=4)
    elif case == "3d":
        p = P.small_3d(nc=4, l_uniform=2, l_max=4)
    else:
        p = P.small_2d(nc=4, l_uniform=2, l_max=4)
    st = p.state
    rng = np.random.default_rng(seed)
    n = st.n_values()
    st.data = [rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n),
               rng.uniform(0.5, 3.0, n), rng.standard_normal(n), rng.standard_normal(n)]
    return st, case == "cyl", 0.25 if case == "cyl" else 0.0


class Boxes:
    """Per-record views of a TreeState: ids, parents, offsets, cell counts,
    child offsets (a$D_get_child_offset, m_aX_t.f90:362-374), cell sizes and,
    for a cylindrical tree, r_min(1) as the oracle sets it."""

    def __init__(self, st, cyl, r_base):
        self.st = st
        self.ids = [int(i) for i in st.ids]
        self.rec = {i: r for r, i in enumerate(self.ids)}
        self.off = st.offsets()
        self.n = [int(x) for x in st.ncell()]
        self.lvl = [int(x) for x in st.lvl]
        self.ix = st.ix
        self.parent = [int(x) for x in st.parent]
        self.dr = [st.dr_base * 0.5 ** (l - 1) for l in self.lvl]
        self.rmin = {}
        if cyl:
            order = sorted(range(st.n_boxes), key=lambda r: self.lvl[r])
            for r in order:
                ix = int(self.ix[r, 0])
                if self.lvl[r] <= 1:
                    self.rmin[r] = r_base + (ix - 1) * st.dr_base * st.nc
                else:
                    q = self.rec[self.parent[r]]
                    self.rmin[r] = self.rmin[q] + 0.5 * self.dr[q] * ((ix - 1) & 1) * self.n[q]

    def view(self, a, r):
        return a[self.off[r]:self.off[r + 1]]

    def child_offset(self, r):
        if self.lvl[r] <= 1:
            return [0] * self.st.dim
        return [((int(self.ix[r, d]) - 1) & 1) * (self.n[r] >> 1) for d in range(self.st.dim)]


def oracle_of(st, cyl, r_base):
    ot = O.OracleTree(st, n_var=6)
    if cyl:
        ot.set_coord(True, r_base)
    return ot, O.default_opts(2, 2, 2, eps=True, lsf=True)


def call(ot, om, what, kind, id_, arg):
    assert O.lib().orc_box_call(ot.h, ctypes.byref(om), what, kind, id_, arg) == 0


CASES = [("2d", [LPL, LPLD, LSF]), ("3d", [LPL, LPLD, LSF]), ("cyl", [CLPL, CLPLD])]


@pytest.mark.parametrize("case,kinds", CASES, ids=[c for c, _ in CASES])
def test_box_op_and_gsrb(library, case, kinds):
    """mg%box_op into tmp on every box, then box_gsrb colours 1, 2 and 7
    (only the parity selects the cells) on every box, per family."""
    st, cyl, r_base = tree_case(case)
    B = Boxes(st, cyl, r_base)
    dim = st.dim
    for kind in kinds:
        ot, om = oracle_of(st, cyl, r_base)
        d = [a.copy() for a in st.data]
        for r, id_ in enumerate(B.ids):
            ir0 = (int(B.ix[r, 0]) - 1) * B.n[r]
            args = dict(eps=B.view(d[EPS - 1], r), lsf=B.view(d[LS - 1], r), bval=B.view(d[BV - 1], r), ir0=ir0)
            M.box_op(dim, B.n[r], B.dr[r], kind, B.view(d[PHI - 1], r), B.view(d[TMP - 1], r), **args)
            call(ot, om, 0, kind, id_, TMP)
        assert_bitexact(d[TMP - 1], ot.get(TMP), f"{case} box op kind {kind}")
        for rb in (1, 2, 7):
            for r, id_ in enumerate(B.ids):
                ir0 = (int(B.ix[r, 0]) - 1) * B.n[r]
                M.box_gsrb(dim, B.n[r], B.dr[r], kind, rb, B.view(d[PHI - 1], r), B.view(d[RHS - 1], r),
                           eps=B.view(d[EPS - 1], r), lsf=B.view(d[LS - 1], r), bval=B.view(d[BV - 1], r),
                           ir0=ir0)
                call(ot, om, 1, kind, id_, rb)
            assert_bitexact(d[PHI - 1], ot.get(PHI), f"{case} box gsrb kind {kind} colour {rb}")


@pytest.mark.parametrize("case", ["2d", "3d", "cyl"])
def test_box_corr_rstr_and_sides_rb(library, case):
    """On every box with a parent: mg%box_corr of its phi from the parent's
    tmp (lpl, lpld, lpllsf), mg%box_rstr of rhs into the parent (rstr_lpl --
    the radial weights on the cylindrical tree -- and rstr_lpllsf), and
    mg$D_sides_rb of every face at a refinement boundary."""
    st, cyl, r_base = tree_case(case)
    B = Boxes(st, cyl, r_base)
    dim = st.dim
    nbs = st.neighbors
    children = [r for r in range(st.n_boxes) if B.parent[r] > 0]
    assert children
    for kind in [LPL, LPLD, LSF]:
        ot, om = oracle_of(st, cyl, r_base)
        d = [a.copy() for a in st.data]
        for r in children:
            q = B.rec[B.parent[r]]
            M.box_corr(dim, B.n[r], B.n[q], kind, B.child_offset(r), B.view(d[TMP - 1], q), B.view(d[PHI - 1], r),
                       p_eps=B.view(d[EPS - 1], q), p_lsf=B.view(d[LS - 1], q), c_lsf=B.view(d[LS - 1], r))
            call(ot, om, 2, kind, B.ids[r], 0)
        assert_bitexact(d[PHI - 1], ot.get(PHI), f"{case} box corr kind {kind}")
    for kind in [CLPL if cyl else LPL, LSF]:
        ot, om = oracle_of(st, cyl, r_base)
        d = [a.copy() for a in st.data]
        for r in children:
            q = B.rec[B.parent[r]]
            M.box_rstr(dim, B.n[r], B.n[q], kind, B.child_offset(r), B.view(d[RHS - 1], r), B.view(d[RHS - 1], q),
                       c_lsf=B.view(d[LS - 1], r), p_lsf=B.view(d[LS - 1], q), p_rmin=B.rmin.get(q, 0.0),
                       c_dr=B.dr[r])
            call(ot, om, 3, LSF if kind == LSF else LPL, B.ids[r], RHS)
        assert_bitexact(d[RHS - 1], ot.get(RHS), f"{case} box rstr kind {kind}")
    ot, om = oracle_of(st, cyl, r_base)
    d = [a.copy() for a in st.data]
    faces = 0
    for r in children:
        q = B.rec[B.parent[r]]
        for f in range(2 * dim):
            if nbs[r, f] == 0:
                M.sides_rb(dim, B.n[r], B.n[q], f + 1, B.child_offset(r), B.view(d[PHI - 1], q),
                           B.view(d[PHI - 1], r))
                call(ot, om, 4, 0, B.ids[r], f + 1)
                faces += 1
    assert faces > 0
    assert_bitexact(d[PHI - 1], ot.get(PHI), f"{case} sides_rb")


def test_box_operators_reject_bad_arguments(library):
    """Invalid sizes, kinds, offsets and missing arrays are errors with a
    message, and leave the output untouched."""
    nc = 4
    a = np.zeros((nc + 2) ** 2)
    out = np.full_like(a, 7.0)
    f64 = a.ctypes.data_as(_lib.F64P)
    assert _lib.load().afg_box_op(2, 3, 1.0, LPL, 0, f64, None, None, None, f64) == 1   # odd nc
    assert b"afg_box_op" in _lib.load().afg_last_error(None)
    with pytest.raises(_lib.AfgError, match="afg_box_op"):
        M.box_op(3, nc, 1.0, CLPL, np.zeros((nc + 2) ** 3), np.zeros((nc + 2) ** 3))
    with pytest.raises(_lib.AfgError, match="afg_box_gsrb"):
        M.box_gsrb(2, nc, 1.0, LPLD, 1, a, a)
    with pytest.raises(_lib.AfgError, match="afg_box_corr"):
        M.box_corr(2, nc, nc, LPL, [3, 0], a, a)
    with pytest.raises(_lib.AfgError, match="afg_box_rstr"):
        M.box_rstr(2, nc, nc, LSF, [0, 0], a, a)
    with pytest.raises(_lib.AfgError, match="afg_box_sides_rb"):
        M.sides_rb(2, nc, nc, 5, [0, 0], a, a)
    assert (out == 7.0).all()
