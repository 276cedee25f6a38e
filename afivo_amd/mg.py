"""Python host API over the C ABI, mirroring afivo's multigrid interface.

Names follow the reference module ``m_a$D_mg`` (``src/m_aX_mg.f90:87-125``) and
the tree reductions of ``m_a$D_utils``; ``DeviceTree`` is the device mirror of
an ``a$D_t`` (topology plus the cell variables).  Errors raise
:class:`afivo_amd._lib.AfgError` with the library's message, where the
reference stops with a message.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import (AFG_BC_DIRICHLET, AFG_BC_NEUMANN, AFG_BOX_CLPL, AFG_BOX_CLPLD, AFG_BOX_LPL,  # noqa: F401
                   AFG_BOX_LPLD, AFG_BOX_LSF, AFG_OP_AUTO, AFG_OP_LPL, AFG_OP_LSF, AFG_RSTR_AUTO,
                   AFG_RSTR_LPL, AFG_RSTR_LSF, MgOpts, check)
from .treefile import TreeState

__all__ = ["DeviceTree", "MgOpts", "mg_init_mg", "mg_fas_fmg", "mg_fas_vcycle",
           "tree_max_cc", "tree_min_cc", "AFG_OP_LPL", "AFG_OP_AUTO",
           "AFG_BC_DIRICHLET", "AFG_BC_NEUMANN"]


def _i32(a):
    return a.ctypes.data_as(_lib.I32P)


def _f64(a):
    return a.ctypes.data_as(_lib.F64P)


class DeviceTree:
    """A tree whose cell data lives in device memory (one afg_ctx)."""

    def __init__(self, dim, n_cell, n_var, records, dr_base, tree_of=None):
        L = _lib.load()
        h = ctypes.c_void_p()
        check(L.afg_ctx_create(dim, n_cell, n_var, ctypes.byref(h)))
        self.h = h
        self._ptrace_cap = 0
        self.dim, self.nc, self.n_var = dim, n_cell, n_var
        self.records = np.ascontiguousarray(records, dtype=np.int32)
        self.ids = np.ascontiguousarray(self.records[:, 0])
        if tree_of is None:
            check(L.afg_set_tree(self.h, self.records.shape[0], _i32(self.records), float(dr_base)), self.h)
        else:
            self.tree_of = np.ascontiguousarray(tree_of, dtype=np.int32)
            check(L.afg_set_forest(self.h, self.records.shape[0], _i32(self.records), float(dr_base),
                                   _i32(self.tree_of)), self.h)
        self.dr_base = float(dr_base)
        self._tree_info()

    def _tree_info(self):
        L = _lib.load()
        lo, hi = ctypes.c_int(), ctypes.c_int()
        check(L.afg_levels(self.h, ctypes.byref(lo), ctypes.byref(hi)), self.h)
        self.min_lvl, self.highest_lvl = lo.value, hi.value
        n = ctypes.c_int64()
        check(L.afg_packed_size(self.h, self.ids.size, _i32(self.ids), ctypes.byref(n)), self.h)
        self.n_values = n.value

    def set_coord(self, cyl, r_base=0.0):
        """Cylindrical coordinates (2D; x the radius; the auto operators become
        mg2_box_clpl / clpld) or Cartesian (afg_set_coord)."""
        self._c(_lib.load().afg_set_coord(self.h, 1 if cyl else 0, float(r_base)))

    # ghost cells and transfers of any variable
    def gc_tree(self, iv, rb=0, bc=0):
        """a$D_gc_tree: rb 0 mg sides_rb, 1 gc_prolong0, 2 gc_interp, 3
        gc_interp_lim; bc 0 the boundary table, AFG_BC_DIRICHLET /
        AFG_BC_NEUMANN that condition with b = 0."""
        self._c(_lib.load().afg_gc_tree(self.h, int(iv), int(rb), int(bc)))

    def gc_level(self, iv, lvl, rb=0, bc=0):
        self._c(_lib.load().afg_gc_level(self.h, int(iv), int(lvl), int(rb), int(bc)))

    def gc2(self, iv, ids=None):
        """a$D_gc2_box of variable iv for every box (or ids): per box
        gc_data(nc_l[, nc_l], 2*dim) in Fortran order, concatenated."""
        from .treefile import level_ncell
        ids = self.ids if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        lvl_of = dict(zip(self.records[:, 0].tolist(), self.records[:, 1].tolist()))
        nc_l = level_ncell(self.nc, np.array([lvl_of.get(int(i), 1) for i in ids], dtype=np.int64))
        out = np.empty(int((2 * self.dim * nc_l ** (self.dim - 1)).sum()), dtype=np.float64)
        self._c(_lib.load().afg_gc2_boxes(self.h, int(iv), ids.size, _i32(ids), _f64(out)))
        return out

    def restrict_tree(self, iv, iv_to=None):
        self._c(_lib.load().afg_restrict_tree(self.h, int(iv), int(iv if iv_to is None else iv_to)))

    def prolong_level(self, iv, iv_to, lvl, order, add=False):
        self._c(_lib.load().afg_prolong_level(self.h, int(iv), int(iv_to), int(lvl), int(order), int(add)))

    def n_face_values(self):
        """Values of one face variable over all boxes (fx, fy (, fz) per box)."""
        from .treefile import level_ncell
        n = level_ncell(self.nc, self.records[:, 1])
        return int((self.dim * (n + 1) * n ** (self.dim - 1)).sum())

    def upload_fc(self, ifc, packed):
        """Face variable ifc of every box (record order), fx, fy (, fz) per box."""
        a = np.ascontiguousarray(packed, dtype=np.float64)
        assert a.size == self.n_face_values()
        self._c(_lib.load().afg_upload_fc(self.h, int(ifc), self.ids.size, _i32(self.ids), _f64(a)))

    def download_fc(self, ifc):
        out = np.empty(self.n_face_values(), dtype=np.float64)
        self._c(_lib.load().afg_download_fc(self.h, int(ifc), self.ids.size, _i32(self.ids), _f64(out)))
        return out

    def consistent_fluxes(self, f_ixs):
        f = np.ascontiguousarray(list(f_ixs), dtype=np.int32)
        self._c(_lib.load().afg_consistent_fluxes(self.h, f.size, _i32(f)))

    def add_cc(self, iv_from, iv_to):
        self._c(_lib.load().afg_add_cc(self.h, int(iv_from), int(iv_to)))

    def sub_cc(self, iv_from, iv_to):
        self._c(_lib.load().afg_sub_cc(self.h, int(iv_from), int(iv_to)))

    def times_cc(self, a, iv):
        self._c(_lib.load().afg_times_cc(self.h, float(a), int(iv)))

    def lincomb_cc(self, a, iv_a, b, iv_b):
        self._c(_lib.load().afg_lincomb_cc(self.h, float(a), int(iv_a), float(b), int(iv_b)))

    def update_tree(self, records, prolong_order=2, vars=(1,), tree_of=None, dr_base=None):
        """Switch to the refined / derefined tree `records` keeping the data
        on the device (afg_update_tree): kept boxes keep all their values,
        new boxes get the variables `vars` prolonged from their parents
        (a$D_prolong0/1/2 for prolong_order 0/1/2, -1 for none)."""
        recs = np.ascontiguousarray(records, dtype=np.int32)
        vs = np.ascontiguousarray(list(vars), dtype=np.int32)
        t_of = None if tree_of is None else np.ascontiguousarray(tree_of, dtype=np.int32)
        dr = self.dr_base if dr_base is None else float(dr_base)
        rc = _lib.load().afg_update_tree(self.h, recs.shape[0], _i32(recs), dr,
                                          None if t_of is None else _i32(t_of), int(prolong_order),
                                          vs.size, _i32(vs))
        self._c(rc)
        self.records = recs
        self.ids = np.ascontiguousarray(recs[:, 0])
        if t_of is not None:
            self.tree_of = t_of
        self.dr_base = dr
        self._tree_info()

    def update_forest(self, records, tree_of, old_tree, dr_base=None):
        """Switch to another forest keeping the trees it shares with this one
        (afg_update_forest): tree t of the new forest is tree old_tree[t] of
        the current one (its boxes keep all values), or new (-1: zeros)."""
        recs = np.ascontiguousarray(records, dtype=np.int32)
        t_of = np.ascontiguousarray(tree_of, dtype=np.int32)
        old = np.ascontiguousarray(old_tree, dtype=np.int32)
        dr = self.dr_base if dr_base is None else float(dr_base)
        self._c(_lib.load().afg_update_forest(self.h, recs.shape[0], _i32(recs), dr, _i32(t_of), _i32(old)))
        self.records = recs
        self.ids = np.ascontiguousarray(recs[:, 0])
        self.tree_of = t_of
        self.dr_base = dr
        self._tree_info()

    def mixed_fp64_base(self, enabled):
        """The FP32 stages' coarsest-grid relaxation in FP64 arithmetic
        (afg_mixed_fp64_base)."""
        self._c(_lib.load().afg_mixed_fp64_base(self.h, 1 if enabled else 0))

    def mixed_fp64_tau(self, enabled):
        """The FP32 stages' update_coarse (the FAS tau) in FP64 arithmetic
        (afg_mixed_fp64_tau)."""
        self._c(_lib.load().afg_mixed_fp64_tau(self.h, 1 if enabled else 0))

    def mixed_first_fmg(self, enabled):
        """Whether the first mixed-precision correction after mixed_begin
        starts with an FMG (afg_mixed_first_fmg; default True)."""
        self._c(_lib.load().afg_mixed_first_fmg(self.h, 1 if enabled else 0))

    @classmethod
    def from_state(cls, st: TreeState, n_var=None, tree_of=None):
        t = cls(st.dim, st.nc, n_var or max(st.n_var, len(st.data)), st.recs, st.dr_base, tree_of)
        for iv, a in enumerate(st.data, start=1):
            t.upload(iv, a)
        return t

    @classmethod
    def from_forest(cls, states, n_var=None):
        """One context holding the trees `states` (afg_set_forest)."""
        from . import forest
        st, tree_of = forest.combine(states)
        return cls.from_state(st, n_var, tree_of)

    def level_ids(self, lvl):
        """The box ids of level lvl in the device's slot order (afg_level_ids)."""
        L = _lib.load()
        n = ctypes.c_int()
        self._c(L.afg_level_ids(self.h, int(lvl), 0, None, ctypes.byref(n)))
        ids = np.zeros(n.value, dtype=np.int32)
        self._c(L.afg_level_ids(self.h, int(lvl), ids.size, _i32(ids), ctypes.byref(n)))
        return ids

    # forests
    def forest_trees(self):
        n = ctypes.c_int()
        self._c(_lib.load().afg_forest_trees(self.h, ctypes.byref(n), None))
        hi = np.zeros(n.value, dtype=np.int32)
        self._c(_lib.load().afg_forest_trees(self.h, ctypes.byref(n), _i32(hi)))
        return hi

    def forest_max_abs(self, iv):
        """Per tree: max |variable iv| over its leaf interiors."""
        n = self.forest_trees().size
        out = np.zeros(n)
        self._c(_lib.load().afg_forest_max_abs(self.h, iv, n, _f64(out)))
        return out

    def set_tile(self, lvl, shape, fp32=False):
        """Tile shape (0, 1, 2) of the colour pass on level lvl."""
        self._c(_lib.load().afg_set_tile(self.h, int(lvl), int(fp32), int(shape)))

    def tile(self, lvl, fp32=False):
        v = ctypes.c_int()
        self._c(_lib.load().afg_get_tile(self.h, int(lvl), int(fp32), ctypes.byref(v)))
        return v.value

    def autotune(self, mg, fp32=False, reps=8):
        """Pick the fastest tile shape and pass mode per level and the
        fastest bottom cell limit; returns (ms per two passes as an array
        [level - min_lvl, mode (0 single passes, 1 fused pair), shape],
        ms per V-cycle for the limits 512, 2048, 8192, 32768); -1 where not
        timed."""
        n = self.highest_lvl - self.min_lvl + 1
        ms = np.zeros(6 * n + 4)
        self._c(_lib.load().afg_autotune(self.h, ctypes.byref(mg), int(fp32), int(reps), _f64(ms)))
        return ms[:6 * n].reshape(n, 2, 3), ms[6 * n:]

    def pass_pairs(self, lvl, fp32=False):
        v = ctypes.c_int()
        self._c(_lib.load().afg_get_pass_pairs(self.h, int(lvl), int(fp32), ctypes.byref(v)))
        return bool(v.value)

    def set_level_pass_pairs(self, lvl, enabled, fp32=False):
        """Fused pass pairs (True) or single passes on one level."""
        self._c(_lib.load().afg_set_level_pass_pairs(self.h, int(lvl), int(fp32), 1 if enabled else 0))

    def bottom_cells(self, fp32=False):
        v = ctypes.c_int64()
        self._c(_lib.load().afg_get_bottom_cells(self.h, int(fp32), ctypes.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            _lib.load().afg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _c(self, rc):
        check(rc, self.h)

    # data movement
    def upload(self, iv, packed, ids=None):
        ids = self.ids if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        a = np.ascontiguousarray(packed, dtype=np.float64)
        self._c(_lib.load().afg_upload_cc(self.h, iv, ids.size, _i32(ids), _f64(a)))

    def download(self, iv, ids=None):
        ids = self.ids if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        n = self.n_values if ids is self.ids else self.packed_size(ids)
        out = np.empty(n, dtype=np.float64)
        self._c(_lib.load().afg_download_cc(self.h, iv, ids.size, _i32(ids), _f64(out)))
        return out

    def packed_size(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        n = ctypes.c_int64()
        self._c(_lib.load().afg_packed_size(self.h, ids.size, _i32(ids), ctypes.byref(n)))
        return n.value

    def fp32_from_cc(self, iv, iv32):
        self._c(_lib.load().afg_fp32_from_cc(self.h, iv, iv32))

    def fp32_to_cc(self, iv32, iv, add=False):
        self._c(_lib.load().afg_fp32_to_cc(self.h, iv32, iv, int(add)))

    def copy_cc(self, iv_from, iv_to):
        self._c(_lib.load().afg_copy_cc(self.h, iv_from, iv_to))

    def set_cc(self, iv, value):
        self._c(_lib.load().afg_set_cc(self.h, iv, float(value)))

    def set_bc(self, id_, nb, bc_type, values=None):
        v = None
        if values is not None:
            arr = np.ascontiguousarray(values, dtype=np.float64)
            v = _f64(arr)
        self._c(_lib.load().afg_set_bc(self.h, int(id_), int(nb), int(bc_type), v))

    def sync(self):
        self._c(_lib.load().afg_sync(self.h))

    # single stages
    def fill_gc(self, iv, lvl):
        self._c(_lib.load().afg_fill_gc(self.h, iv, lvl))

    def gsrb_level(self, mg, lvl, n_cycle):
        self._c(_lib.load().afg_gsrb_level(self.h, ctypes.byref(mg), lvl, n_cycle))

    def residual_level(self, mg, lvl):
        self._c(_lib.load().afg_residual_level(self.h, ctypes.byref(mg), lvl))

    def restrict_level(self, iv, lvl):
        self._c(_lib.load().afg_restrict_level(self.h, iv, lvl))

    def prolong2_level(self, iv, iv_to, lvl, add):
        self._c(_lib.load().afg_prolong2_level(self.h, iv, iv_to, lvl, int(add)))

    def correct_children(self, mg, lvl):
        self._c(_lib.load().afg_correct_children(self.h, ctypes.byref(mg), lvl))

    def update_coarse(self, mg, lvl):
        self._c(_lib.load().afg_update_coarse(self.h, ctypes.byref(mg), lvl))

    # controls and measurements
    def set_bottom_cells(self, cells, wave_cells=0):
        self._c(_lib.load().afg_set_bottom_cells(self.h, int(cells), int(wave_cells)))

    def set_split_passes(self, enabled):
        """Relaxation blocks on a colour-split working copy of phi (3D tiles)."""
        self._c(_lib.load().afg_set_split_passes(self.h, int(bool(enabled))))

    def set_persist(self, wgs, cells=32768):
        """The persistent kernel for the levels of at most `cells` cells each
        (wgs workgroups on one XCD; 0 turns it off)."""
        self._c(_lib.load().afg_set_persist(self.h, int(wgs), int(cells)))

    def persist_trace(self, cap):
        """Ticks (10 ns) of the persistent kernel's stage trace since the last
        call; record up to `cap` per launch from now on (0: stop)."""
        ticks = np.zeros(max(self._ptrace_cap, 1), dtype=np.int64)
        n = ctypes.c_int(0)
        self._c(_lib.load().afg_persist_trace(self.h, int(cap), ticks.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                              ctypes.byref(n)))
        self._ptrace_cap = int(cap)
        return ticks[:n.value].copy()

    def persist(self):
        """(workgroups, cell limit, highest level taken) of the persistent kernel."""
        w, n, top = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int()
        self._c(_lib.load().afg_get_persist(self.h, ctypes.byref(w), ctypes.byref(n), ctypes.byref(top)))
        return w.value, n.value, top.value

    def set_pass_pairs(self, enabled):
        """Fused pairs of colour passes (default) or one pass per launch."""
        self._c(_lib.load().afg_set_pass_pairs(self.h, int(enabled)))

    def set_graphs(self, enabled):
        self._c(_lib.load().afg_set_graphs(self.h, int(enabled)))

    def set_timing(self, enabled, lvl=-1000):
        """Time the GSRB passes of level `lvl` (all levels by default)."""
        self._c(_lib.load().afg_set_timing(self.h, int(enabled), int(lvl)))

    def timing(self):
        n, ms, b = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        self._c(_lib.load().afg_timing(self.h, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)))
        return n.value, ms.value, b.value

    def vcycle_cell_updates(self, mg, highest_lvl=None):
        out = ctypes.c_double()
        hl = self.highest_lvl if highest_lvl is None else highest_lvl
        self._c(_lib.load().afg_vcycle_cell_updates(self.h, ctypes.byref(mg), hl, ctypes.byref(out)))
        return out.value

    def device_bytes(self):
        out = ctypes.c_int64()
        self._c(_lib.load().afg_device_bytes(self.h, ctypes.byref(out)))
        return out.value


def mg_init_mg(mg: MgOpts) -> MgOpts:
    """mg$D_init_mg (m_aX_mg.f90:130-154)."""
    check(_lib.load().afg_mg_init(ctypes.byref(mg)))
    return mg


def mg_fas_fmg(tree: DeviceTree, mg: MgOpts, set_residual: bool, have_guess: bool):
    """mg$D_fas_fmg (m_aX_mg.f90:164-203)."""
    tree._c(_lib.load().afg_fas_fmg(tree.h, ctypes.byref(mg), int(set_residual), int(have_guess)))


def mg_fas_vcycle(tree: DeviceTree, mg: MgOpts, highest_lvl: int, set_residual: bool):
    """mg$D_fas_vcycle (m_aX_mg.f90:208-255)."""
    tree._c(_lib.load().afg_fas_vcycle(tree.h, ctypes.byref(mg), highest_lvl, int(set_residual)))


def tree_max_cc(tree: DeviceTree, iv: int, include_parents: bool = False) -> float:
    """a$D_tree_max_cc (m_aX_utils.f90:468-518)."""
    out = ctypes.c_double()
    tree._c(_lib.load().afg_tree_max_cc(tree.h, iv, int(include_parents), ctypes.byref(out)))
    return out.value


def tree_min_cc(tree: DeviceTree, iv: int, include_parents: bool = False) -> float:
    """a$D_tree_min_cc (m_aX_utils.f90:522-571)."""
    out = ctypes.c_double()
    tree._c(_lib.load().afg_tree_min_cc(tree.h, iv, int(include_parents), ctypes.byref(out)))
    return out.value


def mg_fas_fmg_fp32(tree: DeviceTree, mg: MgOpts, set_residual: bool, have_guess: bool):
    """mg$D_fas_fmg on the FP32 copies of the variables."""
    tree._c(_lib.load().afg_fas_fmg_fp32(tree.h, ctypes.byref(mg), int(set_residual), int(have_guess)))


def mg_fas_vcycle_fp32(tree: DeviceTree, mg: MgOpts, highest_lvl: int, set_residual: bool):
    """mg$D_fas_vcycle on the FP32 copies of the variables."""
    tree._c(_lib.load().afg_fas_vcycle_fp32(tree.h, ctypes.byref(mg), highest_lvl, int(set_residual)))


def mixed_begin(tree: DeviceTree, mg: MgOpts, have_guess=False):
    """Start FP64 defect correction: phi = 0 (unless have_guess), FP64
    residual in tmp."""
    tree._c(_lib.load().afg_mixed_begin(tree.h, ctypes.byref(mg), int(have_guess)))


def mixed_correct(tree: DeviceTree, mg: MgOpts, cycles=2):
    """One FP32 multigrid correction (`cycles` V-cycles; the first after
    mixed_begin starts with an FMG), phi += e, new FP64 residual in tmp.
    Asynchronous: check the residual with forest_max_abs."""
    tree._c(_lib.load().afg_mixed_correct(tree.h, ctypes.byref(mg), int(cycles)))


def mixed_solve(tree: DeviceTree, mg: MgOpts, tol: float, max_iter: int = 50, have_guess=False,
                cycles=2):
    """FP64 defect correction with FP32 multigrid corrections (`cycles`
    V-cycles each) until every tree's relative residual is <= `tol`; returns
    (corrections, largest relative residual over the trees)."""
    it, rr = ctypes.c_int(), ctypes.c_double()
    tree._c(_lib.load().afg_mixed_solve(tree.h, ctypes.byref(mg), int(have_guess), float(tol),
                                        int(max_iter), int(cycles), ctypes.byref(it), ctypes.byref(rr)))
    return it.value, rr.value


def max_abs_cc(tree: DeviceTree, iv: int) -> float:
    """max(|min|, |max|) over leaf interiors, as examples/poisson_basic_2d.f90:90-92."""
    return max(abs(tree_min_cc(tree, iv)), abs(tree_max_cc(tree, iv)))


def default_mg(eps: bool = False, n_down=2, n_up=2, n_base=2, auto: bool = False, lsf: bool = False,
               lsf_set: bool = False, rstr: int = AFG_RSTR_LPL):
    """Options with phi, rhs, tmp (and eps) in variables 1, 2, 3 (4), as the
    problems of afivo_amd.problems store them; the auto operators when eps,
    lsf or auto (auto without eps or level set: every box untagged, as for a
    cylindrical tree in examples/poisson_cyl.f90).  lsf: level set and
    boundary value in variables 5 and 6; lsf_set: the lpllsf operators for
    every box (implies lsf); rstr: the restriction (AFG_RSTR_*)."""
    lsf = lsf or lsf_set
    op = AFG_OP_LSF if lsf_set else AFG_OP_AUTO if (eps or auto or lsf) else AFG_OP_LPL
    mg = MgOpts(i_phi=1, i_rhs=2, i_tmp=3, i_eps=4 if eps else -1, n_cycle_down=n_down,
                n_cycle_up=n_up, n_cycle_base=n_base, op=op, i_lsf=5 if lsf else -1,
                i_bval=6 if lsf else -1, rstr=rstr)
    return mg_init_mg(mg)


# ---------------------------------------------------------------------------
# Per-box operators on the host (afg_box_*): the public box procedures of
# m_a$D_mg applied to one box held in host memory.  `box` arrays are the
# whole cc(0:nc+1, ...) array of one variable (any shape with (nc+2)^dim
# float64 values, x fastest); kind is AFG_BOX_LPL, _LPLD (eps), _LSF (lsf,
# bval), _CLPL or _CLPLD (2D cylindrical, ir0 = (ix(1) - 1) * nc).

def _box(a, n):
    if a is None:
        return None
    a = np.asarray(a)
    if a.dtype != np.float64 or not a.flags.c_contiguous or a.size != n:
        raise ValueError(f"box array must be {n} contiguous float64 values")
    return a


def _opt(a):
    return None if a is None else _f64(a)


def box_op(dim, nc, dr, kind, phi, out, eps=None, lsf=None, bval=None, ir0=0):
    """out(1:nc, ...) = A phi: mg$D_box_lpl and its families (m_aX_mg.f90:763-797,
    866-916, 1183-1253, 1354-1407); out's ghost cells are left as they are."""
    n = (nc + 2) ** dim
    phi, out, eps, lsf, bval = (_box(a, n) for a in (phi, out, eps, lsf, bval))
    check(_lib.load().afg_box_op(dim, nc, dr, kind, ir0, _f64(phi), _opt(eps), _opt(lsf), _opt(bval), _f64(out)))


def box_gsrb(dim, nc, dr, kind, redblack_cntr, phi, rhs, eps=None, lsf=None, bval=None, ir0=0):
    """One colour of Gauss-Seidel red-black on phi, in place: mg$D_box_gsrb_lpl
    and its families (m_aX_mg.f90:718-760, 810-863, 1103-1180, 1325-1351, 1411-1443)."""
    n = (nc + 2) ** dim
    phi, rhs, eps, lsf, bval = (_box(a, n) for a in (phi, rhs, eps, lsf, bval))
    check(_lib.load().afg_box_gsrb(dim, nc, dr, kind, ir0, redblack_cntr, _f64(phi), _f64(rhs), _opt(eps),
                                   _opt(lsf), _opt(bval)))


def _offset(offset):
    o = np.zeros(3, dtype=np.int32)
    o[:len(offset)] = offset
    return o


def box_corr(dim, nc, pnc, kind, offset, p_corr, c_phi, p_eps=None, p_lsf=None, c_lsf=None):
    """Correction of a child's phi (in place) from its parent's correction
    p_corr (the parent's tmp after tmp = phi - tmp): mg$D_box_corr_lpl,
    _lpld, _lpllsf (m_aX_mg.f90:708-715, 918-985, 1020-1100).  offset:
    a$D_get_child_offset of the child (m_aX_t.f90:362-374)."""
    n, m = (nc + 2) ** dim, (pnc + 2) ** dim
    c_phi, c_lsf = _box(c_phi, n), _box(c_lsf, n)
    p_corr, p_eps, p_lsf = _box(p_corr, m), _box(p_eps, m), _box(p_lsf, m)
    o = _offset(offset)
    check(_lib.load().afg_box_corr(dim, nc, pnc, kind, _i32(o), _f64(p_corr), _opt(p_eps), _opt(p_lsf),
                                   _f64(c_phi), _opt(c_lsf)))


def box_rstr(dim, nc, pnc, kind, offset, c_var, p_var, c_lsf=None, p_lsf=None, p_rmin=0.0, c_dr=0.0):
    """Restriction of a child's variable into its parent's (in place):
    mg$D_box_rstr_lpl = a$D_restrict_box (m_aX_restrict.f90:67-132; kind
    AFG_BOX_CLPL for a cylindrical parent, with p_rmin = box_p%r_min(1) and
    c_dr = box_c%dr) and mg$D_box_rstr_lpllsf (m_aX_mg.f90:1256-1321)."""
    n, m = (nc + 2) ** dim, (pnc + 2) ** dim
    c_var, c_lsf = _box(c_var, n), _box(c_lsf, n)
    p_var, p_lsf = _box(p_var, m), _box(p_lsf, m)
    o = _offset(offset)
    check(_lib.load().afg_box_rstr(dim, nc, pnc, kind, _i32(o), _f64(c_var), _opt(c_lsf), _opt(p_lsf),
                                   p_rmin, c_dr, _f64(p_var)))


def sides_rb(dim, nc, pnc, nb, offset, p_var, c_var):
    """mg$D_sides_rb (m_aX_mg.f90:262-350): the ghost cells of face nb
    (1-based) of a child at a refinement boundary, from its parent's array."""
    n, m = (nc + 2) ** dim, (pnc + 2) ** dim
    c_var, p_var = _box(c_var, n), _box(p_var, m)
    o = _offset(offset)
    check(_lib.load().afg_box_sides_rb(dim, nc, pnc, nb, _i32(o), _f64(p_var), _f64(c_var)))
