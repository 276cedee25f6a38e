"""ctypes binding of the C restatement (liboracle.so) and runner of the
reference harness (_ref/ref_harness_{2,3}d).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


class MgOpts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "i_phi", "i_rhs", "i_tmp", "i_eps",
        "n_cycle_down", "n_cycle_up", "n_cycle_base", "op", "i_lsf", "i_bval", "rstr")]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i32p = ctypes.POINTER(ctypes.c_int32)
        f64p = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = P
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_int, i32p]
        L.orc_destroy.argtypes = [P]
        L.orc_n_values.argtypes = [P]
        L.orc_set_cc.argtypes = [P, ctypes.c_int, f64p]
        L.orc_get_cc.argtypes = [P, ctypes.c_int, f64p]
        L.orc_set_bc.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, f64p]
        L.orc_fas_fmg.argtypes = [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int]
        L.orc_fas_vcycle.argtypes = [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int]
        L.orc_tree_max_cc.restype = ctypes.c_double
        L.orc_tree_max_cc.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_tree_min_cc.restype = ctypes.c_double
        L.orc_tree_min_cc.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_fill_gc_level.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_gsrb_level.argtypes = [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int]
        L.orc_residual_level.argtypes = [P, ctypes.POINTER(MgOpts), ctypes.c_int]
        L.orc_restrict_level.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_prolong2_level.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_correct_children_level.argtypes = [P, ctypes.POINTER(MgOpts), ctypes.c_int]
        L.orc_gc_tree.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_set_coord.argtypes = [P, ctypes.c_int, ctypes.c_double]
        L.orc_n_face_values.argtypes = [P]
        L.orc_set_fc.argtypes = [P, ctypes.c_int, f64p]
        L.orc_get_fc.argtypes = [P, ctypes.c_int, f64p]
        L.orc_consistent_fluxes.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_restrict_tree.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.orc_prolong_level.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int]
        L.orc_update_coarse_level.argtypes = [P, ctypes.POINTER(MgOpts), ctypes.c_int]
        L.orc_set_tags.argtypes = [P, ctypes.POINTER(MgOpts)]
        L.orc_gc2.argtypes = [P, ctypes.c_int, f64p]
        L.orc_box_call.argtypes = [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int]
        _lib = L
    return _lib


def _i32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _f64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleTree:
    """A tree in the C restatement, loaded from a TreeState."""

    def __init__(self, st, n_var=None):
        L = lib()
        self.st = st
        self.n_var = n_var or max(st.n_var, len(st.data))
        self._recs = np.ascontiguousarray(st.recs, dtype=np.int32)
        self.h = L.orc_create(st.dim, st.nc, self.n_var, st.dr_base, st.n_boxes,
                              _i32(self._recs))
        if not self.h:
            raise ValueError("orc_create rejected the tree")
        self.nv = L.orc_n_values(self.h)
        for iv, a in enumerate(st.data, start=1):
            self.set(iv, a)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set(self, iv, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == self.nv
        lib().orc_set_cc(self.h, iv, _f64(a))

    def get(self, iv):
        a = np.empty(self.nv, dtype=np.float64)
        lib().orc_get_cc(self.h, iv, _f64(a))
        return a

    def set_bc(self, id_, nb, bc_type, values=None):
        v = None if values is None else _f64(np.ascontiguousarray(values, dtype=np.float64))
        if lib().orc_set_bc(self.h, id_, nb, bc_type, v):
            raise ValueError("orc_set_bc failed")

    def fas_fmg(self, opts, set_residual=True, have_guess=False):
        if lib().orc_fas_fmg(self.h, ctypes.byref(opts), int(set_residual), int(have_guess)):
            raise ValueError("orc_fas_fmg failed")

    def fas_vcycle(self, opts, highest_lvl=None, set_residual=True):
        hl = self.st.highest_lvl if highest_lvl is None else highest_lvl
        if lib().orc_fas_vcycle(self.h, ctypes.byref(opts), hl, int(set_residual)):
            raise ValueError("orc_fas_vcycle failed")

    def set_tags(self, opts):
        """mg$D_set_box_tag on every box (what mg$D_auto_rstr requires)."""
        if lib().orc_set_tags(self.h, ctypes.byref(opts)):
            raise ValueError("orc_set_tags: a box has both a level-set and an eps tag")

    def set_coord(self, cyl, r_base=0.0):
        """Cylindrical coordinates (2D; x the radius) or Cartesian."""
        if lib().orc_set_coord(self.h, int(cyl), float(r_base)):
            raise ValueError("orc_set_coord failed")

    def gc_tree(self, iv, rb, bc=0):
        """a$D_gc_tree: rb 0 mg sides_rb, 1 gc_prolong0, 2 gc_interp, 3
        gc_interp_lim; bc 0 the boundary table, BC_DIRICHLET / BC_NEUMANN
        that condition with b = 0."""
        if lib().orc_gc_tree(self.h, iv, rb, bc):
            raise ValueError("orc_gc_tree failed")

    def gc2(self, iv):
        """a$D_gc2_box (gc2_prolong1, bc2_neumann_zero) of every box, record
        order, per box gc_data(nc_l[, nc_l], 2*dim) in Fortran order."""
        st = self.st
        n = st.ncell()
        out = np.empty(int((2 * st.dim * n ** (st.dim - 1)).sum()), dtype=np.float64)
        lib().orc_gc2(self.h, iv, _f64(out))
        return out

    def restrict_tree(self, iv, iv_to=None):
        lib().orc_restrict_tree(self.h, iv, iv if iv_to is None else iv_to)

    def prolong_level(self, iv, iv_to, lvl, order, add=False):
        if lib().orc_prolong_level(self.h, iv, iv_to, lvl, order, int(add)):
            raise ValueError("orc_prolong_level failed")

    def set_fc(self, ifc, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == lib().orc_n_face_values(self.h)
        if lib().orc_set_fc(self.h, ifc, _f64(a)):
            raise ValueError("orc_set_fc failed")

    def get_fc(self, ifc):
        a = np.empty(lib().orc_n_face_values(self.h), dtype=np.float64)
        if lib().orc_get_fc(self.h, ifc, _f64(a)):
            raise ValueError("orc_get_fc failed")
        return a

    def consistent_fluxes(self, f_ixs):
        f = np.ascontiguousarray(list(f_ixs), dtype=np.intc)
        if lib().orc_consistent_fluxes(self.h, f.size, f.ctypes.data_as(ctypes.POINTER(ctypes.c_int))):
            raise ValueError("orc_consistent_fluxes failed")

    def max_cc(self, iv, include_parents=False):
        return lib().orc_tree_max_cc(self.h, iv, int(include_parents))

    def min_cc(self, iv, include_parents=False):
        return lib().orc_tree_min_cc(self.h, iv, int(include_parents))

    def max_abs(self, iv):
        return max(abs(self.min_cc(iv)), abs(self.max_cc(iv)))


BC_DIRICHLET, BC_NEUMANN = -1, -2             # ORC_BC_* (afivo's a5_bc_dirichlet / a5_bc_neumann)
RB_MG, RB_PROLONG0, RB_INTERP, RB_INTERP_LIM = 0, 1, 2, 3


OP_LPL, OP_AUTO, OP_LSF = 0, 1, 2             # operator sets: lpl, auto, lpllsf
RSTR_LPL, RSTR_AUTO, RSTR_LSF = 0, 1, 2       # box_rstr: rstr_lpl, auto_rstr, rstr_lpllsf


def default_opts(n_down=2, n_up=2, n_base=2, eps=False, auto=False, lsf=False, lsf_set=False,
                 rstr=RSTR_LPL):
    """Options of the lpl operators, or of the auto operators when eps (i_eps
    4), lsf (level set and boundary value in variables 5 and 6) or auto (no
    eps: every box untagged, as examples/poisson_cyl.f90); lsf_set selects
    the lpllsf operators for every box (and implies lsf)."""
    lsf = lsf or lsf_set
    op = OP_LSF if lsf_set else OP_AUTO if (eps or auto or lsf) else OP_LPL
    return MgOpts(1, 2, 3, 4 if eps else -1, n_down, n_up, n_base, op,
                  5 if lsf else -1, 6 if lsf else -1, rstr)


# ---------------------------------------------------------------------------
# reference harness


def ref_available(dim: int) -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"ref_harness_{dim}d"))


def _fmt(v):
    if isinstance(v, bool):
        return ".true." if v else ".false."
    if isinstance(v, str):
        return "'" + v + "'"
    if isinstance(v, (list, tuple, np.ndarray)):
        return ", ".join(_fmt(float(x)) if not isinstance(x, str) else _fmt(x)
                         for x in np.ravel(v))
    if isinstance(v, float):
        return repr(v).replace("e", "d") if "e" in repr(v) else repr(v) + "d0"
    return str(v)


def run_ref(dim: int, cfg: dict, workdir=None, threads=1, timeout=600, bind=False):
    """Run the reference harness with namelist values `cfg` on `threads`
    OpenMP threads (bound one per core, OMP_PROC_BIND=close and
    OMP_PLACES=cores, when `bind`).

    Returns (prefix, stdout) where prefix + '.in' / '.out' / '.res' /
    '.time' are the files it wrote."""
    exe = os.path.join(REF_DIR, f"ref_harness_{dim}d")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} missing: run `make -C oracle ref`")
    workdir = workdir or tempfile.mkdtemp(prefix="afg_ref_")
    prefix = os.path.join(workdir, cfg.get("out_file", "ref"))
    cfg = dict(cfg)
    cfg["out_file"] = prefix
    body = ",\n".join(f" {k}={_fmt(v)}" for k, v in cfg.items())
    nml = os.path.join(workdir, "cfg.nml")
    with open(nml, "w") as f:
        f.write("&cfg\n" + body + "\n/\n")
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    if bind:
        env["OMP_PROC_BIND"] = "close"
        env["OMP_PLACES"] = "cores"
    p = subprocess.run([exe, nml], capture_output=True, text=True, env=env,
                       timeout=timeout)
    if p.returncode != 0 or "STOP" in p.stdout or "STOP" in p.stderr:
        raise RuntimeError(f"reference harness failed: {p.stdout}\n{p.stderr}")
    return prefix, p.stdout


def read_res(prefix):
    return np.loadtxt(prefix + ".res", ndmin=1)


def read_time(prefix):
    v = np.loadtxt(prefix + ".time", ndmin=1)
    return int(v[0]), v[1:]
