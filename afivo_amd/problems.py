"""Problem set-ups of the benchmark configurations.

The right-hand sides restate the reference's example module
``examples/m_gaussians.f90`` (``gauss_lpl`` :90-102, ``gauss_val`` :50-58) and the
tree-building loops of ``examples/poisson_basic_2d.f90:41-68`` /
``poisson_basic_3d.f90``.  All cell data starts at zero, ghost cells included
(the reference leaves ``cc`` uninitialised, ``src/m_aX_core.f90:361-369``), then
the rhs is set on the interior of every box.

Configurations (sizes as stated in the benchmark definition):

* ``c1`` plain FP64 2D Poisson: nc = 8, uniform to level 5, refined to level 7
  where dr^2 max|rhs| > 1e-4; two Gaussians (ampl 1, sigma 0.04) at
  (0.25, 0.25) and (0.75, 0.75); FMG then 10 V-cycles, 2+2 sweeps, 50 base.
* ``c2`` plain FP64 3D Poisson: nc = 8, uniform to level 4, refined to level 7
  around 4 Gaussians of amplitude +-1, sigma 0.03, centres uniform in
  [0.2, 0.8]^3 from the seed; threshold fixed at THR_C2; 8 V-cycles after FMG.
* ``c4`` variable-coefficient 3D Poisson: nc = 16, eps = 1 below z = 0.5 and
  eps_hi ~ U[2, 10] above.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import tree as treemod
from .treefile import TreeState

_exp = np.frompyfunc(math.exp, 1, 1)


def exp_libm(x):
    """exp through the C library, element by element (the routine the
    reference's Fortran exp calls), so right-hand sides are bit-identical."""
    x = np.asarray(x, dtype=np.float64)
    return _exp(x).astype(np.float64)


@dataclasses.dataclass
class Gaussians:
    """gauss_t (examples/m_gaussians.f90:8-15)."""
    ampl: np.ndarray
    sigma: np.ndarray
    r0: np.ndarray           # (n, dim)
    # exp of the C library element by element (bit-identical to the reference
    # harness), or numpy's vectorised exp (faster; for inputs that are not
    # compared with the reference)
    exact: bool = True

    def _exp(self, x):
        return exp_libm(x) if self.exact else np.exp(x)

    @property
    def dim(self):
        return self.r0.shape[1]

    def _sum_sq(self, r, n):
        xrel = (r - self.r0[n]) / self.sigma[n]
        s = xrel[..., 0] * xrel[..., 0]
        for d in range(1, self.dim):
            s = s + xrel[..., d] * xrel[..., d]
        return s

    def lpl(self, r):
        """gauss_lpl (m_gaussians.f90:90-102).  Note: the reference sums
        exp(-|x|^2) without the amplitude; kept as is."""
        r = np.asarray(r, dtype=np.float64)
        out = np.zeros(r.shape[:-1])
        for n in range(len(self.sigma)):
            s = self._sum_sq(r, n)
            g = self._exp(-s)
            out = out + 4 / (self.sigma[n] * self.sigma[n]) * (s - 0.5 * self.dim) * g
        return out

    def val(self, r):
        """gauss_val (m_gaussians.f90:50-58)."""
        r = np.asarray(r, dtype=np.float64)
        out = np.zeros(r.shape[:-1])
        for n in range(len(self.sigma)):
            out = out + self._exp(-self._sum_sq(r, n))
        return out


def gauss_refinement(gs: Gaussians, l_uniform: int, l_max: int, thr: float):
    """Refine below l_uniform everywhere and below l_max where
    dr^2 * max|gauss_lpl| over the box interior exceeds thr."""

    def ref(view: treemod.BoxView, cur):
        out = np.array(cur, dtype=np.int64, copy=True) if np.ndim(cur) else np.full(view.ids.size, cur)
        out[view.lvl < l_uniform] = treemod.REFINE
        cand = np.nonzero((view.lvl >= l_uniform) & (view.lvl < l_max))[0]
        # boxes of equal cell count are evaluated together (element-wise the
        # same arithmetic as one box at a time)
        for n in np.unique(view.n_cell[cand]):
            sel = cand[view.n_cell[cand] == n]
            idx = np.arange(1, int(n) + 1) - 0.5
            dim = view.ix.shape[1]
            grids = np.meshgrid(*([idx] * dim), indexing="ij")
            rc = np.stack([view.r_min[sel, d][:, None] + grids[d].ravel()[None, :] * view.dr[sel][:, None]
                           for d in range(dim)], axis=-1)
            rmax = np.abs(gs.lpl(rc)).max(axis=1)
            out[sel[view.dr[sel] ** 2 * rmax > thr]] = treemod.REFINE
        return out

    return ref


def set_rhs(t: treemod.Tree, st: TreeState, gs: Gaussians) -> np.ndarray:
    """Packed rhs: zero ghosts, gauss_lpl at interior cell centres of every box."""
    ids = st.ids
    view = t.box_view(ids)
    out = np.zeros(st.n_values())
    off = st.offsets()
    sizes = st.ncell()
    # boxes of equal cell count are evaluated together
    for n in np.unique(sizes):
        sel = np.nonzero(sizes == n)[0]
        idx = (np.arange(1, n + 1) - 0.5)
        dim = t.dim
        grids = np.meshgrid(*([idx] * dim), indexing="ij")
        rc = np.stack([view.r_min[sel, d][:, None] + (grids[d].ravel()[None, :] * view.dr[sel][:, None])
                       for d in range(dim)], axis=-1)
        # rc: (boxes, cells in (x, y, z) 'ij' order, dim)
        vals = gs.lpl(rc).reshape((sel.size,) + (int(n),) * dim)
        vals = np.transpose(vals, [0] + list(range(dim, 0, -1)))    # -> (box, z, y, x)
        m = int(n) + 2
        for k, r in enumerate(sel):
            box = np.zeros((m,) * dim)
            box[(slice(1, m - 1),) * dim] = vals[k]
            out[off[r]:off[r + 1]] = box.ravel()
    return out


def set_eps(st: TreeState, t: treemod.Tree, eps_hi: float, cut: float = 0.5) -> np.ndarray:
    """Packed eps: 1 where the last coordinate of the cell centre is below
    cut, eps_hi above, on interior and ghost cells."""
    view = t.box_view(st.ids)
    out = np.zeros(st.n_values())
    off = st.offsets()
    dim = t.dim
    for r, n in enumerate(st.ncell()):
        n = int(n)
        z = view.r_min[r, dim - 1] + (np.arange(0, n + 2) - 0.5) * view.dr[r]
        e = np.where(z < cut, 1.0, eps_hi)
        shape = [n + 2] * dim
        box = np.broadcast_to(e.reshape([n + 2] + [1] * (dim - 1)), shape)
        out[off[r]:off[r + 1]] = np.ascontiguousarray(box).ravel()
    return out


def set_level_set(st: TreeState, t: treemod.Tree, centre, radius: float):
    """Packed level set |r - centre| - radius and boundary value 1 + x (the
    value on the zero contour) on interior and ghost cells, as a program sets
    mg%i_lsf and mg%i_bval for the lpllsf operators."""
    view = t.box_view(st.ids)
    lsf = np.zeros(st.n_values())
    bval = np.zeros(st.n_values())
    off = st.offsets()
    dim = t.dim
    for r, n in enumerate(st.ncell()):
        n = int(n)
        # coordinate d of the cell centres varies along numpy axis dim-1-d (x fastest)
        ax = [view.r_min[r, d] + (np.arange(0, n + 2) - 0.5) * view.dr[r] for d in range(dim)]
        grid = np.meshgrid(*ax[::-1], indexing="ij")[::-1]
        dist = np.sqrt(sum((grid[d] - centre[d]) ** 2 for d in range(dim)))
        lsf[off[r]:off[r + 1]] = (dist - radius).ravel()
        bval[off[r]:off[r + 1]] = (1.0 + grid[0]).ravel()
    return lsf, bval


@dataclasses.dataclass
class Problem:
    name: str
    tree: treemod.Tree
    state: TreeState           # topology + packed [phi, rhs, tmp(, eps)]
    gs: Gaussians
    n_down: int = 2
    n_up: int = 2
    n_base: int = 2
    n_vcycle: int = 8
    eps_hi: float | None = None
    l_uniform: int = 0
    l_max: int = 0
    thr: float = 0.0

    @property
    def n_var(self):
        return len(self.state.data)


def make_problem(name, dim, nc, gs, l_uniform, l_max, thr, n_base, n_vcycle, eps_hi=None,
                 n_down=2, n_up=2):
    t = treemod.build(dim, nc, gauss_refinement(gs, l_uniform, l_max, thr))
    st = t.state()
    rhs = set_rhs(t, st, gs)
    zeros = np.zeros_like(rhs)
    st.data = [zeros.copy(), rhs, zeros.copy()]
    if eps_hi is not None:
        st.data.append(set_eps(st, t, eps_hi))
    st.n_var = len(st.data)
    return Problem(name, t, st, gs, n_down, n_up, n_base, n_vcycle, eps_hi, l_uniform, l_max, thr)


C1_GAUSS = Gaussians(np.array([1.0, 1.0]), np.array([0.04, 0.04]),
                     np.array([[0.25, 0.25], [0.75, 0.75]]))

# Threshold of the 3D adaptive configuration, tuned once for 1e6-4e6 cells
# and frozen.
THR_C2 = 1.0e-3


def c2_gaussians(seed=1):
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(0.2, 0.8, size=(4, 3))
    return Gaussians(np.array([1.0, -1.0, 1.0, -1.0]), np.full(4, 0.03), r0)


def c4_gaussians(seed=1):
    rng = np.random.default_rng(seed)
    r0 = rng.uniform(0.2, 0.8, size=(6, 3))
    ampl = np.where(rng.uniform(size=6) < 0.5, 1.0, -1.0)
    # numpy's exp: this configuration is compared with FP64 solves of the
    # same rhs, not bit for bit with the reference
    return Gaussians(ampl, np.full(6, 0.04), r0, exact=False)


def c4_eps(seed=1):
    """eps above z = 0.5, drawn from U[2, 10]."""
    return float(np.random.default_rng(seed + 1000).uniform(2.0, 10.0))


# Threshold of the 3D variable-eps configuration (n_cell 16, uniform to level
# 3, adaptive to level 8), tuned once for about 1e7 leaf cells and frozen.
THR_C4 = 4.0e-3


def c4(seed=1, thr=THR_C4, l_max=8):
    """3D variable-coefficient Poisson: n_cell 16, uniform to level 3,
    adaptive to level `l_max` around 6 Gaussians, eps = 1 below z = 0.5 and
    U[2, 10] above (interior and ghost cells), auto operators."""
    return make_problem("c4", 3, 16, c4_gaussians(seed), 3, l_max, thr, n_base=50, n_vcycle=8,
                        eps_hi=c4_eps(seed))


def c1():
    return make_problem("c1", 2, 8, C1_GAUSS, 5, 7, 1e-4, n_base=50, n_vcycle=10)


def c2(seed=1, thr=THR_C2, l_max=7):
    return make_problem("c2", 3, 8, c2_gaussians(seed), 4, l_max, thr, n_base=50, n_vcycle=8)


def small_2d(nc=8, l_uniform=3, l_max=5, thr=1e-3, n_base=4):
    return make_problem("small2d", 2, nc, C1_GAUSS, l_uniform, l_max, thr, n_base=n_base, n_vcycle=3)


def small_3d(nc=8, l_uniform=2, l_max=4, thr=1e-3, n_base=4, eps_hi=None):
    gs = Gaussians(np.array([1.0, -1.0]), np.array([0.1, 0.1]),
                   np.array([[0.3, 0.3, 0.4], [0.7, 0.6, 0.5]]))
    return make_problem("small3d", 3, nc, gs, l_uniform, l_max, thr, n_base=n_base, n_vcycle=3,
                        eps_hi=eps_hi)


def harness_cfg(p: Problem, **extra):
    """Namelist values for the reference harness describing the same problem."""
    gs = p.gs
    cfg = dict(nc=p.tree.nc, ref_mode="gauss", l_uniform=p.l_uniform, l_max=p.l_max,
               thr=float(p.thr), n_gauss=len(gs.sigma), ampl=list(gs.ampl),
               sigma=list(gs.sigma), r0=list(gs.r0.ravel()), n_down=p.n_down, n_up=p.n_up,
               n_base=p.n_base, n_vcycle=p.n_vcycle, out_file=p.name)
    if p.eps_hi is not None:
        cfg.update(use_eps=True, eps_hi=float(p.eps_hi))
    cfg.update(extra)
    return cfg
