"""Host-side quadtree/octree construction with afivo's semantics.

Restates the parts of the reference tree core that produce the structure
the multigrid path runs on:

* ``a$D_init`` / ``a$D_set_base`` (``src/m_aX_core.f90:24-235``): base boxes,
  plus the coarse copies on levels ``min_lvl..0`` when ``coarsen_to`` is
  given, each with one child in ``children(1)``;
* ``a$D_adjust_refinement`` (``src/m_aX_core.f90:452-541``) with
  ``consistent_ref_flags`` (2:1 balance, ``:628-725``), ``add_children``
  (``:752-778``), ``remove_children`` (``:728-749``) and ``set_neighbs`` /
  ``find_neighb`` (``:398-440``).

The refinement callback is vectorised: it receives a :class:`BoxView` of many
boxes and returns their flags (``REFINE`` / ``KEEP`` / ``DEREFINE``, the
reference's ``a5_do_ref`` / ``a5_keep_ref`` / ``a5_rm_ref``).  Results do not
depend on box ids, which may differ from the reference's; the set of boxes
and their links are the same.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .treefile import TreeState, level_ncell, record_len

REFINE, KEEP, DEREFINE = 1, 0, -1        # a5_do_ref, a5_keep_ref, a5_rm_ref
_A5_REFINE, _A5_DEREFINE = 2, -2         # internal flags (m_aX_core.f90)


def _child_dix(dim):
    """a$D_child_dix (m_aX_t.f90:64,99-101): offset of child c (0-based)."""
    return np.array([[(c >> d) & 1 for d in range(dim)] for c in range(1 << dim)], dtype=np.int64)


def _neighb_dix(dim):
    """a$D_neighb_dix: offset of neighbour direction f (0-based)."""
    out = np.zeros((2 * dim, dim), dtype=np.int64)
    for f in range(2 * dim):
        out[f, f // 2] = 1 if f % 2 else -1
    return out


def ix_to_ichild(ix):
    """a$D_ix_to_ichild (m_aX_t.f90:378-387), 0-based child index."""
    ix = np.asarray(ix)
    dim = ix.shape[-1]
    c = np.zeros(ix.shape[:-1], dtype=np.int64)
    for d in range(dim):
        c += (1 - (ix[..., d] & 1)) << d
    return c


@dataclasses.dataclass
class BoxView:
    """Geometry of a batch of boxes handed to refinement callbacks."""
    ids: np.ndarray
    lvl: np.ndarray
    ix: np.ndarray          # (n, dim)
    n_cell: np.ndarray
    dr: np.ndarray
    r_min: np.ndarray       # (n, dim)

    def cell_centers(self, i):
        """Coordinates of the cell centres of box i (a$D_r_cc), shape (nc,..,dim)
        in C order (z, y, x) of the cell index."""
        n = int(self.n_cell[i])
        dim = self.ix.shape[1]
        idx = (np.arange(1, n + 1) - 0.5) * self.dr[i]
        grids = np.meshgrid(*([idx] * dim), indexing="ij")   # (x, y, z) order
        pts = np.stack([self.r_min[i, d] + grids[d] for d in range(dim)], axis=-1)
        return np.transpose(pts, list(range(dim))[::-1] + [dim])


class Tree:
    """A quadtree (dim 2) or octree (dim 3) of boxes with n_cell^dim cells."""

    def __init__(self, dim, n_cell, dr, r_min=None, coarsen_to=-1, lvl_limit=30):
        if n_cell < 2 or n_cell % 2:
            raise ValueError("a_init: n_cell should be even and >= 2")
        self.dim = dim
        self.nc = n_cell
        self.dr_base = float(dr)
        self.r_base = np.zeros(dim) if r_min is None else np.asarray(r_min, dtype=float)
        self.lvl_limit = lvl_limit
        if coarsen_to > 0:
            # m_aX_core.f90:74-82
            min_lvl = 1 - int(round(np.log(n_cell / coarsen_to) / np.log(2.0)))
            if 2 ** (1 - min_lvl) * coarsen_to != n_cell:
                raise ValueError("a_set_base: cannot coarsen to given value")
        else:
            min_lvl = 1
        self.min_lvl = min_lvl
        nch, nnb = 1 << dim, 2 * dim
        cap = 64
        self.lvl = np.zeros(cap, dtype=np.int64)
        self.ix = np.zeros((cap, dim), dtype=np.int64)
        self.parent = np.zeros(cap, dtype=np.int64)
        self.children = np.zeros((cap, nch), dtype=np.int64)
        self.nbr = np.zeros((cap, nnb), dtype=np.int64)
        self.in_use = np.zeros(cap, dtype=bool)
        self.highest_id = 0
        self.highest_lvl = 0
        self.ready = False

    # -- storage ----------------------------------------------------------
    def _grow(self, need):
        cap = self.lvl.shape[0]
        if need < cap:
            return
        new = max(need + 1, 2 * cap)
        for name in ("lvl", "ix", "parent", "children", "nbr", "in_use"):
            a = getattr(self, name)
            b = np.zeros((new,) + a.shape[1:], dtype=a.dtype)
            b[:cap] = a
            setattr(self, name, b)

    def _new_ids(self, n):
        ids = np.arange(self.highest_id + 1, self.highest_id + 1 + n, dtype=np.int64)
        self._grow(self.highest_id + n)
        self.highest_id += n
        return ids

    def ids_at(self, lvl):
        return np.nonzero(self.in_use[:self.highest_id + 1] & (self.lvl[:self.highest_id + 1] == lvl))[0]

    def has_children(self, ids):
        return self.children[ids, 0] != 0

    def n_cell_of(self, lvl):
        return level_ncell(self.nc, lvl)

    def box_dr(self, lvl):
        return self.dr_base * 0.5 ** (np.asarray(lvl, dtype=np.float64) - 1)

    def box_view(self, ids):
        ids = np.asarray(ids, dtype=np.int64)
        lvl = self.lvl[ids]
        dr = self.box_dr(lvl)
        # r_min: set_base (m_aX_core.f90:197-198) and add_children (:765-766)
        # both give r_base + (ix - 1) * dr * n_cell on levels >= 1; coarse
        # copies keep the base box origin.
        lv = np.maximum(lvl, 1)
        drl = self.dr_base * 0.5 ** (lv - 1.0)
        r_min = self.r_base + (self.ix[ids] - 1) * (drl * self.nc)[:, None]
        return BoxView(ids, lvl, self.ix[ids], self.n_cell_of(lvl), dr, r_min)

    # -- a$D_set_base -----------------------------------------------------
    def set_base(self, ix_list, nb_list):
        """ix_list (n, dim) and nb_list (n, 2*dim), 1-based box numbers,
        as a$D_set_base (m_aX_core.f90:152-235)."""
        ix_list = np.atleast_2d(np.asarray(ix_list, dtype=np.int64))
        nb_list = np.atleast_2d(np.asarray(nb_list, dtype=np.int64)).copy()
        if (ix_list < 1).any():
            raise ValueError("a_set_base: need all ix_list > 0")
        if self.highest_id > 0:
            raise ValueError("a_set_base: this tree already has boxes")
        n = ix_list.shape[0]
        rev = np.array([f ^ 1 for f in range(2 * self.dim)])
        for i in range(n):
            for f in range(2 * self.dim):
                nb = nb_list[i, f]
                if nb > 0 and nb != i + 1:
                    nb_list[nb - 1, rev[f]] = i + 1
        if (nb_list == 0).any():
            raise ValueError("a_set_base: unresolved neighbors")
        prev = None
        for lvl in range(self.min_lvl, 2):
            ids = self._new_ids(n)
            off = ids[0] - 1
            self.lvl[ids] = lvl
            self.ix[ids] = ix_list
            self.parent[ids] = 0
            self.children[ids] = 0
            self.nbr[ids] = np.where(nb_list > 0, nb_list + off, nb_list)
            self.in_use[ids] = True
            if prev is not None:
                self.children[prev, 0] = ids
                self.parent[ids] = prev
            prev = ids
        self.highest_lvl = 1
        self.ready = True

    # -- a$D_adjust_refinement -------------------------------------------
    def adjust_refinement(self, ref_fn):
        """One refinement pass; returns the number of added boxes (n_add)."""
        if not self.ready:
            raise ValueError("Tree not ready")
        dim, nch, nnb = self.dim, 1 << self.dim, 2 * self.dim
        hid_prev = self.highest_id
        flags = np.zeros(hid_prev + 1, dtype=np.int64)

        # consistent_ref_flags (m_aX_core.f90:628-725)
        for lvl in range(1, self.highest_lvl + 1):
            ids = self.ids_at(lvl)
            leaves = ids[~self.has_children(ids)]
            if leaves.size:
                flags[leaves] = ref_fn(self.box_view(leaves), KEEP)
                p = self.parent[leaves]
                first = (p > 0) & (ix_to_ichild(self.ix[leaves]) == 0)
                if first.any():
                    pp = p[first]
                    flags[pp] = ref_fn(self.box_view(pp), flags[pp])
        if flags.max(initial=0) > REFINE or flags.min(initial=0) < DEREFINE:
            raise ValueError("adjust_refinement: invalid refinement flag given")
        at_limit = self.ids_at(self.lvl_limit)
        flags[at_limit[flags[at_limit] == REFINE]] = KEEP

        for lvl in range(self.highest_lvl, 0, -1):
            ids = self.ids_at(lvl)
            leaves = ids[~self.has_children(ids)]
            ref = leaves[flags[leaves] > KEEP]
            flags[ref] = _A5_REFINE
            for f in range(nnb):
                miss = ref[self.nbr[ref, f] == 0]
                if miss.size:
                    flags[self.nbr[self.parent[miss], f]] = _A5_REFINE
            rm = leaves[flags[leaves] == DEREFINE]
            if rm.size:
                keep = np.zeros(rm.size, dtype=bool)
                for f in range(nnb):
                    nb = self.nbr[rm, f]
                    ok = nb > 0
                    nbc = np.zeros(rm.size, dtype=bool)
                    nbc[ok] = self.has_children(nb[ok]) | (flags[nb[ok]] > KEEP)
                    keep |= nbc
                flags[rm[keep]] = KEEP

        for lvl in range(self.highest_lvl - 1, 0, -1):
            ids = self.ids_at(lvl)
            par = ids[self.has_children(ids)]
            if par.size:
                ch = self.children[par]
                allrm = (flags[ch] == DEREFINE).all(axis=1) & (flags[par] <= KEEP)
                flags[par] = np.where(allrm, _A5_DEREFINE, KEEP)

        n_add = nch * int((flags == _A5_REFINE).sum())
        child_dix = _child_dix(dim)
        rev = np.array([f ^ 1 for f in range(nnb)])
        adj = [np.nonzero(child_dix[:, f // 2] == (f % 2))[0] for f in range(nnb)]

        lvl = 1
        while lvl <= self.lvl_limit - 1:
            ids = self.ids_at(lvl)
            old = ids[ids <= hid_prev]
            for pid in old[flags[old] == _A5_DEREFINE]:
                self._remove_children(pid)
            par = old[flags[old] == _A5_REFINE]
            if par.size:
                cids = self._new_ids(nch * par.size).reshape(par.size, nch)
                self.children[par] = cids
                base = 2 * self.ix[par] - 1
                for c in range(nch):
                    cc = cids[:, c]
                    self.ix[cc] = base + child_dix[c]
                    self.lvl[cc] = lvl + 1
                    self.parent[cc] = par
                    self.children[cc] = 0
                    self.nbr[cc] = 0
                    self.in_use[cc] = True
                for f in range(nnb):
                    phys = self.nbr[par, f] < 0
                    if phys.any():
                        for c in adj[f]:
                            self.nbr[cids[phys, c], f] = self.nbr[par[phys], f]
            # set_neighbs for the new boxes of lvl + 1 (find_neighb, :413-440)
            new = self.ids_at(lvl + 1)
            new = new[new > hid_prev]
            for f in range(nnb):
                d, low = f // 2, f % 2 == 0
                cand = new[self.nbr[new, f] == 0]
                if cand.size == 0:
                    continue
                cix = ix_to_ichild(self.ix[cand])
                child_low = ((cix >> d) & 1) == 0
                p = self.parent[cand].copy()
                same = child_low == low
                p[same] = self.nbr[p[same], f]
                okp = p > 0
                nb_id = np.zeros(cand.size, dtype=np.int64)
                nb_id[okp] = self.children[p[okp], cix[okp] ^ (1 << d)]
                hit = nb_id > 0
                self.nbr[cand[hit], f] = nb_id[hit]
                self.nbr[nb_id[hit], rev[f]] = cand[hit]
            if self.ids_at(lvl + 1).size == 0:
                break
            lvl += 1
        self.highest_lvl = lvl
        return n_add

    def _remove_children(self, pid):
        nnb = 2 * self.dim
        for c in self.children[pid]:
            for f in range(nnb):
                nb = self.nbr[c, f]
                if nb > 0:
                    self.nbr[nb, f ^ 1] = 0
            self.in_use[c] = False
        self.children[pid] = 0

    # -- export -----------------------------------------------------------
    def records(self):
        """int32 records [id, lvl, ix, parent, children, neighbors] in level order."""
        ids = np.concatenate([self.ids_at(l) for l in range(self.min_lvl, self.highest_lvl + 1)])
        rec = np.concatenate([ids[:, None], self.lvl[ids][:, None], self.ix[ids],
                              self.parent[ids][:, None], self.children[ids], self.nbr[ids]], axis=1)
        assert rec.shape[1] == record_len(self.dim)
        return rec.astype(np.int32)

    def state(self, n_var=0):
        st = TreeState(self.dim, self.nc, n_var, self.min_lvl, self.highest_lvl,
                       self.dr_base, self.records(), [])
        return st

    def n_leaf_cells(self):
        n = 0
        for l in range(self.min_lvl, self.highest_lvl + 1):
            ids = self.ids_at(l)
            n += int((~self.has_children(ids)).sum()) * int(self.n_cell_of(l)) ** self.dim
        return n

    def n_boxes(self):
        return int(self.in_use[:self.highest_id + 1].sum())


def build(dim, n_cell, ref_fn, coarsen_to=2, domain_boxes=1, dr=None, max_iter=100):
    """Unit-domain tree from one base box (Dirichlet all round), refined with
    ref_fn until no box is added (examples/poisson_basic_2d.f90:41-68)."""
    t = Tree(dim, n_cell, dr if dr is not None else 1.0 / n_cell, coarsen_to=coarsen_to)
    t.set_base(np.ones((1, dim), dtype=np.int64), -np.ones((1, 2 * dim), dtype=np.int64))
    for _ in range(max_iter):
        if t.adjust_refinement(ref_fn) == 0:
            break
    return t
