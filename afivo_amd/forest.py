"""Forests: several independent trees solved together in one context.

`combine` packs trees (same dim, n_cell, dr_base and variables) into one
TreeState with renumbered box ids plus the tree index of every record, the
input of afg_set_forest (include/afivo_amd.h).  `split` cuts a packed
variable of the forest back into the trees' own packed arrays.  The batched
configuration (`c3_trees`) draws its trees from counter-based hashes of
(seed, tree, level, index), so a tree does not depend on how many others are
built or in which order (the reference's random_refinement_2d draws from a
stateful generator inside its refinement callback, examples/
random_refinement_2d.f90:79-85).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import problems as P
from . import tree as treemod
from .treefile import TreeState


def combine(states):
    """One TreeState holding every tree of `states`, and tree_of per record."""
    s0 = states[0]
    for st in states:
        if (st.dim, st.nc, st.dr_base, len(st.data)) != (s0.dim, s0.nc, s0.dr_base, len(s0.data)):
            raise ValueError("trees of a forest need the same dim, n_cell, dr_base and variables")
    D = s0.dim
    c_par, c_ch, c_nb = 2 + D, 3 + D, 3 + D + (1 << D)
    recs, tree_of, off = [], [], 0
    for t, st in enumerate(states):
        r = st.recs.astype(np.int64).copy()
        r[:, 0] += off
        for cols in (slice(c_par, c_par + 1), slice(c_ch, c_nb), slice(c_nb, c_nb + 2 * D)):
            block = r[:, cols]
            block[block > 0] += off
            r[:, cols] = block
        recs.append(r)
        tree_of.append(np.full(st.n_boxes, t, dtype=np.int32))
        off += int(st.ids.max())
    if off >= 2 ** 31:
        raise ValueError("too many boxes for int32 ids")
    recs = np.concatenate(recs).astype(np.int32)
    data = [np.concatenate([st.data[iv] for st in states]) for iv in range(len(s0.data))]
    out = TreeState(D, s0.nc, s0.n_var, min(st.min_lvl for st in states),
                    max(st.highest_lvl for st in states), s0.dr_base, recs, data)
    return out, np.concatenate(tree_of)


def split(states, packed):
    """The packed forest variable `packed` (record order of `combine`) cut
    into one packed array per tree."""
    out, pos = [], 0
    for st in states:
        n = st.n_values()
        out.append(packed[pos:pos + n])
        pos += n
    return out


def _hash(*keys):
    """A counter-based 64-bit hash (splitmix64 steps over the keys), mapped
    to [0, 1)."""
    x = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        for k in keys:
            x = x ^ np.uint64(int(k) & 0xFFFFFFFFFFFFFFFF)
            x = x + np.uint64(0x9E3779B97F4A7C15)
            z = x
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            x = z ^ (z >> np.uint64(31))
    return float(x >> np.uint64(11)) / float(1 << 53)


@dataclasses.dataclass
class TreeSpec:
    nc: int
    l_max: int
    gs: P.Gaussians


def c3_spec(seed, t, dim=2, nc_choices=(8, 16), l_range=(4, 8)):
    """Tree t of the batched configuration: n_cell, maximum level and
    right-hand side Gaussians, from hashes of (seed, t)."""
    nc = nc_choices[int(_hash(seed, t, 1) * len(nc_choices))]
    l_max = l_range[0] + int(_hash(seed, t, 2) * (l_range[1] - l_range[0] + 1))
    n_g = 1 + int(_hash(seed, t, 3) * 3)
    amp = np.array([1.0 if _hash(seed, t, 10, g) < 0.5 else -1.0 for g in range(n_g)])
    sig = np.array([0.02 + 0.06 * _hash(seed, t, 11, g) for g in range(n_g)])
    r0 = np.array([[0.1 + 0.8 * _hash(seed, t, 12, g, d) for d in range(dim)] for g in range(n_g)])
    return TreeSpec(nc, l_max, P.Gaussians(amp, sig, r0))


def random_refinement(seed, t, l_uniform, l_max, p_ref):
    """Refine every box below l_uniform, and boxes below l_max where the hash
    of (seed, tree, level, index) is below p_ref."""

    def ref(view, cur):
        out = np.array(cur, dtype=np.int64, copy=True) if np.ndim(cur) else np.full(view.ids.size, cur)
        for i in range(view.ids.size):
            lvl = int(view.lvl[i])
            if lvl < l_uniform:
                out[i] = treemod.REFINE
            elif lvl < l_max and _hash(seed, t, 100 + lvl, *[int(v) for v in view.ix[i]]) < p_ref:
                out[i] = treemod.REFINE
        return out

    return ref


def c3_tree(seed, t, dim=2, l_uniform=3, p_ref=0.5, **kw):
    """Tree t of the batched configuration with phi = 0, rhs = gauss_lpl and
    tmp = 0 (and its spec)."""
    spec = c3_spec(seed, t, dim, **kw)
    tr = treemod.build(dim, spec.nc, random_refinement(seed, t, l_uniform, spec.l_max, p_ref))
    st = tr.state()
    rhs = P.set_rhs(tr, st, spec.gs)
    zeros = np.zeros_like(rhs)
    st.data = [zeros.copy(), rhs, zeros.copy()]
    st.n_var = 3
    return st, spec
