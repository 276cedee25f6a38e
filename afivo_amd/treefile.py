"""Binary tree/state file: the exchange format for trees and their cell data.

Layout (little-endian stream, written by the Fortran harness and by
:func:`write_tree_file`):

    char[4] 'AFGT'
    int32   version (1), dim, nc, n_var, min_lvl, highest_lvl, n_boxes
    float64 dr_base
    n_boxes int32 records [id, lvl, ix(dim), parent, children(2^dim),
                           neighbors(2*dim)]         (level order)
    for iv = 1..n_var, for each box in record order:
        float64 cc(0:nc_l+1, ..., iv) in Fortran (x fastest) order

A record carries afivo's box topology (``box$D_t`` components ``lvl``, ``ix``,
``parent``, ``children``, ``neighbors``; reference ``src/m_aX_t.f90:153-177``)
with afivo's conventions: 1-based ids, a neighbour id > 0 is a box, 0
(``a5_no_box``) a refinement boundary and < 0 a physical boundary.  ``nc_l``
is ``nc`` on levels >= 1 and ``nc / 2**(1-lvl)`` below (``a$D_n_cell``,
``src/m_aX_utils.f90:676-686``).
"""
from __future__ import annotations

import dataclasses

import numpy as np

MAGIC = b"AFGT"


def record_len(dim: int) -> int:
    return 3 + dim + (1 << dim) + 2 * dim


def level_ncell(nc: int, lvl: np.ndarray | int):
    lvl = np.asarray(lvl)
    return np.where(lvl >= 1, nc, nc // (2 ** np.clip(1 - lvl, 0, 30)))


@dataclasses.dataclass
class TreeState:
    dim: int
    nc: int
    n_var: int
    min_lvl: int
    highest_lvl: int
    dr_base: float
    recs: np.ndarray              # (n_boxes, record_len) int32
    data: list                    # n_var packed float64 arrays (may be empty)

    @property
    def n_boxes(self) -> int:
        return self.recs.shape[0]

    @property
    def ids(self):
        return self.recs[:, 0]

    @property
    def lvl(self):
        return self.recs[:, 1]

    @property
    def ix(self):
        return self.recs[:, 2:2 + self.dim]

    @property
    def parent(self):
        return self.recs[:, 2 + self.dim]

    @property
    def children(self):
        c0 = 3 + self.dim
        return self.recs[:, c0:c0 + (1 << self.dim)]

    @property
    def neighbors(self):
        n0 = 3 + self.dim + (1 << self.dim)
        return self.recs[:, n0:n0 + 2 * self.dim]

    def ncell(self):
        return level_ncell(self.nc, self.lvl)

    def box_sizes(self):
        return (self.ncell() + 2) ** self.dim

    def offsets(self):
        s = self.box_sizes()
        off = np.zeros(self.n_boxes + 1, dtype=np.int64)
        np.cumsum(s, out=off[1:])
        return off

    def face_sizes(self):
        """Values of one face variable per box: fx, fy (, fz) with
        (nc_l + 1) * nc_l^(dim-1) values each (box%fx.., m_aX_core.f90:359-369)."""
        n = self.ncell()
        return self.dim * (n + 1) * n ** (self.dim - 1)

    def face_offsets(self):
        s = self.face_sizes()
        off = np.zeros(self.n_boxes + 1, dtype=np.int64)
        np.cumsum(s, out=off[1:])
        return off

    def n_values(self) -> int:
        return int(self.box_sizes().sum())

    def box_array(self, iv: int, r: int) -> np.ndarray:
        """cc(..., iv) of the box in record r, as a C-ordered (z, y, x) view."""
        off = self.offsets()
        m = int(self.ncell()[r]) + 2
        a = self.data[iv - 1][off[r]:off[r + 1]]
        return a.reshape((m,) * self.dim)

    def interior_mask(self) -> np.ndarray:
        """Boolean mask over the packed values, true on interior cells."""
        mask = np.zeros(self.n_values(), dtype=bool)
        off = self.offsets()
        for r, n in enumerate(self.ncell()):
            m = int(n) + 2
            b = np.zeros((m,) * self.dim, dtype=bool)
            b[(slice(1, m - 1),) * self.dim] = True
            mask[off[r]:off[r + 1]] = b.ravel()
        return mask

    def leaf_mask(self) -> np.ndarray:
        """Boolean mask over the packed values, true on leaf interior cells."""
        mask = self.interior_mask()
        off = self.offsets()
        for r in np.nonzero(self.children[:, 0] != 0)[0]:
            mask[off[r]:off[r + 1]] = False
        return mask


def read_tree_file(path) -> TreeState:
    raw = open(path, "rb").read()
    if raw[:4] != MAGIC:
        raise ValueError(f"{path}: not a tree file")
    hdr = np.frombuffer(raw, dtype="<i4", count=7, offset=4)
    version, dim, nc, n_var, min_lvl, highest_lvl, n_boxes = (int(x) for x in hdr)
    if version != 1:
        raise ValueError(f"{path}: unsupported version {version}")
    pos = 4 + 7 * 4
    dr_base = float(np.frombuffer(raw, dtype="<f8", count=1, offset=pos)[0])
    pos += 8
    rl = record_len(dim)
    recs = np.frombuffer(raw, dtype="<i4", count=n_boxes * rl, offset=pos)
    recs = recs.reshape(n_boxes, rl).copy()
    pos += n_boxes * rl * 4
    st = TreeState(dim, nc, n_var, min_lvl, highest_lvl, dr_base, recs, [])
    nv = st.n_values()
    for _ in range(n_var):
        if pos >= len(raw):
            break
        st.data.append(np.frombuffer(raw, dtype="<f8", count=nv, offset=pos).copy())
        pos += nv * 8
    return st


def write_tree_file(path, st: TreeState) -> None:
    with open(path, "wb") as f:
        f.write(MAGIC)
        np.array([1, st.dim, st.nc, len(st.data) or st.n_var, st.min_lvl,
                  st.highest_lvl, st.n_boxes], dtype="<i4").tofile(f)
        np.array([st.dr_base], dtype="<f8").tofile(f)
        st.recs.astype("<i4").tofile(f)
        for a in st.data:
            np.asarray(a, dtype="<f8").tofile(f)
