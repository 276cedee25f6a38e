"""VTK output of a tree's leaf cells, for visual checks (ParaView, VisIt).

Writes a legacy-format binary VTK file (``DATASET UNSTRUCTURED_GRID``): one
pixel (2D) or voxel (3D) per leaf cell, with its 2**dim corners, and one
``SCALARS`` block of cell data per variable.  The reference writes its trees
with ``a$D_write_vtk`` (``src/m_aX_io.f90:115-310``), which needs the Silo
library; this writer covers the same use (looking at phi, rhs and the residual
on the leaves) from the tree/state file format of ``afivo_amd.treefile``.

Geometry follows afivo: a box at level ``lvl >= 1`` with index ``ix`` has
``dr = dr_base / 2**(lvl-1)`` and lower corner ``origin + (ix-1) * nc * dr``
(``src/m_aX_core.f90:196-198,764-766``); leaves are boxes without children.

Usage: ``python -m afivo_amd.vtk tree_file out.vtk [name ...]``.
"""
from __future__ import annotations

import sys

import numpy as np

from .treefile import TreeState, read_tree_file

VTK_PIXEL, VTK_VOXEL = 8, 11


def leaf_cells(st: TreeState, origin=None):
    """Lower corners (n, dim), edge lengths (n,) and, per leaf box, the
    packed-value index of each of its cells, in the order of the corners."""
    dim = st.dim
    origin = np.zeros(dim) if origin is None else np.asarray(origin, dtype=float)
    off = st.offsets()
    nc_l = st.ncell()
    corners, sizes, index = [], [], []
    for r in np.nonzero(st.children[:, 0] == 0)[0]:
        lvl, n = int(st.lvl[r]), int(nc_l[r])
        dr = st.dr_base * 0.5 ** (lvl - 1)
        r_min = origin + (st.ix[r] - 1) * st.nc * dr if lvl >= 1 else origin
        m = n + 2
        g = np.arange(1, n + 1)
        if dim == 2:
            j, i = np.meshgrid(g, g, indexing="ij")
            pos = np.stack([i, j], axis=-1).reshape(-1, 2)
            q = (i + m * j).ravel()
        else:
            k, j, i = np.meshgrid(g, g, g, indexing="ij")
            pos = np.stack([i, j, k], axis=-1).reshape(-1, 3)
            q = (i + m * (j + m * k)).ravel()
        corners.append(r_min + (pos - 1) * dr)
        sizes.append(np.full(len(q), dr))
        index.append(off[r] + q)
    if not corners:
        return np.zeros((0, dim)), np.zeros(0), np.zeros(0, dtype=np.int64)
    return np.concatenate(corners), np.concatenate(sizes), np.concatenate(index)


def write_vtk(path, st: TreeState, names=None, origin=None, title="afivo_amd leaf cells"):
    """Write the leaf cells of `st` with its variables (named by `names`,
    default var1, var2, ...) to a legacy binary VTK file."""
    dim = st.dim
    corner, size, index = leaf_cells(st, origin)
    n = len(size)
    npc = 1 << dim
    # the 2**dim corners of each cell, x fastest (pixel / voxel order)
    bits = np.array([[(c >> d) & 1 for d in range(dim)] for c in range(npc)], dtype=float)
    pts = corner[:, None, :] + bits[None, :, :] * size[:, None, None]
    pts3 = np.zeros((n * npc, 3))
    pts3[:, :dim] = pts.reshape(-1, dim)
    conn = np.empty((n, npc + 1), dtype=">i4")
    conn[:, 0] = npc
    conn[:, 1:] = np.arange(n * npc, dtype=np.int64).reshape(n, npc)
    names = list(names) if names else [f"var{iv}" for iv in range(1, st.n_var + 1)]
    with open(path, "wb") as f:
        f.write(f"# vtk DataFile Version 3.0\n{title}\nBINARY\nDATASET UNSTRUCTURED_GRID\n".encode())
        f.write(f"POINTS {n * npc} double\n".encode())
        f.write(pts3.astype(">f8").tobytes())
        f.write(f"\nCELLS {n} {n * (npc + 1)}\n".encode())
        f.write(conn.tobytes())
        f.write(f"\nCELL_TYPES {n}\n".encode())
        f.write(np.full(n, VTK_PIXEL if dim == 2 else VTK_VOXEL, dtype=">i4").tobytes())
        f.write(f"\nCELL_DATA {n}\n".encode())
        for iv, name in enumerate(names[:len(st.data)], start=1):
            f.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n".encode())
            f.write(np.asarray(st.data[iv - 1])[index].astype(">f8").tobytes())
            f.write(b"\n")


def read_vtk(path):
    """Read back what write_vtk writes: (points (n, 3), cell types, cell
    data dict).  Only this writer's layout is understood."""
    raw = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        e = raw.index(b"\n", pos)
        s = raw[pos:e].decode()
        pos = e + 1
        return s

    for _ in range(4):
        line()
    n_pts = int(line().split()[1])
    pts = np.frombuffer(raw, ">f8", 3 * n_pts, pos).reshape(n_pts, 3)
    pos += 24 * n_pts
    line()
    w = line().split()
    n, size = int(w[1]), int(w[2])
    conn = np.frombuffer(raw, ">i4", size, pos)
    pos += 4 * size
    line()
    n_types = int(line().split()[1])
    types = np.frombuffer(raw, ">i4", n_types, pos)
    pos += 4 * n_types
    line()
    n_data = int(line().split()[1])
    data = {}
    while pos < len(raw):
        name = line().split()[1]
        line()
        data[name] = np.frombuffer(raw, ">f8", n_data, pos).copy()
        pos += 8 * n_data + 1
    return pts, conn.reshape(n, -1), types, data


def main(argv):
    if len(argv) < 3:
        print(__doc__)
        return 2
    st = read_tree_file(argv[1])
    write_vtk(argv[2], st, names=argv[3:] or None)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
