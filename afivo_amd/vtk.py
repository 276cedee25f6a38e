"""VTK output of a tree's leaf cells (ParaView, VisIt).

* :func:`write_vtu` writes the reference's own format: ``a$D_write_vtk``
  (``src/m_aX_io.f90:115-310``) through its XML writer ``m_vtk``
  (``src/m_vtk.f90``) -- an ``UnstructuredGrid`` ``.vtu`` file with the time
  and cycle as field data, one pixel (2D) / voxel (3D) per leaf cell of the
  levels >= 1 with its own corner nodes (Float32), the cell variables
  (Float64) and the face variables averaged to the cell centres, all in one
  raw appended block.  The bytes equal the reference writer's for the same
  tree and data (``tests/test_vtk.py`` against fixtures written by the
  reference's ``m_vtk``, ``tests/golden/make_golden.py``).
* :func:`write_vtk` writes a legacy binary ``.vtk`` file of the same cells,
  for tools that do not read XML.

Geometry follows afivo: a box at level ``lvl >= 1`` with index ``ix`` has
``dr = dr_base / 2**(lvl-1)``; ``r_min`` is ``r_base + (ix-1) dr_base nc`` on
level 1 and the parent's ``r_min + 0.5 dr_p dix nc`` below it
(``src/m_aX_core.f90:194-198, 773-775``); leaves are boxes without children.

Usage: ``python -m afivo_amd.vtk tree_file out.vtk [name ...]`` (legacy), or
``python -m afivo_amd.vtk --vtu tree_file out [name ...]`` (writes out.vtu).
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


def box_r_min(st: TreeState, r_base=0.0):
    """r_min of every box (records order, levels >= 1; lower levels get the
    base box's origin), as a$D_set_base and add_children compute it."""
    dim, nc = st.dim, st.nc
    r_base = np.broadcast_to(np.asarray(r_base, dtype=np.float64), (dim,))
    rec = {int(i): r for r, i in enumerate(st.ids)}
    lvl = st.lvl
    dr = st.dr_base * 0.5 ** (lvl.astype(np.float64) - 1)
    out = np.tile(r_base, (st.n_boxes, 1))
    for r in np.argsort(lvl, kind="stable"):
        if lvl[r] == 1:
            out[r] = r_base + (st.ix[r] - 1) * st.dr_base * nc
        elif lvl[r] > 1:
            q = rec[int(st.parent[r])]
            dix = st.ix[r] - (2 * st.ix[q] - 1)
            out[r] = out[q] + 0.5 * dr[q] * dix * nc
    return out, dr


def fortran_list_real(x: float) -> str:
    """A real as flang's list-directed output writes it (what m_vtk writes
    for TIME): the shortest round-trip digits, in fixed form without a
    leading zero when the value rounded to one significant digit lies in
    [0.1, 1e15), else as d.dddE+XX; zero as "0."."""
    from decimal import Decimal

    x = float(x)
    sign = "-" if np.signbit(x) else ""
    if x == 0:
        return sign + "0."
    digits, exp = _shortest_digits(abs(x))      # abs(x) = 0.digits * 10**exp
    if Decimal("0.1") <= _round1(abs(x)) < Decimal("1e15"):
        if exp <= 0:
            return sign + "." + "0" * (-exp) + digits
        if len(digits) <= exp:
            return sign + digits + "0" * (exp - len(digits)) + "."
        return sign + digits[:exp] + "." + digits[exp:]
    e = exp - 1
    return f"{sign}{digits[0]}.{digits[1:]}E{'-' if e < 0 else '+'}{abs(e):02d}"


def _shortest_digits(a: float):
    """Shortest round-trip decimal digits of a > 0 and the exponent e with
    a = 0.digits * 10**e."""
    m, _, e = f"{a!r}".partition("e")
    e = int(e) if e else 0
    ip, _, fp = m.partition(".")
    digits = (ip + fp).lstrip("0")
    lead = len(ip.lstrip("0")) if ip.strip("0") else -(len(fp) - len(fp.lstrip("0")))
    digits = digits.rstrip("0") or "0"
    return digits, lead + e


def _round1(a: float):
    """The exact binary value of a > 0 rounded half-up to one significant digit."""
    from decimal import ROUND_HALF_UP, Decimal

    d = Decimal(a)
    return d.quantize(Decimal(1).scaleb(d.adjusted()), rounding=ROUND_HALF_UP)


def write_vtu(st: TreeState, filename, n_cycle=0, time=0.0, ixs_cc=None, ixs_fc=None, dir=None, *,
              cc_names=None, fc_names=None, fc=None, r_base=0.0, cyl=False):
    """a$D_write_vtk (src/m_aX_io.f90:115-310) of the tree `st`: writes
    filename + ".vtu" (into dir if given) and returns its path.

    ixs_cc: the cell variables (1-based; default all), named by cc_names
    (default cc_var1, cc_var2, ... as a$D_init names them); ixs_fc: face
    variables of `fc` (fc[ifc - 1] packed per box in record order as
    fx, fy[, fz], each (nc+1) nc^(dim-1) values in Fortran order; default
    none), named fc_names (default flux1, ...) with suffixes _x, _y, _z (_r,
    _z on a cylindrical 2D tree).  As in the reference, the value written for
    each face variable of a cell is 0.5 * the sum over the cell's two faces
    of all the listed face variables (the reference sums the whole section
    fx(i:i+1, j, ixs_fc)), which is the face average when one is listed."""
    import os

    dim, nc = st.dim, st.nc
    icc = list(range(1, len(st.data) + 1)) if ixs_cc is None else [int(i) for i in ixs_cc]
    ifc = [] if ixs_fc is None else [int(i) for i in ixs_fc]
    cc_names = list(cc_names) if cc_names else [f"cc_var{iv}" for iv in range(1, len(st.data) + 1)]
    fc_names = list(fc_names) if fc_names else [f"flux{i}" for i in range(1, (len(fc) if fc else 0) + 1)]
    sfx = (["_r", "_z"] if cyl else ["_x", "_y"]) if dim == 2 else ["_x", "_y", "_z"]
    names = [cc_names[i - 1] for i in icc] + [fc_names[i - 1] + s for i in ifc for s in sfx]

    r_min, dr = box_r_min(st, r_base)
    off = st.offsets()
    foff = st.face_offsets() if ifc else None
    bc, bn = nc, nc + 1
    npb, cpb, nch = bn ** dim, bc ** dim, 1 << dim
    leaves = [r for lv in range(1, st.highest_lvl + 1) for r in np.nonzero(st.lvl == lv)[0]
              if st.children[r, 0] == 0]
    ng = len(leaves)
    n_nodes, n_cells = npb * ng, cpb * ng
    coords = np.zeros((ng, npb, 3), dtype=np.float32)
    cc_vars = np.empty((len(names), ng, cpb), dtype=np.float64)
    g = np.arange(1, bc + 1)
    gn = np.arange(bn)
    if dim == 2:
        jj, ii = np.meshgrid(gn, gn, indexing="ij")
        node_pos = np.stack([ii.ravel(), jj.ravel()], axis=1)
        cj, ci = np.meshgrid(g, g, indexing="ij")
        q = (ci + (bc + 2) * cj).ravel()
        base = ((cj - 1) * bn + ci - 1).ravel()
        corner = np.array([0, 1, bn, bn + 1])
    else:
        kk, jj, ii = np.meshgrid(gn, gn, gn, indexing="ij")
        node_pos = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
        ck, cj, ci = np.meshgrid(g, g, g, indexing="ij")
        q = (ci + (bc + 2) * (cj + (bc + 2) * ck)).ravel()
        base = ((ck - 1) * bn * bn + (cj - 1) * bn + ci - 1).ravel()
        b2 = bn * bn
        corner = np.array([0, 1, bn, bn + 1, b2, b2 + 1, b2 + bn, b2 + bn + 1])
    for n, r in enumerate(leaves):
        coords[n, :, :dim] = (r_min[r][None, :] + node_pos * dr[r]).astype(np.float32)
        for k, iv in enumerate(icc):
            cc_vars[k, n] = st.data[iv - 1][off[r] + q]
        if ifc:
            faces = _face_sums(dim, nc, [fc[i - 1][foff[r]:foff[r + 1]] for i in ifc])
            for m in range(dim):
                for k in range(len(ifc)):
                    cc_vars[len(icc) + k * dim + m, n] = faces[m]
    connects = (np.arange(ng)[:, None, None] * npb + base[None, :, None] + corner[None, None, :])
    connects = connects.astype("<i4").ravel()
    offsets = (nch * np.arange(1, n_cells + 1)).astype("<i4")
    types = np.full(n_cells, 8 if dim == 2 else 11, dtype="<i4")

    blocks, lines, offset = [], [], 0

    def appended(arr, n_bytes):
        nonlocal offset
        blocks.append(np.array([n_bytes], dtype="<i4").tobytes() + arr.tobytes())
        at = offset
        offset += 4 + n_bytes
        return at

    ind = 0

    def put(text):
        lines.append(" " * ind + text + "\n")

    def open_tag(name):
        nonlocal ind
        put(f"<{name}>")
        ind += 2

    def close_tag(name):
        nonlocal ind
        ind -= 2
        put(f"</{name}>")

    lines.append('<?xml version="1.0"?>\n')
    lines.append('<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">\n')
    ind = 2
    open_tag("UnstructuredGrid")
    open_tag("FieldData")
    put('<DataArray type="Float64" Name="TIME" NumberOfTuples="1" format="ascii">')
    ind += 2
    lines.append((" " + " " * ind + " " + fortran_list_real(time)).rstrip() + "\n")
    close_tag("DataArray")
    put('<DataArray type="Int32" Name="CYCLE" NumberOfTuples="1" format="ascii">')
    ind += 2
    lines.append((" " + " " * ind + " " + str(int(n_cycle))).rstrip() + "\n")
    close_tag("DataArray")
    close_tag("FieldData")
    put(f'<Piece NumberOfPoints="{n_nodes}" NumberOfCells="{n_cells}">')
    ind += 2
    open_tag("Points")
    at = appended(coords.astype("<f4"), 3 * n_nodes * 4)
    put(f'<DataArray type="Float32" NumberOfComponents="3" Name="Points" format="appended" offset="{at}"/>')
    close_tag("Points")
    open_tag("Cells")
    at = appended(connects, int(offsets[-1]) * 4 if n_cells else 0)
    put(f'<DataArray type="Int32" Name="connectivity" format="appended" offset="{at}"/>')
    at = appended(offsets, n_cells * 4)
    put(f'<DataArray type="Int32" Name="offsets" format="appended" offset="{at}"/>')
    at = appended(types, n_cells * 4)
    put(f'<DataArray type="Int32" Name="types" format="appended" offset="{at}"/>')
    close_tag("Cells")
    open_tag("CellData")
    for k, name in enumerate(names):
        at = appended(cc_vars[k].astype("<f8"), n_cells * 8)
        put(f'<DataArray type="Float64" Name="{name}" NumberOfComponents="1" format="appended" offset="{at}"/>')
    close_tag("CellData")
    close_tag("Piece")
    close_tag("UnstructuredGrid")
    put('<AppendedData encoding="raw">')
    path = str(filename) + ".vtu"
    if dir:
        path = os.path.join(dir, path)
    with open(path, "wb") as f:
        f.write("".join(lines).encode())
        f.write(b"_" + b"".join(blocks) + b"\n")
        f.write((" " * ind + "</AppendedData>\n").encode())
        f.write(b"</VTKFile>\n")
    return path


def _face_sums(dim, nc, boxes):
    """Per direction, 0.5 * (sum over the two faces of each cell and over the
    listed face variables, in the reference's column order), one value per
    interior cell (x fastest)."""
    n = nc
    out = []
    for m in range(dim):
        tot = None
        for arr in boxes:
            size = (n + 1) * n ** (dim - 1)
            a = arr[m * size:(m + 1) * size]
            shape = [n] * dim
            shape[m] = n + 1
            f = a.reshape(shape[::-1])            # Fortran order: last index slowest
            lo = [slice(0, n)] * dim
            hi = [slice(0, n)] * dim
            lo[dim - 1 - m] = slice(0, n)
            hi[dim - 1 - m] = slice(1, n + 1)
            first, second = f[tuple(lo)].ravel(), f[tuple(hi)].ravel()
            tot = first + second if tot is None else (tot + first) + second
        out.append(0.5 * tot)
    return out


def main(argv):
    if len(argv) >= 2 and argv[1] == "--vtu":
        if len(argv) < 4:
            print(__doc__)
            return 2
        st = read_tree_file(argv[2])
        print(write_vtu(st, argv[3], cc_names=argv[4:] or None))
        return 0
    if len(argv) < 3:
        print(__doc__)
        return 2
    st = read_tree_file(argv[1])
    write_vtk(argv[2], st, names=argv[3:] or None)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
