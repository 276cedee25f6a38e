"""VTK output of the leaf cells (afivo_amd/vtk.py): the XML .vtu writer
against the reference's own writer (fixtures), and the legacy .vtk file read
back (one cell per leaf cell, the right corners, the leaves' values)."""
import numpy as np
import pytest

from afivo_amd import problems as P
from afivo_amd import vtk


def check(st, tmp_path):
    path = tmp_path / "t.vtk"
    vtk.write_vtk(path, st, names=["phi", "rhs", "tmp"][:len(st.data)])
    pts, conn, types, data = vtk.read_vtk(path)
    leaf = st.leaf_mask()
    n = int(leaf.sum())
    npc = 1 << st.dim
    assert conn.shape == (n, npc + 1) and np.all(conn[:, 0] == npc)
    assert np.all(types == (8 if st.dim == 2 else 11))
    # every variable: the leaf values, as a multiset and in corner order
    _, _, index = vtk.leaf_cells(st)
    assert np.array_equal(np.sort(index), np.nonzero(leaf)[0])
    for iv, name in enumerate(["phi", "rhs", "tmp"][:len(st.data)], start=1):
        assert np.array_equal(data[name], st.data[iv - 1][index])
    # cells tile the unit domain: total measure 1, all corners inside it
    c = pts.reshape(n, npc, 3)
    edge = c[:, -1, :st.dim] - c[:, 0, :st.dim]
    assert np.allclose(edge, edge[:, :1])
    assert abs(np.prod(edge, axis=1).sum() - 1.0) < 1e-12
    assert pts[:, :st.dim].min() >= 0 and pts[:, :st.dim].max() <= 1 + 1e-12


def test_vtk_2d(tmp_path):
    check(P.small_2d().state, tmp_path)


def test_vtk_3d(tmp_path):
    check(P.small_3d().state, tmp_path)


VTKS = ["vtk_g2d", "vtk_g2d_cyl", "vtk_g3d"]


@pytest.mark.parametrize("name", VTKS)
def test_vtu_matches_reference_writer(name, tmp_path):
    """write_vtu (a$D_write_vtk, src/m_aX_io.f90:115-310) writes the same
    bytes as the reference's XML writer m_vtk did for the same tree, cell
    data and face variable (tests/golden/make_golden.py, VTKS): 2D with the
    domain at r0 = 0.25, 2D cylindrical (names _r/_z), 3D at r0 = -0.5;
    three cell variables and one face variable averaged to the cells."""
    import hashlib
    import json
    import os

    from afivo_amd.treefile import TreeState
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    meta = json.loads(str(z["meta"]))
    st = TreeState(meta["dim"], meta["nc"], 3, meta["min_lvl"], meta["highest_lvl"], meta["dr_base"],
                   z["recs"], list(z["data"]))
    path = vtk.write_vtu(st, tmp_path / name, n_cycle=meta["n_cycle"], time=meta["time"],
                         ixs_fc=meta["ixs_fc"], cc_names=meta["cc_names"], fc_names=meta["fc_names"],
                         fc=[z["fc"]], r_base=meta["r0"], cyl=meta["cyl"])
    got = open(path, "rb").read()
    assert got[:got.index(b"<AppendedData")].decode() == meta["vtu_head"]
    assert len(got) == meta["vtu_len"]
    assert hashlib.sha256(got).hexdigest() == meta["vtu_sha256"]


# flang's list-directed output of these reals (what m_vtk writes for TIME)
LIST_DIRECTED = [
    (0.0, "0."), (0.125, ".125"), (1e-5, "1.E-05"), (123.456, "123.456"), (1e10, "10000000000."),
    (-2.5, "-2.5"), (1 / 3, ".3333333333333333"), (1e-300, "1.E-300"), (0.01, "1.E-02"), (0.099, ".099"),
    (1e15, "1.E+15"), (1.5e20, "1.5E+20"), (2.0 ** 60, "1.152921504606847E+18"), (0.09, "9.E-02"),
    (0.095, ".095"), (0.0949, "9.49E-02"), (0.0951, ".0951"), (9.4e14, "940000000000000."),
    (9.6e14, "9.6E+14"), (0.09499999, "9.499999E-02"), (0.095000001, ".095000001"), (-0.0951, "-.0951"),
    (12345678901234.5, "12345678901234.5"), (-0.0, "-0."), (999999999999999.0, "9.99999999999999E+14"),
    (1.5e14, "150000000000000."),
]


def test_time_is_written_as_the_reference_compiler_does():
    for x, want in LIST_DIRECTED:
        assert vtk.fortran_list_real(x) == want, x


def read_vtu(path):
    """Points, connectivity, offsets, types and cell arrays of a .vtu file
    with a raw appended block (the layout m_vtk writes)."""
    import re
    raw = open(path, "rb").read()
    head = raw[:raw.index(b"<AppendedData")].decode()
    blob = raw[raw.index(b"_", raw.index(b'encoding="raw">')) + 1:]
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)"[^>]*Name="(\w+)"[^>]*offset="(\d+)"', head):
        typ, name, off = m.group(1), m.group(2), int(m.group(3))
        n = int(np.frombuffer(blob, "<i4", 1, off)[0])
        dt = {"Float32": "<f4", "Float64": "<f8", "Int32": "<i4"}[typ]
        arrays[name] = np.frombuffer(blob, dt, n // np.dtype(dt).itemsize, off + 4)
    return arrays


@pytest.mark.parametrize("dim", [2, 3])
def test_vtu_cells_are_the_leaf_cells(dim, tmp_path):
    """Independent of the reference: each cell's corners (through its
    connectivity) form a VTK pixel / voxel of the leaf's cell size in VTK's
    corner order, the cells tile the unit domain, and the cell data are the
    leaves' interior values, leaves of levels >= 1 in record order."""
    st = (P.small_2d() if dim == 2 else P.small_3d()).state
    path = vtk.write_vtu(st, tmp_path / "t", cc_names=["phi", "rhs", "tmp"])
    a = read_vtu(path)
    nch = 1 << dim
    pts = a["Points"].reshape(-1, 3)[:, :dim].astype(np.float64)
    conn = a["connectivity"].reshape(-1, nch)
    assert np.array_equal(a["offsets"], nch * np.arange(1, len(conn) + 1))
    assert np.all(a["types"] == (8 if dim == 2 else 11))
    c = pts[conn]                                            # (cells, corners, dim)
    edge = c[:, 1, 0] - c[:, 0, 0]
    bits = np.array([[(k >> d) & 1 for d in range(dim)] for k in range(nch)], dtype=float)
    assert np.allclose(c, c[:, :1, :] + bits[None] * edge[:, None, None], atol=1e-6)
    assert abs((edge ** dim).sum() - 1.0) < 1e-5
    leaf = st.leaf_mask()
    off = st.offsets()
    lv = st.lvl
    order = [r for level in range(1, st.highest_lvl + 1) for r in np.nonzero(lv == level)[0]
             if st.children[r, 0] == 0]
    idx = np.concatenate([np.arange(off[r], off[r + 1])[leaf[off[r]:off[r + 1]]] for r in order])
    for iv, name in enumerate(["phi", "rhs", "tmp"]):
        assert np.array_equal(a[name], st.data[iv][idx])
