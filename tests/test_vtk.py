"""VTK output of the leaf cells (afivo_amd/vtk.py): the file reads back with
one cell per leaf cell, the right corners, and the variables' leaf values."""
import numpy as np

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
