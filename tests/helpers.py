"""Shared helpers of the test suite."""
import numpy as np


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def assert_bitexact(got, want, what=""):
    """Bit-for-bit equality of two float64 arrays (sign of zero included)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    bad = np.nonzero(bits(got) != bits(want))[0]
    if bad.size:
        k = bad[0]
        raise AssertionError(
            f"{what}: {bad.size} of {got.size} values differ in their bits; first at {k}: "
            f"{got[k]!r} vs {want[k]!r} (max |diff| {np.nanmax(np.abs(got - want)):.3e})")


def random_state(st, n_var, seed):
    """Random values in every variable, ghosts and corners included."""
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(st.n_values()) for _ in range(n_var)]


# The routines of the transfer fixtures (tests/golden/xfer_*.npz, written by
# make_golden.py from the reference's own a$D_gc_tree, a$D_restrict_tree and
# a$D_prolong*_from): op -> (source variable, what it runs on variable 4).
# `t` is an OracleTree or a DeviceTree (same method names).
RB_MG, RB_PROLONG0, RB_INTERP, RB_INTERP_LIM = 0, 1, 2, 3
BC_DIRICHLET, BC_NEUMANN = -1, -2


def run_xfer(t, op, min_lvl, highest_lvl):
    if op == "gci":
        t.gc_tree(4, RB_INTERP, BC_DIRICHLET)
    elif op == "gcl":
        t.gc_tree(4, RB_INTERP_LIM, BC_NEUMANN)
    elif op == "gc0":
        t.gc_tree(4, RB_PROLONG0, BC_NEUMANN)
    elif op == "rst":
        t.restrict_tree(4)
    else:
        order = {"pr0": 0, "pr1": 1, "pr2": 2}[op]
        for lvl in range(min_lvl + 1, highest_lvl + 1):
            t.prolong_level(1, 4, lvl, order, add=(op == "pr1"))


def xfer_expected(source, pos, val):
    """Variable 4 after a fixture's routine: the source variable with the
    routine's changes."""
    out = np.array(source, dtype=np.float64, copy=True)
    out[pos] = val
    return out


def face_gradient(st, phi):
    """Face variable holding (phi(i) - phi(i-1)) / dr on every face of every
    box (fx, fy (, fz) per box, Fortran order), packed as the face layout of
    afg_upload_fc: the flux the transfer fixtures start from."""
    out = np.zeros(int(st.face_offsets()[-1]))
    off, foff = st.offsets(), st.face_offsets()
    for r in range(st.n_boxes):
        n = int(st.ncell()[r])
        m = n + 2
        dr = st.dr_base * 0.5 ** (int(st.lvl[r]) - 1)
        a = phi[off[r]:off[r + 1]].reshape((m,) * st.dim)    # C order: (z,) y, x
        parts = []
        for d in range(st.dim):
            ax = st.dim - 1 - d                               # the array axis of dimension d
            hi = [slice(1, n + 1)] * st.dim
            lo = [slice(1, n + 1)] * st.dim
            hi[ax] = slice(1, n + 2)
            lo[ax] = slice(0, n + 1)
            parts.append(((a[tuple(hi)] - a[tuple(lo)]) / dr).ravel())
        out[foff[r]:foff[r + 1]] = np.concatenate(parts)
    return out
