"""The oracle (C restatement, oracle/afivo_oracle.c) pinned three ways:

1. against the golden fixtures written by the reference solver itself
   (tests/golden, made by tests/golden/make_golden.py): bit for bit;
2. against the reference built from its sources (oracle/_ref), on the full
   2D configuration and on randomised states, when that build is present;
3. against what the mathematics fixes (exactness of the transfer operators
   on polynomials, the discrete Laplacian of a quadratic, the GSRB fixed
   point, the boundary formula), so that a dropped term, a wrong sign or a
   transposed index fails a test.
"""
import gzip
import json
import os

import numpy as np
import pytest

import oracle as O
from afivo_amd import problems as P
from afivo_amd import tree as T
from afivo_amd.treefile import read_tree_file
from helpers import assert_bitexact, run_xfer, xfer_expected

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(f[:-5] for f in os.listdir(GOLDEN) if f.endswith(".json"))


def load_golden(name):
    cfg = json.load(open(os.path.join(GOLDEN, name + ".json")))
    st = {}
    for ext in ("in", "out"):
        raw = gzip.open(os.path.join(GOLDEN, f"{name}.{ext}.gz")).read()
        # one file per process: tests of several workers load the same case
        path = os.path.join(GOLDEN, f"_tmp_{name}.{os.getpid()}.{ext}")
        try:
            with open(path, "wb") as f:
                f.write(raw)
            st[ext] = read_tree_file(path)
        finally:
            os.remove(path)
    res = np.loadtxt(os.path.join(GOLDEN, name + ".res"), ndmin=1)
    return cfg, st["in"], st["out"], res


XFERS = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.startswith("xfer_") and f.endswith(".npz"))


def load_xfer(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    return meta, {op: (z[op + "_pos"], z[op + "_val"]) for op in meta["ops"]}


def rstr_of(c):
    """The restriction of a harness configuration: rstr_lpllsf with the
    lpllsf set (lsf_mode 2), mg$D_auto_rstr when rstr_auto, else rstr_lpl."""
    return O.RSTR_LSF if c.get("lsf_mode", 0) == 2 else O.RSTR_AUTO if c.get("rstr_auto") else O.RSTR_LPL


def opts_of(cfg):
    c = cfg["cfg"]
    lsf = c.get("lsf_mode", 0)
    return O.default_opts(c.get("n_down", 2), c.get("n_up", 2), c["n_base"], eps=c.get("use_eps", False),
                          auto=c.get("cyl", False), lsf=lsf == 1, lsf_set=lsf == 2, rstr=rstr_of(c))


def run_solve(solver, cfg, res_fn):
    c = cfg["cfg"]
    opts = opts_of(cfg)
    res = []
    for it in range(c.get("n_fmg", 1)):
        solver.fas_fmg(opts, True, it > 0)
        res.append(res_fn())
    for _ in range(c.get("n_vcycle", 0)):
        solver.fas_vcycle(opts)
        res.append(res_fn())
    return np.array(res)


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_golden(name):
    cfg, st_in, st_out, res_ref = load_golden(name)
    ot = O.OracleTree(st_in)
    if cfg["cfg"].get("cyl"):
        ot.set_coord(True)
    if cfg["cfg"].get("rstr_auto"):   # the harness tags every box first (mg$D_set_box_tag)
        ot.set_tags(opts_of(cfg))
    res = run_solve(ot, cfg, lambda: ot.max_abs(3))
    for iv in range(1, st_out.n_var + 1):
        assert_bitexact(ot.get(iv), st_out.data[iv - 1], f"{name} var {iv}")
    assert np.array_equal(res, res_ref), (res, res_ref)


needs_ref2 = pytest.mark.skipif(not O.ref_available(2), reason="reference build oracle/_ref absent")
needs_ref3 = pytest.mark.skipif(not O.ref_available(3), reason="reference build oracle/_ref absent")


@needs_ref2
def test_oracle_matches_reference_c1(tmp_path):
    """The full 2D configuration: FMG + 10 V-cycles, 2+2 sweeps, 50 base sweeps."""
    p = P.c1()
    prefix, _ = O.run_ref(2, P.harness_cfg(p), workdir=str(tmp_path))
    st_in, st_out = read_tree_file(prefix + ".in"), read_tree_file(prefix + ".out")
    ot = O.OracleTree(st_in)
    res = run_solve(ot, {"cfg": P.harness_cfg(p)}, lambda: ot.max_abs(3))
    for iv in (1, 2, 3):
        assert_bitexact(ot.get(iv), st_out.data[iv - 1], f"c1 var {iv}")
    assert np.array_equal(res, O.read_res(prefix))


@needs_ref3
@pytest.mark.parametrize("eps", [False, True])
def test_oracle_matches_reference_3d_guess(tmp_path, eps):
    """3D, warm-started FMG (set_coarse_phi_rhs path), odd sweep counts."""
    p = P.small_3d(nc=4, l_uniform=2, l_max=4, eps_hi=3.0 if eps else None)
    cfg = P.harness_cfg(p, n_fmg=2, n_vcycle=2, n_down=3, n_up=1, n_base=5)
    prefix, _ = O.run_ref(3, cfg, workdir=str(tmp_path))
    st_in, st_out = read_tree_file(prefix + ".in"), read_tree_file(prefix + ".out")
    ot = O.OracleTree(st_in)
    res = run_solve(ot, {"cfg": cfg}, lambda: ot.max_abs(3))
    for iv in range(1, st_out.n_var + 1):
        assert_bitexact(ot.get(iv), st_out.data[iv - 1], f"var {iv}")
    assert np.array_equal(res, O.read_res(prefix))


# ---------------------------------------------------------------------------
# mathematical pins


def uniform_tree(dim, nc, lvl):
    t = T.build(dim, nc, lambda v, cur: np.where(v.lvl < lvl, T.REFINE, T.KEEP))
    return t, t.state()


def sample(t, st, f):
    """Packed array of f at the centres of every cell, ghosts included."""
    view = t.box_view(st.ids)
    out = np.zeros(st.n_values())
    off = st.offsets()
    for r, n in enumerate(st.ncell()):
        n = int(n)
        c = view.r_min[r][:, None] + (np.arange(0, n + 2) - 0.5)[None, :] * view.dr[r]
        grids = np.meshgrid(*[c[d] for d in range(t.dim)][::-1], indexing="ij")  # (z, y, x)
        out[off[r]:off[r + 1]] = f(*grids[::-1]).ravel()
    return out


def interior_of_level(st, lvl):
    mask = st.interior_mask()
    off = st.offsets()
    for r in range(st.n_boxes):
        if st.lvl[r] != lvl:
            mask[off[r]:off[r + 1]] = False
    return mask


@pytest.mark.parametrize("dim", [2, 3])
def test_restriction_exact_for_linear(dim):
    t, st = uniform_tree(dim, 4, 3)
    f = (lambda x, y: 1 + 2 * x - 3 * y) if dim == 2 else (lambda x, y, z: 1 + 2 * x - 3 * y + 0.5 * z)
    ot = O.OracleTree(st, n_var=1)
    ot.set(1, sample(t, st, f))
    O.lib().orc_restrict_level(ot.h, 1, 3)
    m = interior_of_level(st, 2)
    got, want = ot.get(1)[m], sample(t, st, f)[m]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_prolong2_exact_for_separable_quadratic(dim):
    t, st = uniform_tree(dim, 4, 3)
    if dim == 2:
        f = lambda x, y: 0.3 + x - 2 * y + 3 * x * x - 0.7 * y * y
    else:
        f = lambda x, y, z: 0.3 + x - 2 * y + 0.1 * z + 3 * x * x - 0.7 * y * y + 1.3 * z * z
    ot = O.OracleTree(st, n_var=2)
    ot.set(1, sample(t, st, f))
    O.lib().orc_prolong2_level(ot.h, 1, 2, 3, 0)
    m = interior_of_level(st, 3)
    np.testing.assert_allclose(ot.get(2)[m], sample(t, st, f)[m], rtol=0, atol=2e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_laplacian_of_quadratic(dim):
    """box_lpl of |x|^2 is 2*dim; residual = rhs - lpl = 0 with rhs = 2*dim."""
    t, st = uniform_tree(dim, 4, 2)
    f = (lambda x, y: x * x + y * y) if dim == 2 else (lambda x, y, z: x * x + y * y + z * z)
    ot = O.OracleTree(st, n_var=3)
    ot.set(1, sample(t, st, f))
    ot.set(2, np.full(st.n_values(), 2.0 * dim))
    O.lib().orc_residual_level(ot.h, O.default_opts(), 2)
    m = interior_of_level(st, 2)
    np.testing.assert_allclose(ot.get(3)[m], 0.0, atol=1e-9)


def test_dirichlet_ghost_is_minus_interior():
    t, st = uniform_tree(2, 4, 1)
    ot = O.OracleTree(st, n_var=1)
    rng = np.random.default_rng(3)
    a = rng.standard_normal(st.n_values())
    ot.set(1, a)
    O.lib().orc_fill_gc_level(ot.h, 1, 1)
    r = int(np.nonzero(st.lvl == 1)[0][0])
    box = ot.get(1)[st.offsets()[r]:st.offsets()[r + 1]].reshape(6, 6)
    src = a[st.offsets()[r]:st.offsets()[r + 1]].reshape(6, 6)
    np.testing.assert_array_equal(box[1:5, 0], -src[1:5, 1])    # low x
    np.testing.assert_array_equal(box[1:5, 5], -src[1:5, 4])    # high x
    np.testing.assert_array_equal(box[0, 1:5], -src[1, 1:5])    # low y
    np.testing.assert_array_equal(box[5, 1:5], -src[4, 1:5])    # high y
    np.testing.assert_array_equal(box[0, 0], src[0, 0])         # corners untouched


def test_neumann_ghost_formula():
    t, st = uniform_tree(2, 4, 1)
    ot = O.OracleTree(st, n_var=1)
    a = np.random.default_rng(4).standard_normal(st.n_values())
    ot.set(1, a)
    r = int(np.nonzero(st.lvl == 1)[0][0])
    id_ = int(st.ids[r])
    b = np.array([0.5, -1.0, 2.0, 0.25])
    ot.set_bc(id_, 2, -2, b)             # high x: ghost = dr * b + interior
    O.lib().orc_fill_gc_level(ot.h, 1, 1)
    box = ot.get(1)[st.offsets()[r]:st.offsets()[r + 1]].reshape(6, 6)
    src = a[st.offsets()[r]:st.offsets()[r + 1]].reshape(6, 6)
    np.testing.assert_array_equal(box[1:5, 5], 0.25 * b + src[1:5, 4])


def test_gsrb_converges_to_dense_solution():
    """Many GSRB sweeps on a one-box level with Dirichlet-0 ghosts reach the
    solution of the dense 5-point system."""
    t, st = uniform_tree(2, 4, 1)
    n = 4
    h = 1.0 / n
    rhs = np.random.default_rng(5).standard_normal((n, n))
    ot = O.OracleTree(st, n_var=3)
    r = int(np.nonzero(st.lvl == 1)[0][0])
    packed = np.zeros(st.n_values())
    box = np.zeros((n + 2, n + 2))
    box[1:-1, 1:-1] = rhs
    packed[st.offsets()[r]:st.offsets()[r + 1]] = box.ravel()
    ot.set(2, packed)
    O.lib().orc_gsrb_level(ot.h, O.default_opts(), 1, 400)
    phi = ot.get(1)[st.offsets()[r]:st.offsets()[r + 1]].reshape(n + 2, n + 2)[1:-1, 1:-1]
    # dense: (sum of 4 neighbours - 4 phi) / h^2 = rhs, ghosts = -interior
    N = n * n
    A = np.zeros((N, N))
    for j in range(n):
        for i in range(n):
            k = j * n + i
            A[k, k] = -4
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < n and 0 <= jj < n:
                    A[k, jj * n + ii] += 1
                else:
                    A[k, k] -= 1
    sol = np.linalg.solve(A / (h * h), rhs.ravel())
    np.testing.assert_allclose(phi.ravel(), sol, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_vcycle_residual_reduction(dim):
    """A 2+2 V-cycle reduces the residual by about an order of magnitude."""
    p = P.small_2d() if dim == 2 else P.small_3d()
    ot = O.OracleTree(p.state)
    opts = O.default_opts(2, 2, p.n_base)
    ot.fas_fmg(opts, True, False)
    r0 = ot.max_abs(3)
    ot.fas_vcycle(opts)
    r1 = ot.max_abs(3)
    ot.fas_vcycle(opts)
    r2 = ot.max_abs(3)
    assert r1 < 0.3 * r0 and r2 < 0.3 * r1, (r0, r1, r2)


@pytest.mark.parametrize("what", ["ulp", "sweeps"])
def test_planted_mistake_is_caught(what):
    """Planted mistakes -- one interior value off by a relative 1e-9 after
    FMG, or the wrong number of up-sweeps -- change the bits of the final
    state, so the bit-exact comparisons used throughout catch them."""
    cfg, st_in, st_out, _ = load_golden("g2d_nc4")
    ot = O.OracleTree(st_in)
    opts = opts_of(cfg)
    if what == "sweeps":
        opts.n_cycle_up += 1
    ot.fas_fmg(opts, True, False)
    if what == "ulp":
        a = ot.get(1)
        off = st_in.offsets()
        r = int(np.nonzero(st_in.lvl == st_in.highest_lvl)[0][0])
        q = off[r] + (int(st_in.ncell()[r]) + 2) + 1          # cell (1, 1)
        a[q] = a[q] * (1 + 1e-9)
        ot.set(1, a)
    for _ in range(cfg["cfg"]["n_vcycle"]):
        ot.fas_vcycle(opts)
    assert not np.array_equal(ot.get(1), st_out.data[0])


@pytest.mark.parametrize("name", XFERS)
def test_oracle_transfers_match_reference(name):
    """a$D_gc_tree (gc_interp, gc_interp_lim, gc_prolong0 with Dirichlet or
    Neumann zero), a$D_restrict_tree and a$D_prolong0/1/2_from: the oracle
    reproduces the reference's variable 4 bit for bit (fixture written by the
    reference build)."""
    meta, ops = load_xfer(name)
    _, _, st, _ = load_golden(meta["base"])
    for op, (pos, val) in ops.items():
        src = meta["ops"][op]
        ot = O.OracleTree(st, n_var=4)
        if meta["cfg"].get("cyl"):
            ot.set_coord(True)
        ot.set(4, st.data[src - 1])
        run_xfer(ot, op, st.min_lvl, st.highest_lvl)
        assert_bitexact(ot.get(4), xfer_expected(st.data[src - 1], pos, val), f"{name} {op}")
        for iv in (1, 2, 3):
            assert_bitexact(ot.get(iv), st.data[iv - 1], f"{name} {op}: variable {iv}")


@pytest.mark.parametrize("name", XFERS)
def test_oracle_second_ghost_layer_matches_reference(name):
    """a$D_gc2_box with a$D_gc2_prolong1 and a$D_bc2_neumann_zero
    (m_aX_gc.f90:420-634) of variable 4 after the gc0 step (phi with
    a$D_gc_prolong0 and Neumann-zero ghosts), every box: the oracle equals
    the reference's dump bit for bit.  The fixture trees hold refinement
    boundaries, same-level neighbours and physical faces."""
    meta, ops = load_xfer(name)
    _, _, st, _ = load_golden(meta["base"])
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    ot = O.OracleTree(st, n_var=4)
    if meta["cfg"].get("cyl"):
        ot.set_coord(True)
    pos, val = ops["gc0"]
    ot.set(4, xfer_expected(st.data[0], pos, val))
    assert (st.neighbors == 0).any() and (st.neighbors > 0).any() and (st.neighbors < 0).any()
    assert_bitexact(ot.gc2(4), z["gc2"], name)


@pytest.mark.parametrize("name", XFERS)
def test_oracle_consistent_fluxes_match_reference(name):
    """a$D_consistent_fluxes on face variable 1 = the gradient of phi: the
    oracle reproduces the reference's faces bit for bit."""
    from helpers import face_gradient
    meta, _ = load_xfer(name)
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    _, _, st, _ = load_golden(meta["base"])
    before = face_gradient(st, st.data[0])
    ot = O.OracleTree(st)
    if meta["cfg"].get("cyl"):
        ot.set_coord(True)
    ot.set_fc(1, before)
    ot.set_fc(2, before)
    ot.consistent_fluxes([1])
    assert_bitexact(ot.get_fc(1), xfer_expected(before, z["flx_pos"], z["flx_val"]), name)
    assert_bitexact(ot.get_fc(2), before, "face variable 2 untouched")


def test_cylindrical_flux_fixture_needs_the_weights():
    """The cylindrical flux fixture pins the radius weights: the Cartesian
    mean (the same tree without cylindrical coordinates) differs from it on
    y faces only, and a weight taken at the fine cell's radius instead of
    the coarse one would differ too (checked by the weights' closed form)."""
    from helpers import face_gradient
    name = "xfer_g2d_nc4_cyl"
    meta, _ = load_xfer(name)
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    _, _, st, _ = load_golden(meta["base"])
    before = face_gradient(st, st.data[0])
    want = xfer_expected(before, z["flx_pos"], z["flx_val"])
    ot = O.OracleTree(st)
    ot.set_fc(1, before)
    ot.consistent_fluxes([1])
    cart = ot.get_fc(1)
    diff = np.nonzero(cart != want)[0]
    assert diff.size > 0
    # every differing value is a y face: offsets within a box past its fx block
    foff, nc = st.face_offsets(), st.ncell()
    box = np.searchsorted(foff, diff, side="right") - 1
    local = diff - foff[box]
    assert np.all(local >= (nc[box] + 1) * nc[box])
    # weights w = 1 -/+ 0.25 dr / r: their mean is 1, so a flux constant
    # along x restricts to itself, and the weighted mean exceeds the plain
    # one where the outer flux is the larger
    a, b, dr, r = 1.0, 3.0, 0.125, 0.3125
    t = 0.25 * dr / r
    assert 0.5 * ((1 - t) * a + (1 + t) * b) > 0.5 * (a + b)
    assert 0.5 * ((1 - t) * 2.0 + (1 + t) * 2.0) == 2.0
