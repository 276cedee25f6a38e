"""HIP path against the oracle, through the C ABI, bit for bit.

Stage tests start every variable (ghosts and corners included) from random
values and compare all variables after one stage on one level; solve tests
run FMG and V-cycles on the benchmark trees and on the reference's own golden
fixtures.  The comparison is bitwise (sign of zero included): the kernels keep
the reference's operation order and are compiled without FMA contraction.
"""
import zlib

import numpy as np
import pytest

import oracle as O
from afivo_amd import mg as M
from afivo_amd import problems as P
from afivo_amd import tree as T
from afivo_amd._lib import AfgError
from helpers import assert_bitexact, random_state
from test_oracle import CASES, load_golden, opts_of, rstr_of

pytestmark = pytest.mark.gpu


def pair(st, n_var, data):
    dt = M.DeviceTree(st.dim, st.nc, n_var, st.recs, st.dr_base)
    ot = O.OracleTree(st, n_var=n_var)
    for iv, a in enumerate(data, start=1):
        dt.upload(iv, a)
        ot.set(iv, a)
    return dt, ot


def compare_all(dt, ot, n_var, what):
    for iv in range(1, n_var + 1):
        assert_bitexact(dt.download(iv), ot.get(iv), f"{what} var {iv}")


def problem(kind):
    if kind == "2d_nc8":
        return P.small_2d(nc=8)
    if kind == "2d_nc4":
        return P.small_2d(nc=4, l_uniform=3, l_max=6)
    if kind == "2d_nc16":
        return P.small_2d(nc=16, l_uniform=2, l_max=4)
    if kind == "3d_nc8":
        return P.small_3d(nc=8)
    if kind == "3d_nc4_eps":
        return P.small_3d(nc=4, l_uniform=2, l_max=4, eps_hi=4.0)
    if kind == "2d_nc8_eps":
        return P.make_problem("e2", 2, 8, P.C1_GAUSS, 3, 5, 1e-3, n_base=4, n_vcycle=2, eps_hi=7.0)
    if kind == "2d_nc32":          # a cell count without a specialised kernel
        return P.small_2d(nc=32, l_uniform=2, l_max=4)
    if kind == "3d_nc16_eps":
        return P.small_3d(nc=16, l_uniform=1, l_max=3, eps_hi=3.0)
    raise KeyError(kind)


KINDS = ["2d_nc8", "2d_nc4", "2d_nc16", "3d_nc8", "3d_nc4_eps", "2d_nc8_eps", "2d_nc32", "3d_nc16_eps"]
STAGES = ["fill_gc", "gsrb", "residual", "restrict", "prolong2", "prolong2_add", "correct", "update_coarse"]


def stage_state(p, seed):
    st = p.state
    n_var = len(st.data)
    data = random_state(st, n_var, seed)
    if n_var == 4:              # keep eps positive and piecewise constant
        data[3] = st.data[3].copy()
    return n_var, data


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("stage", STAGES)
def test_stage_bitexact(kind, stage):
    p = problem(kind)
    st = p.state
    n_var, data = stage_state(p, seed=zlib.crc32(f"{kind}{stage}".encode()) % 1000)
    eps = n_var == 4
    mg = M.default_mg(eps=eps)
    om = O.default_opts(eps=eps)
    L = O.lib()
    for lvl in range(st.min_lvl, st.highest_lvl + 1):
        dt, ot = pair(st, n_var, data)
        has_parent = lvl > st.min_lvl
        if stage == "fill_gc":
            dt.fill_gc(1, lvl)
            L.orc_fill_gc_level(ot.h, 1, lvl)
        elif stage == "gsrb":
            dt.gsrb_level(mg, lvl, 2)
            L.orc_gsrb_level(ot.h, om, lvl, 2)
        elif stage == "residual":
            dt.residual_level(mg, lvl)
            L.orc_residual_level(ot.h, om, lvl)
        elif not has_parent:
            continue
        elif stage == "restrict":
            dt.restrict_level(2, lvl)
            L.orc_restrict_level(ot.h, 2, lvl)
        elif stage == "prolong2":
            dt.prolong2_level(1, 3, lvl, False)
            L.orc_prolong2_level(ot.h, 1, 3, lvl, 0)
        elif stage == "prolong2_add":
            dt.prolong2_level(3, 1, lvl, True)
            L.orc_prolong2_level(ot.h, 3, 1, lvl, 1)
        elif stage == "correct":
            dt.correct_children(mg, lvl)
            L.orc_correct_children_level(ot.h, om, lvl)
        elif stage == "update_coarse":
            dt.update_coarse(mg, lvl)
            L.orc_update_coarse_level(ot.h, om, lvl)
        compare_all(dt, ot, n_var, f"{kind} {stage} lvl {lvl}")
        dt.close()


def solve_both(st, n_var, mg, om, n_fmg, n_vcycle, dt=None, ot=None):
    dt = dt or M.DeviceTree.from_state(st, n_var)
    ot = ot or O.OracleTree(st, n_var=n_var)
    rd, ro = [], []
    for it in range(n_fmg):
        M.mg_fas_fmg(dt, mg, True, it > 0)
        ot.fas_fmg(om, True, it > 0)
        rd.append(M.max_abs_cc(dt, 3))
        ro.append(ot.max_abs(3))
    for _ in range(n_vcycle):
        M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
        ot.fas_vcycle(om)
        rd.append(M.max_abs_cc(dt, 3))
        ro.append(ot.max_abs(3))
    return dt, ot, np.array(rd), np.array(ro)


@pytest.mark.parametrize("name", CASES)
def test_golden_solve_bitexact(name):
    """The reference's own output (golden fixture) reproduced on the device."""
    cfg, st_in, st_out, res_ref = load_golden(name)
    c = cfg["cfg"]
    eps = c.get("use_eps", False)
    lsf = c.get("lsf_mode", 0)
    mg = M.default_mg(eps=eps, n_down=c.get("n_down", 2), n_up=c.get("n_up", 2), n_base=c["n_base"],
                      auto=c.get("cyl", False), lsf=lsf == 1, lsf_set=lsf == 2, rstr=rstr_of(c))
    dt = M.DeviceTree.from_state(st_in)
    if c.get("cyl"):
        dt.set_coord(True)
    res = []
    for it in range(c.get("n_fmg", 1)):
        M.mg_fas_fmg(dt, mg, True, it > 0)
        res.append(M.max_abs_cc(dt, 3))
    for _ in range(c.get("n_vcycle", 0)):
        M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
        res.append(M.max_abs_cc(dt, 3))
    for iv in range(1, st_out.n_var + 1):
        assert_bitexact(dt.download(iv), st_out.data[iv - 1], f"{name} var {iv}")
    assert np.array_equal(np.array(res), res_ref)


@pytest.mark.parametrize("kind", KINDS)
def test_solve_bitexact(kind):
    p = problem(kind)
    eps = p.eps_hi is not None
    mg = M.default_mg(eps=eps, n_base=p.n_base)
    om = O.default_opts(2, 2, p.n_base, eps=eps)
    dt, ot, rd, ro = solve_both(p.state, len(p.state.data), mg, om, 2, 2)
    compare_all(dt, ot, len(p.state.data), kind)
    assert np.array_equal(rd, ro)


def test_c1_full_config_bitexact():
    """The 2D benchmark configuration: FMG + 10 V-cycles, 2+2 sweeps, 50 base."""
    p = P.c1()
    mg = M.default_mg(n_base=50)
    om = O.default_opts(2, 2, 50)
    dt, ot, rd, ro = solve_both(p.state, 3, mg, om, 1, 10)
    compare_all(dt, ot, 3, "c1")
    assert np.array_equal(rd, ro)
    assert rd[-1] < 1e-9


def test_c2_full_config_bitexact():
    """The 3D benchmark configuration (2.6e6 leaf cells): FMG + 8 V-cycles."""
    p = P.c2()
    mg = M.default_mg(n_base=50)
    om = O.default_opts(2, 2, 50)
    dt, ot, rd, ro = solve_both(p.state, 3, mg, om, 1, 8)
    compare_all(dt, ot, 3, "c2")
    assert np.array_equal(rd, ro)
    # per-cycle residual reduction factors
    f = rd[1:] / rd[:-1]
    assert (f < 0.2).all(), f


def test_graphs_and_direct_launches_agree():
    p = P.small_3d()
    mg = M.default_mg(n_base=p.n_base)
    out = []
    for graphs in (True, False):
        dt = M.DeviceTree.from_state(p.state)
        dt.set_graphs(graphs)
        M.mg_fas_fmg(dt, mg, True, False)
        for _ in range(3):
            M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
        out.append([dt.download(iv) for iv in (1, 2, 3)])
    for a, b in zip(*out):
        assert_bitexact(a, b, "graph vs direct")


def launches_per_block(n_cycle, pairs):
    """Launches of a relaxation block of 2 n_cycle passes: fused pairs while
    an even number of pairs is left, then single passes."""
    n_pass = 2 * n_cycle
    return 2 * (n_pass // 4) + n_pass % 4 if pairs else n_pass


@pytest.mark.parametrize("one_level", [False, True])
@pytest.mark.parametrize("pairs", [True, False])
def test_timing_mode_is_exact(one_level, pairs):
    """Event timing (graphs with event records, used in turn from a ring of
    copies) leaves results unchanged and counts every timed launch."""
    p = P.small_2d()
    mg = M.default_mg(n_base=p.n_base)
    om = O.default_opts(2, 2, p.n_base)
    dt = M.DeviceTree.from_state(p.state)
    dt.set_bottom_cells(0)      # every level through the timed per-level passes
    dt.set_pass_pairs(pairs)
    top = p.state.highest_lvl
    dt.set_timing(True, top if one_level else -1000)
    n_cyc = 11                  # more V-cycles than graph copies in the ring
    dt, ot, rd, ro = solve_both(p.state, 3, mg, om, 1, n_cyc, dt=dt)
    n, ms, b = dt.timing()
    assert ms > 0 and b > 0
    per_cycle = launches_per_block(mg.n_cycle_down, pairs) + launches_per_block(mg.n_cycle_up, pairs)
    if one_level:
        # FMG's last V-cycle, then the V-cycles
        assert n == (n_cyc + 1) * per_cycle
    else:
        assert n > (n_cyc + 1) * per_cycle
    compare_all(dt, ot, 3, "timing")
    assert np.array_equal(rd, ro)


def test_boundary_values_and_neumann():
    """Non-zero Dirichlet values and Neumann faces from the boundary table."""
    p = P.small_2d()
    st = p.state
    dt = M.DeviceTree.from_state(st)
    ot = O.OracleTree(st)
    rng = np.random.default_rng(7)
    nc_of = dict(zip(st.ids.tolist(), st.ncell().tolist()))
    for r in range(st.n_boxes):
        for f in range(4):
            if st.neighbors[r, f] < 0:
                id_ = int(st.ids[r])
                b = rng.standard_normal(nc_of[id_])
                typ = -2 if (f == 0 and r % 2 == 0) else -1
                dt.set_bc(id_, f + 1, typ, b)
                ot.set_bc(id_, f + 1, typ, b)
    mg = M.default_mg(n_base=p.n_base)
    om = O.default_opts(2, 2, p.n_base)
    dt, ot, rd, ro = solve_both(st, 3, mg, om, 1, 2, dt=dt, ot=ot)
    compare_all(dt, ot, 3, "bc")
    assert np.array_equal(rd, ro)


def test_periodic_base():
    """Two base boxes periodic in x (neighbours through the base list)."""
    t = T.Tree(2, 8, 1.0 / 16, coarsen_to=2)
    t.set_base([[1, 1], [2, 1]], [[2, 2, -1, -1], [1, 1, -1, -1]])
    gs = P.Gaussians(np.array([1.0]), np.array([0.1]), np.array([[0.9, 0.5]]))
    for _ in range(10):
        if t.adjust_refinement(P.gauss_refinement(gs, 2, 4, 1e-3)) == 0:
            break
    st = t.state()
    rhs = P.set_rhs(t, st, gs)
    z = np.zeros_like(rhs)
    st.data = [z.copy(), rhs, z.copy()]
    st.n_var = 3
    mg = M.default_mg(n_base=8)
    om = O.default_opts(2, 2, 8)
    dt, ot, rd, ro = solve_both(st, 3, mg, om, 1, 2)
    compare_all(dt, ot, 3, "periodic")


@pytest.mark.parametrize("kind", ["3d", "3d_eps", "2d"])
@pytest.mark.parametrize("cells,wave_cells", [(0, 0), (512, 0), (512, 512), (1 << 20, 0),
                                              (1 << 20, 64), (1 << 20, 1 << 20)])
def test_bottom_kernel_limit_is_exact(kind, cells, wave_cells):
    """The one-workgroup bottom kernel (with or without its one-wavefront
    part) and per-level launches give the same bits, whichever levels each
    takes."""
    if kind == "2d":
        p = P.small_2d()
    else:
        p = P.small_3d(eps_hi=4.0 if kind == "3d_eps" else None)
    eps = p.eps_hi is not None
    mg = M.default_mg(eps=eps, n_base=p.n_base)
    om = O.default_opts(2, 2, p.n_base, eps=eps)
    n_var = len(p.state.data)
    dt = M.DeviceTree.from_state(p.state)
    dt.set_bottom_cells(cells, wave_cells)
    dt, ot, rd, ro = solve_both(p.state, n_var, mg, om, 1, 2, dt=dt)
    compare_all(dt, ot, n_var, f"{kind} bottom {cells}/{wave_cells}")
    assert np.array_equal(rd, ro)


@pytest.mark.parametrize("kind", ["2d_nc8", "2d_nc16", "3d_nc8", "3d_nc4_eps", "3d_nc16_eps"])
@pytest.mark.parametrize("wgs,cells", [(1, 1 << 30), (4, 512), (16, 4096), (32, 1 << 30)])
def test_persistent_kernel_is_exact(kind, wgs, cells):
    """The persistent kernel (the lower levels' stages in one launch of wgs
    workgroups meeting at barriers) gives the oracle's bits, whichever levels
    it takes: a few low ones (per-level launches above it, the residual of its
    levels in a launch of its own) or the whole tree (the residual inside it);
    FP32 stages give the same bits with it as without it."""
    p = problem(kind)
    eps = p.eps_hi is not None
    mg = M.default_mg(eps=eps, n_base=p.n_base)
    om = O.default_opts(2, 2, p.n_base, eps=eps)
    n_var = len(p.state.data)
    dt = M.DeviceTree.from_state(p.state)
    dt.set_persist(wgs, cells)
    w, c, top = dt.persist()
    assert (w, c) == (wgs, cells) and top >= dt.min_lvl
    assert top == dt.highest_lvl or cells < 1 << 30
    dt, ot, rd, ro = solve_both(p.state, n_var, mg, om, 1, 2, dt=dt)
    compare_all(dt, ot, n_var, f"{kind} persist {wgs}/{cells}")
    assert np.array_equal(rd, ro)
    dt.close()
    # FP32 stages: without and with the persistent kernel
    out = []
    for on in (False, True):
        d = M.DeviceTree.from_state(p.state, n_var + 1)
        d.set_persist(wgs if on else 0, cells)
        for iv in range(1, n_var + 1):
            d.fp32_from_cc(iv, iv)
        M.mg_fas_fmg_fp32(d, mg, True, False)
        M.mg_fas_vcycle_fp32(d, mg, d.highest_lvl, True)
        d.fp32_to_cc(1, n_var + 1)
        out.append(d.download(n_var + 1))
        d.fp32_to_cc(3, n_var + 1)
        out.append(d.download(n_var + 1))
        d.close()
    assert_bitexact(out[2], out[0], f"{kind} FP32 persist phi")
    assert_bitexact(out[3], out[1], f"{kind} FP32 persist residual")


@pytest.mark.parametrize("kind", ["3d_nc8", "3d_nc16_eps"])
@pytest.mark.parametrize("shape", [0, 1, 2])
@pytest.mark.parametrize("n_down,n_up", [(2, 2), (1, 3), (3, 1)])
def test_split_passes_are_exact(kind, shape, n_down, n_up):
    """Relaxation blocks on the colour-split working copy of phi (first,
    middle and last passes) give the oracle's bits with every tile shape and
    block length, in FP64; FP32 stages give the same bits as without it."""
    p = problem(kind)
    eps = p.eps_hi is not None
    mg = M.default_mg(eps=eps, n_down=n_down, n_up=n_up, n_base=p.n_base)
    om = O.default_opts(n_down, n_up, p.n_base, eps=eps)
    n_var = len(p.state.data)
    dt = M.DeviceTree.from_state(p.state)
    dt.set_bottom_cells(0)            # every level through the tiled pass
    for lvl in range(dt.min_lvl, dt.highest_lvl + 1):
        if T.level_ncell(p.state.nc, lvl) in (8, 16):
            dt.set_tile(lvl, shape)
            dt.set_level_pass_pairs(lvl, False)
    dt, ot, rd, ro = solve_both(p.state, n_var, mg, om, 1, 2, dt=dt)
    compare_all(dt, ot, n_var, f"{kind} split shape {shape} {n_down}/{n_up}")
    assert np.array_equal(rd, ro)
    dt.close()
    out = []
    for split in (False, True):
        d = M.DeviceTree.from_state(p.state, n_var + 1)
        d.set_split_passes(split)
        d.set_bottom_cells(0)
        for lvl in range(d.min_lvl, d.highest_lvl + 1):
            if T.level_ncell(p.state.nc, lvl) in (8, 16):
                d.set_tile(lvl, shape, fp32=True)
                d.set_level_pass_pairs(lvl, False, fp32=True)
        for iv in range(1, n_var + 1):
            d.fp32_from_cc(iv, iv)
        M.mg_fas_fmg_fp32(d, mg, True, False)
        M.mg_fas_vcycle_fp32(d, mg, d.highest_lvl, True)
        d.fp32_to_cc(1, n_var + 1)
        out.append(d.download(n_var + 1))
        d.close()
    assert_bitexact(out[1], out[0], f"{kind} FP32 split shape {shape}")


@pytest.mark.parametrize("kind", ["2d_nc8", "2d_nc16", "3d_nc8", "3d_nc16_eps"])
@pytest.mark.parametrize("shape", [1, 2])
def test_tile_shapes_are_exact(kind, shape):
    """Every tile shape of the colour pass (workgroup size, boxes per
    workgroup) on every level gives the oracle's bits, in FP64; FP32 stages
    give the same bits with every shape as with the default one."""
    p = problem(kind)
    eps = p.eps_hi is not None
    mg = M.default_mg(eps=eps, n_base=p.n_base)
    om = O.default_opts(2, 2, p.n_base, eps=eps)
    n_var = len(p.state.data)
    dt = M.DeviceTree.from_state(p.state)
    dt.set_bottom_cells(0)            # every level through the tiled pass
    for lvl in range(dt.min_lvl, dt.highest_lvl + 1):
        if T.level_ncell(p.state.nc, lvl) in (8, 16):
            dt.set_tile(lvl, shape)
            assert dt.tile(lvl) == shape
    dt, ot, rd, ro = solve_both(p.state, n_var, mg, om, 1, 2, dt=dt)
    compare_all(dt, ot, n_var, f"{kind} shape {shape}")
    assert np.array_equal(rd, ro)
    # FP32 stages: shape 0 against this shape
    out = []
    for sh in (0, shape):
        d = M.DeviceTree.from_state(p.state, n_var + 1)
        d.set_bottom_cells(0)
        for lvl in range(d.min_lvl, d.highest_lvl + 1):
            if T.level_ncell(p.state.nc, lvl) in (8, 16):
                d.set_tile(lvl, sh, fp32=True)
        for iv in range(1, n_var + 1):
            d.fp32_from_cc(iv, iv)
        M.mg_fas_fmg_fp32(d, mg, False, False)
        M.mg_fas_vcycle_fp32(d, mg, d.highest_lvl, False)
        d.fp32_to_cc(1, n_var + 1)
        out.append(d.download(n_var + 1))
        d.close()
    assert_bitexact(out[1], out[0], f"{kind} FP32 shape {shape}")
    dt.close()


@pytest.mark.parametrize("kind", ["2d_nc8", "3d_nc16_eps"])
def test_autotune_keeps_state(kind):
    """The autotuner times every shape on a copy of phi: the tree's data is
    unchanged, each tuned level gets a valid shape and solves stay exact."""
    p = problem(kind)
    eps = p.eps_hi is not None
    mg = M.default_mg(eps=eps, n_base=p.n_base)
    om = O.default_opts(2, 2, p.n_base, eps=eps)
    n_var = len(p.state.data)
    dt = M.DeviceTree.from_state(p.state)
    before = [dt.download(iv) for iv in range(1, n_var + 1)]
    for fp32 in (False, True):
        ms, bms = dt.autotune(mg, fp32=fp32, reps=2)
        assert np.all(bms > 0) and dt.bottom_cells(fp32) in (512, 2048, 8192, 32768)
        for lvl in range(dt.min_lvl, dt.highest_lvl + 1):
            row = ms[lvl - dt.min_lvl]
            if T.level_ncell(p.state.nc, lvl) in (8, 16):
                chosen = row[int(dt.pass_pairs(lvl, fp32)), dt.tile(lvl, fp32)]
                assert np.all(row >= 0) and chosen <= 1.03 * row.min()
            else:
                assert np.all(row == -1) and dt.tile(lvl, fp32) == 0
    for iv in range(1, n_var + 1):
        assert_bitexact(dt.download(iv), before[iv - 1], f"{kind} var {iv} after autotune")
    dt, ot, rd, ro = solve_both(p.state, n_var, mg, om, 1, 2, dt=dt)
    compare_all(dt, ot, n_var, f"{kind} tuned")
    assert np.array_equal(rd, ro)


@pytest.mark.parametrize("kind", ["2d_nc8", "2d_nc16", "3d_nc8", "3d_nc16_eps", "2d_nc8_eps"])
@pytest.mark.parametrize("n_down,n_up", [(2, 2), (1, 3), (4, 2)])
def test_pass_pairs_are_exact(kind, n_down, n_up):
    """Fused pass pairs (phi to the second pool and back) and single passes
    give the oracle's bits, for blocks of 2, 4, 6 and 8 passes (pairs and
    single passes mixed); FP32 stages give the same bits either way."""
    p = problem(kind)
    eps = p.eps_hi is not None
    n_var = len(p.state.data)
    mg = M.default_mg(eps=eps, n_down=n_down, n_up=n_up, n_base=p.n_base)
    om = O.default_opts(n_down, n_up, p.n_base, eps=eps)
    res = []
    for pairs in (True, False):
        dt = M.DeviceTree.from_state(p.state)
        dt.set_bottom_cells(0)
        dt.set_pass_pairs(pairs)
        dt, ot, rd, ro = solve_both(p.state, n_var, mg, om, 1, 2, dt=dt)
        compare_all(dt, ot, n_var, f"{kind} pairs {pairs}")
        assert np.array_equal(rd, ro)
        d = M.DeviceTree.from_state(p.state, n_var + 1)
        d.set_bottom_cells(0)
        d.set_pass_pairs(pairs)
        for iv in range(1, n_var + 1):
            d.fp32_from_cc(iv, iv)
        M.mg_fas_fmg_fp32(d, mg, False, False)
        M.mg_fas_vcycle_fp32(d, mg, d.highest_lvl, False)
        d.fp32_to_cc(1, n_var + 1)
        res.append(d.download(n_var + 1))
        d.close()
        dt.close()
    assert_bitexact(res[0], res[1], f"{kind} FP32 pairs")


@pytest.mark.parametrize("eps", [False, True])
def test_cylindrical_solve_bitexact(eps):
    """The 2D benchmark tree (951 boxes of 8^2, levels -1..7) in cylindrical
    coordinates with the auto operators (mg2_box_clpl; mg2_box_clpld on the
    boxes of an eps jump at y = 0.5), FMG (twice, the second from the guess)
    and 3 V-cycles: every variable and the residual history bit for bit
    against the oracle.  FP32 stages are refused on such a tree."""
    p = P.c1()
    st = p.state
    if eps:
        st.data.append(P.set_eps(st, p.tree, 3.0))
    nv = len(st.data)
    dt = M.DeviceTree.from_state(st, nv)
    ot = O.OracleTree(st, n_var=nv)
    dt.set_coord(True)
    ot.set_coord(True)
    mg = M.default_mg(eps=eps, n_base=50, auto=True)
    om = O.default_opts(2, 2, 50, eps=eps, auto=True)
    dt, ot, rd, ro = solve_both(st, nv, mg, om, 2, 3, dt, ot)
    compare_all(dt, ot, nv, f"cylindrical eps={eps}")
    assert np.array_equal(rd, ro)
    assert rd[-1] < 1e-3 * rd[0]
    with pytest.raises(AfgError):
        M.mixed_solve(dt, mg, 1e-8, max_iter=5)
    dt.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("mode", ["auto", "auto_rstr", "set"])
def test_level_set_solve_bitexact(dim, mode):
    """An internal boundary on a circle (2D benchmark tree, 951 boxes of 8^2)
    or a sphere (3D, 8^3 boxes, levels -1..4): level set and boundary value
    in variables 5 and 6.  Modes: the auto operators (boxes tagged
    mg_lsf_box take gsrb/box/corr_lpllsf) with rstr_lpl or with
    mg$D_auto_rstr, and the lpllsf set for every box.  FMG (twice, the second
    from the guess) and 2 V-cycles: every variable and the residual history
    bit for bit against the oracle.  FP32 stages are refused."""
    p = P.c1() if dim == 2 else P.small_3d(nc=8, l_uniform=3, l_max=4)
    st = p.state
    while len(st.data) < 4:
        st.data.append(np.zeros(st.n_values()))
    lsf, bval = P.set_level_set(st, p.tree, [0.45, 0.55, 0.5][:dim], 0.23)
    st.data += [lsf, bval]
    rstr = {"auto": O.RSTR_LPL, "auto_rstr": O.RSTR_AUTO, "set": O.RSTR_LSF}[mode]
    mg = M.default_mg(n_base=p.n_base, lsf=True, lsf_set=mode == "set", rstr=rstr)
    om = O.default_opts(2, 2, p.n_base, lsf=True, lsf_set=mode == "set", rstr=rstr)
    dt = M.DeviceTree.from_state(st, 6)
    ot = O.OracleTree(st, n_var=6)
    if mode == "auto_rstr":
        ot.set_tags(om)
    dt, ot, rd, ro = solve_both(st, 6, mg, om, 2, 2, dt, ot)
    compare_all(dt, ot, 6, f"level set {dim}D {mode}")
    assert np.array_equal(rd, ro)
    assert rd[-1] < 1e-2 * rd[0]
    with pytest.raises(AfgError):
        M.mixed_solve(dt, mg, 1e-8, max_iter=5)
    dt.close()


def test_level_set_and_eps_tag_conflict():
    """A box with both a level-set and an eps tag is an error, where
    mg$D_set_box_tag stops (m_aX_mg.f90:700-701)."""
    p = P.small_3d(nc=8, l_uniform=2, l_max=3, eps_hi=3.0)
    st = p.state
    lsf, bval = P.set_level_set(st, p.tree, [0.5, 0.5, 0.5], 0.3)
    st.data += [lsf, bval]
    dt = M.DeviceTree.from_state(st, 6)
    mg = M.default_mg(eps=True, n_base=p.n_base, lsf=True)
    with pytest.raises(AfgError, match="Cannot set lsf and eps tag"):
        M.mg_fas_fmg(dt, mg, True, False)
    dt.close()


def test_rhs_changes_between_solves():
    """The tiled colour passes read a colour-split copy of rhs; every way of
    changing rhs between solves (upload, scaling, a constant, a copy from
    another variable) and moving levels between the bottom kernel and the
    per-level passes must leave the results equal to the oracle's, bit for
    bit, on a 3D tree of 8^3 boxes (tiled passes on levels 2..4)."""
    p = P.small_3d(nc=8, l_uniform=3, l_max=4)
    st = p.state
    while len(st.data) < 4:
        st.data.append(np.zeros(st.n_values()))
    dt = M.DeviceTree.from_state(st, 4)
    ot = O.OracleTree(st, n_var=4)
    mg = M.default_mg(n_base=p.n_base)
    om = O.default_opts(2, 2, p.n_base)
    rng = np.random.default_rng(11)

    def cycle(what):
        M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
        ot.fas_vcycle(om)
        compare_all(dt, ot, 4, what)

    M.mg_fas_fmg(dt, mg, True, False)
    ot.fas_fmg(om, True, False)
    compare_all(dt, ot, 4, "fmg")
    cycle("first cycle")
    new_rhs = st.data[1] * (1 + 0.1 * rng.standard_normal(st.n_values()))
    dt.upload(2, new_rhs)
    ot.set(2, new_rhs)
    cycle("after upload")
    dt.times_cc(0.5, 2)
    ot.set(2, 0.5 * ot.get(2))
    cycle("after times_cc")
    dt.upload(4, 2 * new_rhs)
    ot.set(4, 2 * new_rhs)
    dt.copy_cc(4, 2)
    ot.set(2, ot.get(4))
    cycle("after copy_cc")
    dt.set_bottom_cells(1 << 20)   # every level that fits LDS into the bottom kernel
    cycle("bottom limit raised")
    dt.set_bottom_cells(0)         # no bottom kernel
    cycle("bottom limit 0")
    dt.set_cc(2, 1.0)
    ot.set(2, np.ones(st.n_values()))
    cycle("after set_cc")
    M.mg_fas_fmg(dt, mg, True, False)
    ot.fas_fmg(om, True, False)
    compare_all(dt, ot, 4, "second fmg")
    dt.close()
