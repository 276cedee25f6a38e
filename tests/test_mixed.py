"""Mixed precision: the FP32 multigrid stages track the FP64 reference
arithmetic to FP32 accuracy, and FP64 defect correction with FP32
corrections reaches the tolerance with the FP64 solution.  Device tests;
also run through the CPU emulation (test_emul.py)."""
import numpy as np
import pytest

import oracle as O
from afivo_amd import mg as M
from afivo_amd import problems as P

pytestmark = pytest.mark.gpu


def mixed_problem(kind):
    if kind == "2d":
        return P.small_2d()
    if kind == "3d":
        return P.small_3d()
    return P.small_3d(nc=8, l_uniform=2, l_max=4, eps_hi=5.0)


@pytest.mark.parametrize("kind", ["2d", "3d", "3d_eps"])
def test_fp32_multigrid_tracks_fp64(kind):
    """FMG + 2 V-cycles on the FP32 copies agree with the oracle's FP64
    result to FP32 accuracy (and differ from it: the arithmetic is FP32)."""
    p = mixed_problem(kind)
    st = p.state
    nv = len(st.data)
    eps = p.eps_hi is not None
    dt = M.DeviceTree.from_state(st, nv + 1)
    mg = M.default_mg(eps=eps, n_base=p.n_base)
    for iv in range(1, nv + 1):
        dt.fp32_from_cc(iv, iv)
    M.mg_fas_fmg_fp32(dt, mg, False, False)
    for _ in range(2):
        M.mg_fas_vcycle_fp32(dt, mg, dt.highest_lvl, False)
    dt.fp32_to_cc(1, nv + 1)
    phi32 = dt.download(nv + 1)
    ot = O.OracleTree(st, n_var=nv)
    om = O.default_opts(2, 2, p.n_base, eps=eps)
    ot.fas_fmg(om, False, False)
    for _ in range(2):
        ot.fas_vcycle(om, set_residual=False)
    phi64 = ot.get(1)
    mask = st.interior_mask()
    scale = np.abs(phi64[mask]).max()
    err = np.abs(phi32[mask] - phi64[mask]).max() / scale
    assert 0 < err < 1e-5, err
    dt.close()


@pytest.mark.parametrize("kind", ["2d", "3d_eps"])
def test_mixed_solve_matches_fp64_solution(kind):
    """Defect correction reaches 1e-10 and its phi agrees with FP64 solves
    to the same tolerance (the device FP64 solver and the oracle) within
    1e-9 of max|phi|."""
    p = mixed_problem(kind)
    st = p.state
    eps = p.eps_hi is not None
    mg = M.default_mg(eps=eps, n_base=p.n_base)
    tol = 1e-10
    dt = M.DeviceTree.from_state(st)
    it, rr = M.mixed_solve(dt, mg, tol, max_iter=30)
    assert rr <= tol and it >= 2
    phi_mixed = dt.download(1)
    dt.close()
    # FP64 on the device: FMG, then V-cycles to the tolerance
    d64 = M.DeviceTree.from_state(st)
    rhs_norm = M.max_abs_cc(d64, 2)
    M.mg_fas_fmg(d64, mg, True, False)
    k = 0
    while M.max_abs_cc(d64, 3) > tol * rhs_norm and k < 30:
        M.mg_fas_vcycle(d64, mg, d64.highest_lvl, True)
        k += 1
    phi64 = d64.download(1)
    d64.close()
    # the oracle, with as many V-cycles
    ot = O.OracleTree(st)
    om = O.default_opts(2, 2, p.n_base, eps=eps)
    ot.fas_fmg(om, True, False)
    for _ in range(k):
        ot.fas_vcycle(om)
    mask = st.interior_mask()
    scale = np.abs(phi64[mask]).max()
    assert np.abs(phi_mixed[mask] - phi64[mask]).max() <= 1e-9 * scale
    assert np.abs(phi_mixed[mask] - ot.get(1)[mask]).max() <= 1e-9 * scale


@pytest.mark.parametrize("kind", ["2d", "3d_eps"])
def test_mixed_solve_with_fp64_coarsest_solve(kind):
    """The FP32 corrections' coarsest-grid solve in FP64 arithmetic
    (afg_mixed_fp64_base): the defect correction still reaches 1e-10, its phi
    agrees with the default (FP32 coarsest solve) within 1e-9 of max|phi|,
    and switching the option back gives the default's phi bit for bit."""
    p = mixed_problem(kind)
    st = p.state
    mg = M.default_mg(eps=p.eps_hi is not None, n_base=p.n_base)
    tol = 1e-10
    dt = M.DeviceTree.from_state(st)
    M.mixed_solve(dt, mg, tol, max_iter=30)
    phi_default = dt.download(1)
    for iv, a in enumerate(st.data, start=1):
        dt.upload(iv, a)
    dt.mixed_fp64_base(True)
    it, rr = M.mixed_solve(dt, mg, tol, max_iter=30)
    assert rr <= tol and it >= 2
    phi_base64 = dt.download(1)
    for iv, a in enumerate(st.data, start=1):
        dt.upload(iv, a)
    dt.mixed_fp64_base(False)
    M.mixed_solve(dt, mg, tol, max_iter=30)
    assert np.array_equal(dt.download(1).view(np.int64), phi_default.view(np.int64))
    dt.close()
    mask = st.interior_mask()
    scale = np.abs(phi_default[mask]).max()
    assert np.abs(phi_base64[mask] - phi_default[mask]).max() <= 1e-9 * scale


@pytest.mark.parametrize("kind", ["2d", "3d_eps"])
def test_mixed_solve_with_fp64_tau(kind):
    """The FP32 corrections' update_coarse (the FAS tau) in FP64 arithmetic
    (afg_mixed_fp64_tau): the defect correction still reaches 1e-10, its phi
    agrees with the default (FP32 tau) within 1e-9 of max|phi| but not bit for
    bit (the option changes the arithmetic), and switching the option back
    gives the default's phi bit for bit."""
    p = mixed_problem(kind)
    st = p.state
    mg = M.default_mg(eps=p.eps_hi is not None, n_base=p.n_base)
    tol = 1e-10
    dt = M.DeviceTree.from_state(st)
    M.mixed_solve(dt, mg, tol, max_iter=30)
    phi_default = dt.download(1)
    for iv, a in enumerate(st.data, start=1):
        dt.upload(iv, a)
    dt.mixed_fp64_tau(True)
    it, rr = M.mixed_solve(dt, mg, tol, max_iter=30)
    assert rr <= tol and it >= 2
    phi_tau64 = dt.download(1)
    for iv, a in enumerate(st.data, start=1):
        dt.upload(iv, a)
    dt.mixed_fp64_tau(False)
    M.mixed_solve(dt, mg, tol, max_iter=30)
    assert np.array_equal(dt.download(1).view(np.int64), phi_default.view(np.int64))
    dt.close()
    mask = st.interior_mask()
    scale = np.abs(phi_default[mask]).max()
    assert np.abs(phi_tau64[mask] - phi_default[mask]).max() <= 1e-9 * scale
    assert not np.array_equal(phi_tau64.view(np.int64), phi_default.view(np.int64))


@pytest.mark.parametrize("kind", ["2d", "3d_eps"])
def test_mixed_correction_rhs_paths_agree(kind):
    """A correction takes its FP32 right-hand side straight from the FP64
    residual kernel that ended the step before, or, after any write of cell
    data in between, by converting tmp: both give the same phi bit for bit."""
    p = mixed_problem(kind)
    st = p.state
    mg = M.default_mg(eps=p.eps_hi is not None, n_base=p.n_base)
    dt = M.DeviceTree.from_state(st)
    M.mixed_begin(dt, mg)
    for _ in range(3):
        M.mixed_correct(dt, mg, 2)
    phi_direct = dt.download(1)
    for iv, a in enumerate(st.data, start=1):
        dt.upload(iv, a)
    M.mixed_begin(dt, mg)
    for _ in range(3):
        dt.upload(2, st.data[1])   # the same rhs again: cell data written
        M.mixed_correct(dt, mg, 2)
    phi_converted = dt.download(1)
    dt.close()
    assert np.array_equal(phi_converted.view(np.int64), phi_direct.view(np.int64))


def test_c4_configuration_against_the_oracle():
    """The 3D variable-eps configuration itself (nc 16, 2804 boxes, 1.00e7
    leaf cells, levels -2..7): the FP64 GPU solve (FMG + 11 V-cycles, the
    count that reaches 1e-10) equals the oracle's bit for bit, every variable
    and every residual norm; the mixed-precision solve to 1e-10 (3 FP32
    V-cycles per correction) is within 1e-9 (relative, max norm) of the
    oracle's phi.  The oracle's solve takes about 20 s on one host core."""
    p = P.c4(seed=1)
    st = p.state
    nv = len(st.data)
    mg = M.default_mg(eps=True, n_base=p.n_base)
    om = O.default_opts(2, 2, p.n_base, eps=True)
    dt = M.DeviceTree.from_state(st, nv)
    ot = O.OracleTree(st, n_var=nv)
    rd, ro = [], []
    M.mg_fas_fmg(dt, mg, True, False)
    ot.fas_fmg(om, True, False)
    rd.append(M.max_abs_cc(dt, 3))
    ro.append(ot.max_abs(3))
    for _ in range(11):
        M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
        ot.fas_vcycle(om)
        rd.append(M.max_abs_cc(dt, 3))
        ro.append(ot.max_abs(3))
    assert rd == ro
    for iv in range(1, nv + 1):
        a, b = dt.download(iv), ot.get(iv)
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), f"variable {iv} differs from the oracle"
    mask = st.leaf_mask()
    rhs_norm = float(np.abs(st.data[1][mask]).max())
    assert rd[-1] <= 1e-10 * rhs_norm
    phi_oracle = ot.get(1)
    for iv, a in enumerate(st.data, start=1):   # start again from the problem's data
        dt.upload(iv, a)
    it, rel = M.mixed_solve(dt, mg, 1e-10, max_iter=10, cycles=3)
    assert rel <= 1e-10
    phi_mixed = dt.download(1)
    err = float(np.abs(phi_mixed[mask] - phi_oracle[mask]).max() / np.abs(phi_oracle[mask]).max())
    print(f"C4: mixed phi against the oracle's FP64 phi: {err:.2e} (relative), {it} corrections")
    assert err <= 1e-9
    dt.close()
