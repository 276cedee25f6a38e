"""Serving loop (continuous batching of Poisson-solve requests at request
granularity).  Device tests; the small ones also run through the CPU
emulation (test_emul.py).

* every served request's phi equals the same request solved alone through
  the same operations, bit for bit (mixed precision and FP64), with forests
  small enough that trees join and leave running forests many times;
* at the serving workload's sizes (2D and 3D trees of 1e4 to ~1e6 leaf
  cells, tolerances 1e-8 and 1e-10): FP64-served phi is bit-identical to
  the oracle's FAS V-cycles from phi = 0 (the same count), and
  mixed-precision-served phi meets its tolerance when the oracle evaluates
  its residual, within a stated error of the oracle's FP64 solution;
* admission keeps the device bytes within the HBM budget, and the budget
  estimate bounds what a forest allocates.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from afivo_amd import mg as M
from afivo_amd import serving as S

pytestmark = pytest.mark.gpu


def small_trace(n=8, rate=2000.0, seed=4):
    return S.make_trace(seed, n, rate, cells=(1e3, 2e4), tols=(1e-8, 1e-10))


def c5_requests(seed=11):
    """Requests of the serving workload's kinds up to ~1e6 leaf cells: 2D and
    3D, both tolerances, 1-4 Gaussians (request_tree)."""
    spec = [(2, 1e4, 1e-8), (2, 2e5, 1e-10), (2, 1.1e6, 1e-8), (3, 3e4, 1e-10), (3, 1.1e6, 1e-8)]
    out = []
    for k, (dim, cells, tol) in enumerate(spec):
        st = S.request_tree(seed, k, dim, cells)
        leaf = st.leaf_mask()
        out.append(S.Request(k, 0.0005 * k, st, tol, float(np.abs(st.data[1][leaf]).max()), int(leaf.sum())))
    return out


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_served_phi_equals_solo_runs(precision):
    """One lane of at most 3 trees, compacted whenever a member leaves (so
    that requests leave the running forest one by one and waiting ones of
    its class join it): every request meets its tolerance, takes as many
    steps as alone and gets bit-identical phi."""
    log = {}
    mix = S.serve(small_trace(rate=1e6), max_lanes=1, max_trees=3, compact_at=0.0, n_base=20,
                  precision=precision, log=log)
    solo = S.one_at_a_time(small_trace(), n_base=20, precision=precision)
    assert len({r.state.dim for r in mix}) == 2, "the trace should mix 2D and 3D"
    assert log["joined"] > 0, "requests should join running forests"
    for a, b in zip(sorted(mix, key=lambda r: r.rid), sorted(solo, key=lambda r: r.rid)):
        assert a.finish >= a.start >= a.arrival >= 0
        assert a.residual <= a.tol and b.residual <= b.tol
        assert a.cycles == b.cycles, f"request {a.rid}: {a.cycles} steps served, {b.cycles} alone"
        assert np.array_equal(bits(a.phi), bits(b.phi)), f"request {a.rid}"


def oracle_solution(r, n_base, cycles):
    """The oracle on request r's tree from phi = 0: `cycles` FAS V-cycles
    (mg$D_fas_vcycle with the residual stored), then phi and the relative
    residual."""
    ot = O.OracleTree(r.state)
    opts = O.default_opts(2, 2, n_base)
    for _ in range(cycles):
        ot.fas_vcycle(opts)
    return ot.get(1), ot.max_abs(3) / r.rhs_norm


def oracle_residual(r, phi, n_base):
    """max |rhs - A phi| / max |rhs| over leaf interiors of phi, evaluated by
    the oracle (ghost fill from the lowest level up, then residual_box)."""
    st = r.state
    ot = O.OracleTree(st)
    ot.set(1, phi)
    L = O.lib()
    om = O.default_opts(2, 2, n_base)
    for lvl in range(st.min_lvl, st.highest_lvl + 1):
        L.orc_fill_gc_level(ot.h, 1, lvl)
        L.orc_residual_level(ot.h, om, lvl)
    return ot.max_abs(3) / r.rhs_norm


def test_fp64_serving_matches_the_oracle_at_workload_sizes():
    """FP64-served phi at the serving workload's sizes (2D and 3D, up to ~1e6
    leaf cells, tolerances 1e-8 and 1e-10) is bit-identical to the oracle's
    phi after as many V-cycles from zero, every cell and ghost; the residual
    the oracle stores then meets the tolerance."""
    reqs = S.serve(c5_requests(), max_trees=3, n_base=50, precision="fp64")
    assert max(r.leaf_cells for r in reqs) > 9e5 and {r.state.dim for r in reqs} == {2, 3}
    for r in reqs:
        phi, res = oracle_solution(r, 50, r.cycles)
        print(f"request {r.rid}: {r.state.dim}D, {r.leaf_cells} leaf cells, tol {r.tol:g}, "
              f"{r.cycles} V-cycles, residual {res:.3e}")
        assert np.array_equal(bits(r.phi), bits(phi)), f"request {r.rid}"
        assert res == r.residual and res <= r.tol


def test_mixed_serving_against_the_oracle_at_workload_sizes():
    """Mixed-precision-served phi at the serving workload's sizes: the
    oracle's own evaluation of its residual meets the request's tolerance
    (within 5% for the residual's rounding in the FP64 defect correction),
    and its error against the oracle's FP64 solution -- V-cycles from zero
    until the oracle's residual is below tol / 100 -- is below
    10 * tol * max|phi| (measured 0.02-0.4 tol: printed per request)."""
    reqs = S.serve(c5_requests(), max_trees=3, n_base=50, precision="mixed")
    for r in reqs:
        res = oracle_residual(r, r.phi, 50)
        ot = O.OracleTree(r.state)
        opts = O.default_opts(2, 2, 50)
        for k in range(S.MAX_STEPS):
            ot.fas_vcycle(opts)
            if ot.max_abs(3) / r.rhs_norm <= r.tol / 100:
                break
        ref = ot.get(1)
        leaf = r.state.leaf_mask()
        err = np.abs(r.phi[leaf] - ref[leaf]).max() / np.abs(ref[leaf]).max()
        print(f"request {r.rid}: {r.state.dim}D, {r.leaf_cells} leaf cells, tol {r.tol:g}, "
              f"{r.cycles} corrections, oracle residual {res:.3e}, error vs oracle FP64 {err:.3e}")
        assert res <= 1.05 * r.tol, f"request {r.rid}: oracle residual {res:.3e}"
        assert err <= 10 * r.tol, f"request {r.rid}: error {err:.3e}"


def test_serving_admits_by_memory_budget():
    """With a budget that holds the largest single-tree forest but no two
    trees, requests are still all served, one at a time; a request larger
    than the budget is an error."""
    tr = small_trace(n=4)
    for r in tr[1:]:                      # four requests of one tree, arriving together
        r.state, r.tol, r.rhs_norm = tr[0].state, tr[0].tol, tr[0].rhs_norm
    one = S.forest_bytes([tr[0].state])
    assert S.forest_bytes([tr[0].state] * 2) > one
    out = S.serve(tr, hbm_budget=one, max_lanes=4, max_trees=4, n_base=20)
    assert all(r.residual <= r.tol for r in out)
    spans = sorted((r.start, r.finish) for r in out)
    for (s0, f0), (s1, f1) in zip(spans, spans[1:]):
        assert s1 >= f0, "two requests were resident at once"
    with pytest.raises(RuntimeError, match="does not fit"):
        S.serve(small_trace(n=2), hbm_budget=one // 4, n_base=20)


def _mem_free():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_forest_bytes_bounds_the_allocations(precision):
    """The admission estimate bounds a forest's device bytes (the context's
    exact allocation ledger, afg_device_bytes) after set-up and steps, and
    the device memory the forest takes (hipMemGetInfo before and after)."""
    states = [r.state for r in small_trace(n=6) if r.state.dim == 3][:3]
    states += [r.state for r in c5_requests()[4:5]]
    from afivo_amd import _lib
    emulated = "emul" in getattr(_lib._lib, "_name", "")
    free0 = None if emulated else _mem_free()
    pool = S.Lane((3, 8, False), M.default_mg(n_base=20), precision, 2, True)
    reqs = [S.Request(k, 0.0, st, 1e-8, 1.0, 1) for k, st in enumerate(states)]
    pool.rebuild([], reqs)
    for _ in range(2):
        pool.step()
        pool.check()
    used = pool.device_bytes()
    est = S.forest_bytes(states, precision)
    taken = None if emulated else free0 - _mem_free()
    pool.rebuild([], [])
    print(f"{precision}: ledger {used} B, estimate {est} B, hipMemGetInfo {taken} B")
    assert used <= est
    if taken is not None:
        assert taken <= est
