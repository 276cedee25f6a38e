"""Serving loop (continuous batching of Poisson-solve requests).  Device
tests; also run through the CPU emulation (test_emul.py)."""
import numpy as np
import pytest

from afivo_amd import serving as S

pytestmark = pytest.mark.gpu


def small_trace(n=8, rate=2000.0, seed=4):
    return S.make_trace(seed, n, rate, cells=(1e3, 2e4), tols=(1e-8, 1e-10))


def check_mix_against_solo(mix, solo):
    """Every request meets its tolerance in the mix and alone; a request that
    took as many steps in the mix as alone has bit-identical phi, the others
    agree within their tolerance.  Returns the number of identical ones."""
    same = 0
    for a, b in zip(sorted(mix, key=lambda r: r.rid), sorted(solo, key=lambda r: r.rid)):
        assert a.finish >= a.start >= a.arrival >= 0
        assert a.residual <= a.tol and b.residual <= b.tol
        if a.cycles == b.cycles:
            assert np.array_equal(a.phi.view(np.int64), b.phi.view(np.int64)), f"request {a.rid}"
            same += 1
        else:
            scale = np.abs(b.phi).max()
            assert np.abs(a.phi - b.phi).max() <= 100 * a.tol * scale, f"request {a.rid}"
    return same


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_serving_meets_tolerances_and_matches_solo_runs(precision):
    """The same requests served in forests of up to 3 trees and solved one at
    a time through the same steps (mixed-precision corrections, or FP64
    V-cycles) give the same phi."""
    mix = S.serve(small_trace(), max_lanes=2, max_trees=3, n_base=20, precision=precision)
    solo = S.one_at_a_time(small_trace(), n_base=20, precision=precision)
    assert len({(r.state.dim) for r in mix}) == 2, "the trace should mix 2D and 3D"
    assert check_mix_against_solo(mix, solo) > 0


def test_mixed_serving_matches_fp64_solver():
    """Mixed-precision lanes reach each tolerance with the FP64 solver's
    solution (within 100 x tolerance of max|phi|)."""
    mix = S.serve(small_trace(), max_lanes=2, max_trees=3, n_base=20)
    ref = S.one_at_a_time(small_trace(), n_base=20, precision="fp64")
    for a, b in zip(sorted(mix, key=lambda r: r.rid), sorted(ref, key=lambda r: r.rid)):
        assert a.residual <= a.tol
        scale = np.abs(b.phi).max()
        assert np.abs(a.phi - b.phi).max() <= 100 * a.tol * scale, f"request {a.rid}"


def test_serving_admits_by_memory_budget():
    """With a budget that holds one forest of one tree at a time, requests
    are still all served (one lane at a time)."""
    tr = small_trace(n=4)
    budget = max(S.forest_bytes([r]) for r in tr)
    out = S.serve(tr, hbm_budget=budget, max_lanes=4, max_trees=4, n_base=20)
    assert all(r.residual <= r.tol for r in out)
    # no two lanes overlapped in time
    spans = sorted((r.start, r.finish) for r in out)
    for (s0, f0), (s1, f1) in zip(spans, spans[1:]):
        assert s1 >= f0 or (s1, f1) == (s0, f0)
