"""The GPU parity suite run through the CPU emulation build (tests/emul).

The emulation compiles the library's host code and the kernels' per-thread
bodies with g++ and runs every stage's threads one after the other (a stage
is race-free, so the order does not matter), with guard pages after every
device allocation.  It checks the stage logic, indexing and orchestration on
the CPU, bit for bit against the oracle; the same tests run on the GPU under
the gpu marker.  (Test infrastructure: the product never loads it.)
"""
import os

import pytest

import conftest
from afivo_amd import _lib

pytestmark = pytest.mark.skipif(not os.path.exists(conftest.EMUL_LIB),
                                reason="emulation library not built (make -C tests/emul)")


@pytest.fixture(autouse=True)
def emulated_library():
    previous = conftest.load_emulation()
    yield
    _lib._lib = previous


from test_gpu_parity import (  # noqa: E402,F401
    test_boundary_values_and_neumann,
    test_bottom_kernel_limit_is_exact,
    test_persistent_kernel_is_exact,
    test_split_passes_are_exact,
    test_c1_full_config_bitexact,
    test_c2_full_config_bitexact,
    test_cylindrical_solve_bitexact,
    test_golden_solve_bitexact,
    test_level_set_and_eps_tag_conflict,
    test_level_set_solve_bitexact,
    test_graphs_and_direct_launches_agree,
    test_periodic_base,
    test_rhs_changes_between_solves,
    test_solve_bitexact,
    test_stage_bitexact,
    test_tile_shapes_are_exact,
    test_pass_pairs_are_exact,
    test_autotune_keeps_state,
    test_timing_mode_is_exact,
)
from test_forest_gpu import (  # noqa: E402,F401
    test_forest_equals_single_solves,
    test_forest_mixed_precision_equals_single_solves,
    test_forest_partial_vcycle_and_oracle,
    test_forest_rejects_links_across_trees,
    test_mixed_correct_needs_begin,
)
from test_serving import (  # noqa: E402,F401
    test_forest_bytes_bounds_the_allocations,
    test_served_phi_equals_solo_runs,
    test_serving_admits_by_memory_budget,
)
from test_mixed import (  # noqa: E402,F401
    test_fp32_multigrid_tracks_fp64,
    test_mixed_solve_matches_fp64_solution,
    test_mixed_solve_with_fp64_tau,
    test_mixed_correction_rhs_paths_agree,
)
from test_tree_update import (  # noqa: E402,F401
    test_update_full_size_properties,
    test_update_matches_boxes_not_ids,
    test_update_matches_reference,
)
from test_transfers_gpu import (  # noqa: E402,F401
    test_gc_rejects_bad_arguments,
    test_transfers_match_oracle_full_size,
    test_transfers_match_reference,
    test_tree_arithmetic,
)
from test_transfers_gpu import (  # noqa: E402,F401
    test_box_array_transfers,
    test_second_ghost_layer_matches_oracle_full_size,
    test_second_ghost_layer_matches_reference,
    test_consistent_fluxes_match_oracle_full_size,
    test_consistent_fluxes_match_reference,
)


def test_failed_launch_is_an_error():
    """A kernel launch that fails (emulated: the next kernels do not run and
    hipGetLastError reports the failure, as for a launch with a bad
    configuration on the device) makes the entry point return AFG_ERR_HIP,
    naming the entry point, instead of status 0 with stale data; the next
    call works again."""
    import ctypes

    import numpy as np

    from afivo_amd import mg as M
    from afivo_amd import problems as P
    from afivo_amd._lib import AfgError

    hook = ctypes.CDLL(conftest.EMUL_LIB).emul_fail_launches
    hook.argtypes = [ctypes.c_int]
    p = P.small_2d()
    dt = M.DeviceTree.from_state(p.state)
    dt.set_graphs(False)
    mg = M.default_mg(n_base=p.n_base)
    M.mg_fas_fmg(dt, mg, True, False)
    for call, what in ((lambda: M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True), "afg_fas_vcycle"),
                       (lambda: dt.restrict_tree(1), "afg_restrict_tree"),
                       (lambda: M.max_abs_cc(dt, 3), "tree extreme")):
        hook(1)
        with pytest.raises(AfgError, match=what + ".*launch failed"):
            call()
    M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
    assert np.isfinite(M.max_abs_cc(dt, 3))
    dt.close()


def brute_morton(ix0, dim):
    """Bit interleave of ix - 1 done bit by bit (bit b of dimension d goes to
    bit b*dim + d): morton_from_ix2/3 (m_morton.f90:33-59) on ix - 1, the key
    of a$D_tidy_up (m_aX_core.f90:272-279)."""
    key = 0
    for b in range(21):
        for d in range(dim):
            key |= ((int(ix0[d]) >> b) & 1) << (b * dim + d)
    return key


@pytest.mark.parametrize("dim", [2, 3])
def test_slot_order_is_morton(dim):
    """Every level's slots are in ascending Morton order of ix - 1 (brute-force
    interleave), on the benchmark tree of the dimension; in a forest they are
    grouped by tree (higher highest level first, then tree index) and in
    Morton order within a tree.  The interleave itself is pinned on values
    whose key is known: ix - 1 = (1, 0[, 0]) -> 1, (0, 1[, 0]) -> 2,
    (0, 0, 1) -> 4, (3, 3) -> 15, all ones -> 2^(21 dim) - 1."""
    import numpy as np

    from afivo_amd import forest as F
    from afivo_amd import mg as M
    from afivo_amd import problems as P

    assert brute_morton((1, 0, 0), dim) == 1 and brute_morton((0, 1, 0), dim) == 2
    assert brute_morton((3, 3, 0), 2) == 15 and brute_morton((0, 0, 1), 3) == 4
    assert brute_morton(((1 << 21) - 1,) * 3, dim) == (1 << (21 * dim)) - 1
    st = (P.c1() if dim == 2 else P.c2()).state
    dt = M.DeviceTree(st.dim, st.nc, 1, st.recs, st.dr_base)
    row = {int(i): r for r, i in enumerate(st.ids)}
    for lvl in range(dt.min_lvl, dt.highest_lvl + 1):
        ids = dt.level_ids(lvl)
        assert ids.size == int(np.sum(st.lvl == lvl))
        keys = [brute_morton(st.ix[row[int(i)]] - 1, dim) for i in ids]
        assert all(a < b for a, b in zip(keys, keys[1:])), f"level {lvl} not in Morton order"
    dt.close()
    if dim == 2:
        states = [F.c3_tree(5, k)[0] for k in range(6)]
        states = [s for s in states if s.nc == states[0].nc][:4]
        comb, tree_of = F.combine(states)
        dt = M.DeviceTree(comb.dim, comb.nc, 1, comb.recs, comb.dr_base, tree_of)
        hi = [int(s.lvl.max()) for s in states]
        row = {int(i): r for r, i in enumerate(comb.ids)}
        for lvl in range(dt.min_lvl, dt.highest_lvl + 1):
            ids = dt.level_ids(lvl)
            order = [(-hi[tree_of[row[int(i)]]], int(tree_of[row[int(i)]]),
                      brute_morton(comb.ix[row[int(i)]] - 1, 2)) for i in ids]
            assert order == sorted(order) and len(set(order)) == len(order)
        dt.close()
