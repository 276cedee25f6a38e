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
    test_mixed_serving_matches_fp64_solver,
    test_serving_admits_by_memory_budget,
    test_serving_meets_tolerances_and_matches_solo_runs,
)
from test_mixed import (  # noqa: E402,F401
    test_fp32_multigrid_tracks_fp64,
    test_mixed_solve_matches_fp64_solution,
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
    test_consistent_fluxes_match_oracle_full_size,
    test_consistent_fluxes_match_reference,
)
