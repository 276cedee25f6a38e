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
