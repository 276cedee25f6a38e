"""The allocation-free entry point afg_tree_extreme_noalloc, through the C
ABI: the same value as afg_tree_max_cc / afg_tree_min_cc bit for bit, in
device memory the caller owns, on two calls in a row (the second on a
smaller tree, whose reduction uses fewer of the scratch array's partials
than the first left behind); a variable without device data is refused.
Device test (the output buffer is device memory from hipMalloc)."""
import ctypes

import numpy as np
import pytest

from afivo_amd import _lib
from afivo_amd import mg as M
from afivo_amd import problems as P
from helpers import random_state

pytestmark = pytest.mark.gpu


class DeviceDouble:
    """One double of device memory, allocated by the test (the entry point
    under test allocates nothing)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(8)) == 0

    def value(self):
        out = ctypes.c_double()
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(ctypes.byref(out), self.ptr, ctypes.c_size_t(8), 2) == 0   # device to host
        return out.value

    def free(self):
        self.hip.hipFree(self.ptr)


def host_value(dt, iv, include_parents, want_max):
    out = ctypes.c_double()
    fn = _lib.load().afg_tree_max_cc if want_max else _lib.load().afg_tree_min_cc
    dt._c(fn(dt.h, iv, include_parents, ctypes.byref(out)))
    return out.value


def test_noalloc_extreme_equals_the_allocating_call():
    out = DeviceDouble()
    L = _lib.load()
    big = P.c2().state                 # 5875 boxes of 8^3 (about 9000 partials)
    small = P.small_2d().state         # 87 boxes of 8^2
    for st in (big, small):            # two calls in a row, the second smaller
        st.data = random_state(st, 3, seed=st.n_boxes)
        dt = M.DeviceTree.from_state(st)
        for iv in (1, 2):
            for include_parents in (0, 1):
                for want_max in (1, 0):
                    want = host_value(dt, iv, include_parents, want_max)
                    dt._c(L.afg_tree_extreme_noalloc(dt.h, iv, include_parents, want_max, out.ptr))
                    got = out.value()
                    assert np.float64(got).view(np.int64) == np.float64(want).view(np.int64), \
                        f"{st.n_boxes} boxes, var {iv}, parents {include_parents}, max {want_max}"
        dt.close()
    # a variable without device data would have to be allocated: refused
    st = P.small_2d().state
    dt = M.DeviceTree(st.dim, st.nc, 4, st.recs, st.dr_base)
    dt.upload(1, st.data[1])
    assert L.afg_tree_extreme_noalloc(dt.h, 4, 0, 1, out.ptr) == 2   # AFG_ERR_STATE
    assert "allocating" in L.afg_last_error(dt.h).decode()
    assert L.afg_tree_extreme_noalloc(dt.h, 1, 0, 1, None) == 1      # AFG_ERR_ARG: null output
    dt.close()
    out.free()
