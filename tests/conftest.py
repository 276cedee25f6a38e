import ctypes
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

EMUL_LIB = os.path.join(ROOT, "tests", "emul", "_build", "libafivo_emul.so")


def pytest_addoption(parser):
    parser.addoption("--emul", action="store_true",
                     help="run the tests through the CPU emulation build of the library "
                          "(tests/emul) instead of the HIP library")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs through the C ABI)")
    if config.getoption("--emul"):
        load_emulation()


def load_emulation():
    """Point the binding at the CPU emulation library (test process only)."""
    from afivo_amd import _lib
    if not os.path.exists(EMUL_LIB):
        raise RuntimeError(f"{EMUL_LIB} missing: make -C tests/emul")
    L = ctypes.CDLL(EMUL_LIB)
    for name, (res, args) in _lib.SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    previous = _lib._lib
    _lib._lib = L
    return previous
