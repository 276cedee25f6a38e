"""The C-ABI library loads and exports every entry point include/afivo_amd.h
declares (no device compute: these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

from afivo_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "afivo_amd.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(afg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "afg_fas_fmg" in names and "afg_fas_vcycle" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_mg_init_defaults_and_errors():
    """afg_mg_init restates mg$D_init_mg (m_aX_mg.f90:130-154): defaults 2/2/2
    and a stop for missing indices (here: an error code and message)."""
    L = _lib.load()
    o = _lib.MgOpts(i_phi=1, i_rhs=2, i_tmp=3)
    assert L.afg_mg_init(ctypes.byref(o)) == 0
    assert (o.n_cycle_down, o.n_cycle_up, o.n_cycle_base, o.initialized) == (2, 2, 2, 1)
    bad = _lib.MgOpts(i_phi=1, i_rhs=2)
    assert L.afg_mg_init(ctypes.byref(bad)) != 0
    assert b"i_tmp not set" in L.afg_last_error(None)
    # the auto operators without eps are allowed (every box untagged, as
    # examples/poisson_cyl.f90 uses them); an invalid eps index is not
    ok = _lib.MgOpts(i_phi=1, i_rhs=2, i_tmp=3, i_eps=-1, op=_lib.AFG_OP_AUTO)
    assert L.afg_mg_init(ctypes.byref(ok)) == 0
    bad = _lib.MgOpts(i_phi=1, i_rhs=2, i_tmp=3, i_eps=0, op=_lib.AFG_OP_AUTO)
    assert L.afg_mg_init(ctypes.byref(bad)) != 0
    assert b"i_eps" in L.afg_last_error(None)


def test_ctx_create_argument_checks():
    """a$D_init's checks (m_aX_core.f90:61-63) come before any device call."""
    L = _lib.load()
    h = ctypes.c_void_p()
    assert L.afg_ctx_create(2, 7, 3, ctypes.byref(h)) == 1
    assert b"even" in L.afg_last_error(None)
    assert L.afg_ctx_create(4, 8, 3, ctypes.byref(h)) == 1
    assert L.afg_ctx_create(2, 8, 0, ctypes.byref(h)) == 1


def test_null_context_is_rejected():
    L = _lib.load()
    assert L.afg_sync(None) != 0
    assert L.afg_fas_fmg(None, None, 1, 0) != 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.AfgError):
        _lib.load()


def test_mg_opts_layout_matches_header(tmp_path):
    """The ctypes mirror of afg_mg_opts has the header's field offsets and
    size (the header compiled by gcc prints them)."""
    fields = [f[0] for f in _lib.MgOpts._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "afivo_amd.h"\nint main(void) {\n' +
                   "".join(f'    printf("%zu\\n", offsetof(afg_mg_opts, {n}));\n' for n in fields) +
                   '    printf("%zu\\n", sizeof(afg_mg_opts));\n    return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [getattr(_lib.MgOpts, n).offset for n in fields] + [ctypes.sizeof(_lib.MgOpts)]
    assert got == want


def test_mg_init_level_set_indices():
    """i_lsf and i_bval go together (m_aX_mg.f90:136-137); 0 reads as unset
    (the zero-filled spare words of callers that predate the fields)."""
    L = _lib.load()
    both = _lib.MgOpts(i_phi=1, i_rhs=2, i_tmp=3, op=_lib.AFG_OP_AUTO, i_lsf=5, i_bval=6)
    assert L.afg_mg_init(ctypes.byref(both)) == 0
    one = _lib.MgOpts(i_phi=1, i_rhs=2, i_tmp=3, op=_lib.AFG_OP_AUTO, i_lsf=5, i_bval=-1)
    assert L.afg_mg_init(ctypes.byref(one)) != 0
    assert b"both i_lsf and i_bval" in L.afg_last_error(None)
    zero = _lib.MgOpts(i_phi=1, i_rhs=2, i_tmp=3, i_lsf=0, i_bval=0)
    assert L.afg_mg_init(ctypes.byref(zero)) == 0
    assert (zero.i_lsf, zero.i_bval, zero.rstr) == (-1, -1, _lib.AFG_RSTR_LPL)
