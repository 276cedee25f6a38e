"""The Fortran drop-in (afivo_amd/fortran/m_aX_mg.f90).

The reference's own harness program (oracle/ref_harness_X.f90, which uses
afivo's public multigrid API) is compiled unchanged against the drop-in
m_a2_mg / m_a3_mg and linked with libafivo_amd (tests/fortran/Makefile).
Run with the same configuration as the reference build, it must write the
same tree, the same cell data bit for bit, and the same residual history.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from afivo_amd import _lib
from afivo_amd.treefile import read_tree_file
from helpers import assert_bitexact

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "fortran", "_build")
G2 = dict(n_gauss=2, ampl=[1.0, 1.0], sigma=[0.04, 0.04], r0=[0.25, 0.25, 0.75, 0.75])
G3 = dict(n_gauss=2, ampl=[1.0, -1.0], sigma=[0.1, 0.1], r0=[0.3, 0.3, 0.4, 0.7, 0.6, 0.5])


def exe(dim):
    return os.path.join(BUILD, f"gpu_harness_{dim}d")


def have_gpu():
    h = ctypes.c_void_p()
    L = _lib.load()
    if L.afg_ctx_create(2, 8, 1, ctypes.byref(h)) == 0:
        L.afg_ctx_destroy(h)
        return True
    return False


def run_dropin(dim, cfg, workdir):
    """Run the drop-in harness; returns (returncode, output, prefix)."""
    prefix = os.path.join(workdir, "gpu")
    cfg = dict(cfg, out_file=prefix)
    nml = os.path.join(workdir, "gpu.nml")
    with open(nml, "w") as f:
        f.write("&cfg\n" + ",\n".join(f" {k}={O._fmt(v)}" for k, v in cfg.items()) + "\n/\n")
    p = subprocess.run([exe(dim), nml], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    return p.returncode, p.stdout + p.stderr, prefix


needs_build = pytest.mark.skipif(not (os.path.exists(exe(2)) and O.ref_available(2)),
                                 reason="drop-in harness or reference build absent")


@needs_build
def test_dropin_fails_loudly_without_gpu(tmp_path):
    """No CPU fallback: without a device the drop-in stops with the
    library's error message."""
    if have_gpu():
        pytest.skip("a GPU is present")
    rc, out, _ = run_dropin(2, dict(nc=8, ref_mode="gauss", l_uniform=2, l_max=3, thr=1e-3,
                                    n_base=2, n_vcycle=0, **G2), str(tmp_path))
    assert rc != 0
    assert "afivo_amd error" in out


CASES = {
    "2d_c1": (2, dict(nc=8, ref_mode="gauss", l_uniform=5, l_max=7, thr=1e-4, n_base=50,
                      n_fmg=1, n_vcycle=4, **G2)),
    "2d_guess": (2, dict(nc=4, ref_mode="gauss", l_uniform=3, l_max=6, thr=2e-3, n_base=6,
                         n_fmg=3, n_vcycle=1, n_down=1, n_up=3, **G2)),
    "3d": (3, dict(nc=8, ref_mode="gauss", l_uniform=2, l_max=4, thr=1e-3, n_base=4,
                   n_fmg=1, n_vcycle=2, **G3)),
    "3d_eps": (3, dict(nc=4, ref_mode="gauss", l_uniform=2, l_max=4, thr=1e-3, n_base=4,
                       n_fmg=2, n_vcycle=1, use_eps=True, eps_hi=6.0, **G3)),
    # cylindrical trees with the auto operators (mg2_box_clpl, and clpld on eps boxes)
    "2d_cyl": (2, dict(nc=8, ref_mode="gauss", l_uniform=4, l_max=7, thr=1e-4, n_base=50,
                       n_fmg=1, n_vcycle=3, cyl=True, **G2)),
    "2d_cyl_eps": (2, dict(nc=8, ref_mode="gauss", l_uniform=4, l_max=6, thr=1e-4, n_base=50,
                           n_fmg=1, n_vcycle=3, cyl=True, use_eps=True, eps_hi=4.0, **G2)),
    # internal boundary on a circle / sphere (level set and boundary value):
    # the auto operators with mg2_auto_rstr (tags set by mg2_set_box_tag
    # first) and with rstr_lpl, and the lpllsf set
    "2d_lsf_rstr": (2, dict(nc=8, ref_mode="gauss", l_uniform=4, l_max=7, thr=1e-4, n_base=50,
                            n_fmg=1, n_vcycle=3, lsf_mode=1, rstr_auto=True, lsf_c=[0.45, 0.55],
                            lsf_rad=0.23, **G2)),
    "2d_lsf": (2, dict(nc=8, ref_mode="gauss", l_uniform=4, l_max=6, thr=1e-4, n_base=50,
                       n_fmg=2, n_vcycle=1, lsf_mode=1, lsf_c=[0.45, 0.55], lsf_rad=0.23, **G2)),
    "3d_lsf_set": (3, dict(nc=8, ref_mode="gauss", l_uniform=2, l_max=4, thr=1e-3, n_base=4,
                           n_fmg=1, n_vcycle=2, lsf_mode=2, lsf_c=[0.45, 0.55, 0.5], lsf_rad=0.27, **G3)),
}


@pytest.mark.gpu
@needs_build
@pytest.mark.parametrize("case", sorted(CASES))
def test_dropin_matches_reference(tmp_path, case):
    dim, cfg = CASES[case]
    ref_dir = tmp_path / "ref"
    gpu_dir = tmp_path / "gpu"
    ref_dir.mkdir()
    gpu_dir.mkdir()
    prefix_ref, _ = O.run_ref(dim, dict(cfg, out_file="ref"), workdir=str(ref_dir))
    rc, out, prefix_gpu = run_dropin(dim, cfg, str(gpu_dir))
    assert rc == 0, out
    a, b = read_tree_file(prefix_gpu + ".out"), read_tree_file(prefix_ref + ".out")
    assert np.array_equal(a.recs, b.recs)
    for iv in range(1, b.n_var + 1):
        assert_bitexact(a.data[iv - 1], b.data[iv - 1], f"{case} var {iv}")
    assert np.array_equal(O.read_res(prefix_gpu), O.read_res(prefix_ref))


MIXED = os.path.join(BUILD, "mixed_2d")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(MIXED), reason="drop-in test program absent")
def test_dropin_extensions_autotune_and_mixed_precision():
    """The drop-in's extensions from Fortran (tests/fortran/mixed_2d.f90):
    mg2_autotune, then mg2_mixed_solve to 1e-10, against the module's plain
    FP64 cycles on a copy of the tree."""
    p = subprocess.run([MIXED], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert p.returncode == 0, p.stdout + p.stderr
    w = p.stdout.split()
    vals = dict(zip(w[0::2], w[1::2]))
    assert int(vals["corrections"]) >= 2
    assert float(vals["rel_res"]) <= 1e-10
    assert int(vals["fp64_vcycles"]) >= 2
    assert float(vals["phi_diff"]) <= 1e-9


UPDATE = os.path.join(BUILD, "update_2d")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(UPDATE), reason="drop-in test program absent")
def test_dropin_device_update_matches_host_refinement():
    """mg2_device_update from Fortran (tests/fortran/update_2d.f90): one
    refinement step of a device-resident solve gives, bit for bit, what
    afivo's own step on the host gives (kept boxes unchanged, added boxes
    a$D_prolong1 of phi), and the solve continues on the new tree."""
    p = subprocess.run([UPDATE], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert p.returncode == 0, p.stdout + p.stderr
    w = p.stdout.split()
    vals = dict(zip(w[0::2], w[1::2]))
    assert int(vals["added"]) > 0 and int(vals["after"]) != int(vals["before"])
    assert int(vals["kept_bad"]) == 0
    assert int(vals["new_bad"]) == 0
    assert float(vals["rel_res"]) < 1e-3


@pytest.mark.gpu
@needs_build
def test_dropin_auto_rstr_without_tags_stops_like_reference(tmp_path):
    """mg2_auto_rstr with no tag set: the reference stops in the FMG's first
    restriction (m_aX_mg.f90:639), and so does the drop-in, with the same
    message; neither writes a solved tree."""
    cfg = dict(nc=4, ref_mode="gauss", l_uniform=2, l_max=4, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=1,
               lsf_mode=1, rstr_auto=True, tag_boxes=False, lsf_c=[0.45, 0.55], lsf_rad=0.23, **G2)
    msg = "mg2_auto_rstr: box_c tag not set"
    nml = tmp_path / "ref.nml"
    nml.write_text("&cfg\n" + ",\n".join(f" {k}={O._fmt(v)}" for k, v in
                                         dict(cfg, out_file=str(tmp_path / "ref")).items()) + "\n/\n")
    p = subprocess.run([os.path.join(O.REF_DIR, "ref_harness_2d"), str(nml)], capture_output=True,
                       text=True, timeout=600, env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert msg in p.stdout + p.stderr
    assert not (tmp_path / "ref.out").exists()
    gpu_dir = tmp_path / "gpu"
    gpu_dir.mkdir()
    rc, out, prefix = run_dropin(2, cfg, str(gpu_dir))
    assert msg in out
    assert not os.path.exists(prefix + ".out")


@pytest.mark.gpu
@needs_build
def test_dropin_end_to_end_on_the_3d_benchmark_tree(tmp_path):
    """The reference's harness program through the drop-in on the 3D
    benchmark tree (5875 boxes of 8^3, FMG + 8 V-cycles, 50 base sweeps),
    in the default mode (every mg3_fas_fmg / mg3_fas_vcycle call uploads phi,
    rhs and tmp from box%cc and downloads them) and device-resident
    (mg%device_resident, one mg3_to_host after each solve): both write the
    reference build's tree bit for bit, and the default mode its residual
    history (the harness reads it from box%cc after each call, which the
    resident mode leaves to mg3_to_host); the solve times
    (5 timed solves each) give the share of the host copies."""
    from afivo_amd import problems as P
    cfg = P.harness_cfg(P.c2(), n_time=5, dump_input=False)
    prefix_ref, _ = O.run_ref(3, dict(cfg, out_file="ref", n_time=0), workdir=str(tmp_path), threads=8)
    ref = read_tree_file(prefix_ref + ".out")
    times = {}
    for resident in (False, True):
        d = tmp_path / ("resident" if resident else "default")
        d.mkdir()
        rc, out, prefix = run_dropin(3, dict(cfg, resident=resident), str(d))
        assert rc == 0, out
        got = read_tree_file(prefix + ".out")
        assert np.array_equal(got.recs, ref.recs)
        for iv in range(1, ref.n_var + 1):
            assert_bitexact(got.data[iv - 1], ref.data[iv - 1], f"resident={resident} var {iv}")
        if not resident:   # the harness reads its residual history from box%cc
            assert np.array_equal(O.read_res(prefix), O.read_res(prefix_ref))
        times[resident] = float(np.median(O.read_time(prefix)[1]))
    share = 1 - times[True] / times[False]
    print(f"drop-in, 3D benchmark tree: {1e3 * times[False]:.2f} ms per solve with host copies on every call, "
          f"{1e3 * times[True]:.2f} ms device-resident; host copies {100 * share:.0f}% of the default mode")
    assert times[True] < times[False]


BOX_OPS = {d: (os.path.join(BUILD, f"box_ops_{d}d"), os.path.join(O.REF_DIR, f"ref_box_ops_{d}d")) for d in (2, 3)}


@pytest.mark.skipif(not all(os.path.exists(p) for pair in BOX_OPS.values() for p in pair),
                    reason="box procedure programs absent (make -C oracle ref; make -C tests/fortran)")
@pytest.mark.parametrize("dim,coord", [(2, ""), (2, "cyl"), (3, "")])
def test_dropin_box_procedures_match_reference(tmp_path, dim, coord):
    """The box procedures called directly (oracle/box_ops_X.f90): every
    mg$D_box_* family, mg$D_sides_rb and the auto operators on every box of
    a tree with normal, constant-eps, varying-eps and level-set boxes (and
    mg2_box_clpl / clpld on the cylindrical tree), written out after each
    procedure.  The drop-in's build writes the same bytes as the build
    against the reference's m_a$D_mg.  The procedures are host code (the
    library's afg_box_* functions): no device is used."""
    outs = []
    for prog in BOX_OPS[dim]:
        out = str(tmp_path / (os.path.basename(prog) + ".bin"))
        p = subprocess.run([prog, out] + ([coord] if coord else []), capture_output=True, text=True,
                           timeout=120, env=dict(os.environ, OMP_NUM_THREADS="1"))
        assert p.returncode == 0, p.stdout + p.stderr
        outs.append(open(out, "rb").read())
    new, ref = outs
    n = int(np.frombuffer(ref[:4], dtype=np.int32)[0])
    lsf_tags = np.frombuffer(ref[-4 * n:], dtype=np.int32)
    assert n > 10 and set(lsf_tags) == {1, 2}, "the tree should hold level-set and normal boxes"
    assert len(new) == len(ref)
    a, b = np.frombuffer(new[4:-4 * n], dtype=np.uint8), np.frombuffer(ref[4:-4 * n], dtype=np.uint8)
    assert np.array_equal(a, b), f"first differing byte at {int(np.nonzero(a != b)[0][0]) + 4}"
    assert new[-4 * n:] == ref[-4 * n:]
