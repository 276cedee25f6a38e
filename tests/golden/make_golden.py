"""Regenerate the golden fixtures of this directory from the reference.

Runs the reference solver (oracle/_ref/ref_harness_{2,3}d, built from the
reference sources by `make -C oracle ref`) on small problems and stores, per
case, the reference's own input dump (tree + initial cell data), output dump
(after FMG and V-cycles) and residual history.  Fixtures are data only.

Tree updates (the UPDATES cases): the solved tree of a golden case is
refined / derefined once more by the harness's refine2 step
(a$D_adjust_refinement, m_aX_core.f90:452-541, then the variables 1..2 of
every added box prolonged from its parent with a$D_prolong0/1/2).  Stored per
case (`<name>.npz`, no pickles): the new tree's records, the record indices
of the added boxes and the interiors of their prolonged variables.  The
script checks that the reference left every kept box's values unchanged and
that the solved tree equals the committed fixture of the base case.

Ghost cells and transfers (the XFERS cases): variable 4 of a solved golden
tree set from phi or rhs and changed by a$D_gc_tree (with a$D_gc_interp,
a$D_gc_interp_lim, a$D_gc_prolong0), a$D_restrict_tree or
a$D_prolong0/1/2_from, and face variable 1 (the gradient of phi) changed
by a$D_consistent_fluxes; stored per routine (`<name>.npz`): the packed
positions it changed and their values; and the second ghost layer of
a$D_gc2_box (a$D_gc2_prolong1, a$D_bc2_neumann_zero) of variable 4 after the
gc0 step, every box (`gc2`).

VTK output (the VTKS cases): oracle/_ref/vtk_ref_{2,3}d (the reference's XML
writer m_vtk, with the arrays assembled as a$D_write_vtk assembles them) on a
small tree with three cell variables and one face variable; stored per case
(`<name>.npz`): the tree and its cell data, the face variable, and the
written .vtu file's length, SHA-256 and XML header (the text before the raw
appended block).

Usage: python tests/golden/make_golden.py [case ...]   (default: all)
"""
import gzip
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

G2 = dict(n_gauss=2, ampl=[1.0, 1.0], sigma=[0.04, 0.04], r0=[0.25, 0.25, 0.75, 0.75])
G3 = dict(n_gauss=2, ampl=[1.0, -1.0], sigma=[0.1, 0.1], r0=[0.3, 0.3, 0.4, 0.7, 0.6, 0.5])

CASES = {
    # 2D, 4x4 boxes, FMG + 2 V-cycles
    "g2d_nc4": (2, dict(nc=4, l_uniform=2, l_max=5, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=2, **G2)),
    # 2D, 8x8 boxes, two FMG calls (the second with have_guess) + 1 V-cycle
    "g2d_nc8_guess": (2, dict(nc=8, l_uniform=2, l_max=4, thr=2e-3, n_base=6, n_fmg=2, n_vcycle=1,
                              n_down=1, n_up=3, **G2)),
    # 3D, 4^3 boxes, auto operators with a jump in eps at z = 0.5
    "g3d_nc4_eps": (3, dict(nc=4, l_uniform=2, l_max=3, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=1,
                            use_eps=True, eps_hi=5.0, **G3)),
    # 2D, 4x4 boxes, cylindrical coordinates (x the radius), auto operators:
    # mg2_box_clpl / gsrb_clpl, cylindrical restriction
    "g2d_nc4_cyl": (2, dict(nc=4, l_uniform=2, l_max=5, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=2, cyl=True, **G2)),
    # as above with eps = 1 below y = 0.5 and 5 above: mg2_box_clpld / gsrb_clpld
    "g2d_nc4_cyl_eps": (2, dict(nc=4, l_uniform=2, l_max=5, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=2, cyl=True,
                                use_eps=True, eps_hi=5.0, **G2)),
    # internal boundary on the circle |r - (0.45, 0.55)| = 0.23 (level set in
    # variable 5, boundary value 1 + x in variable 6): auto operators
    # (mg_lsf_box tags: gsrb/box/corr_lpllsf) with rstr_lpl ...
    "g2d_nc4_lsf": (2, dict(nc=4, l_uniform=2, l_max=5, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=2, lsf_mode=1,
                            lsf_c=[0.45, 0.55], lsf_rad=0.23, **G2)),
    # ... with mg2_auto_rstr (every tag set first) ...
    "g2d_nc4_lsf_rstr": (2, dict(nc=4, l_uniform=2, l_max=5, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=2, lsf_mode=1,
                                 rstr_auto=True, lsf_c=[0.45, 0.55], lsf_rad=0.23, **G2)),
    # ... and the lpllsf operator set for every box
    "g2d_nc4_lsf_set": (2, dict(nc=4, l_uniform=2, l_max=5, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=2, lsf_mode=2,
                                lsf_c=[0.45, 0.55], lsf_rad=0.23, **G2)),
    # 3D, the sphere |r - (0.45, 0.55, 0.5)| = 0.27: auto operators with
    # mg3_auto_rstr, and the lpllsf set
    "g3d_nc4_lsf_rstr": (3, dict(nc=4, l_uniform=2, l_max=3, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=1, lsf_mode=1,
                                 rstr_auto=True, lsf_c=[0.45, 0.55, 0.5], lsf_rad=0.27, **G3)),
    "g3d_nc4_lsf_set": (3, dict(nc=4, l_uniform=2, l_max=3, thr=2e-3, n_base=4, n_fmg=1, n_vcycle=1, lsf_mode=2,
                                lsf_c=[0.45, 0.55, 0.5], lsf_rad=0.27, **G3)),
    # 3D, 4^3 boxes, constant coefficient
    "g3d_nc4": (3, dict(nc=4, l_uniform=1, l_max=3, thr=5e-3, n_base=4, n_fmg=1, n_vcycle=1, **G3)),
    # 3D, 4^3 boxes, adaptive (42 boxes, 32 faces at refinement boundaries)
    "g3d_nc4_ad": (3, dict(nc=4, l_uniform=2, l_max=3, thr=0.5, n_base=4, n_fmg=1, n_vcycle=1, **G3)),
}


# name: (base case, refine2 namelist values); box counts as the reference makes them
UPDATES = {
    # 86 -> 118 boxes: 32 added, a$D_prolong1
    "update_g2d_nc4_p1": ("g2d_nc4", dict(refine2=1, refine2_lmax=6, refine2_thr=0.5)),
    # 86 -> 54 boxes: 32 removed
    "update_g2d_nc4_rm": ("g2d_nc4", dict(refine2=1, refine2_lmax=6, refine2_thr=2.0)),
    # 55 -> 87 boxes: 32 added, a$D_prolong0
    "update_g2d_nc8_guess_p0": ("g2d_nc8_guess", dict(refine2=3, refine2_lmax=5, refine2_thr=0.5)),
    # 74 -> 66 boxes: 16 removed, 8 added, a$D_prolong2
    "update_g3d_nc4_p2": ("g3d_nc4", dict(refine2=2, refine2_lmax=4, refine2_thr=2.0)),
    # 74 -> 130 boxes: 56 added, a$D_prolong1
    "update_g3d_nc4_p1": ("g3d_nc4", dict(refine2=1, refine2_lmax=4, refine2_thr=0.5)),
}
UPDATE_NVAR = 2
PROLONG_ORDER = {1: 1, 2: 2, 3: 0}   # refine2 value -> order of a$D_prolong<order>


def box_keys(st):
    return {(int(st.lvl[r]),) + tuple(int(x) for x in st.ix[r]): r for r in range(st.n_boxes)}


def make_update(name, base, extra):
    from afivo_amd.treefile import read_tree_file
    g = json.load(open(os.path.join(HERE, base + ".json")))
    dim = g["dim"]
    cfg = dict(g["cfg"], refine2_rm_thr=1e9, refine2_nvar=UPDATE_NVAR, **extra)
    work = tempfile.mkdtemp(prefix="afg_golden_")
    prefix, _ = oracle.run_ref(dim, cfg, workdir=work)
    old, new = read_tree_file(prefix + ".out"), read_tree_file(prefix + ".ref2")
    with gzip.open(os.path.join(HERE, base + ".out.gz")) as f:
        assert f.read() == open(prefix + ".out", "rb").read(), "solved tree differs from the base fixture"
    ko, kn = box_keys(old), box_keys(new)
    oo, on = old.offsets(), new.offsets()
    for key, r in kn.items():
        if key in ko:
            ro = ko[key]
            for iv in range(old.n_var):
                a = old.data[iv][oo[ro]:oo[ro + 1]]
                b = new.data[iv][on[r]:on[r + 1]]
                assert a.tobytes() == b.tobytes(), f"kept box {key} changed (var {iv + 1})"
    added = np.array(sorted(r for key, r in kn.items() if key not in ko), dtype=np.int32)
    inner = (slice(1, -1),) * dim
    vals = np.zeros((UPDATE_NVAR, len(added), new.nc ** dim))
    for iv in range(1, UPDATE_NVAR + 1):
        for k, r in enumerate(added):
            vals[iv - 1, k] = new.box_array(iv, int(r))[inner].ravel()
    meta = dict(base=base, dim=dim, cfg=cfg, prolong_order=PROLONG_ORDER[cfg["refine2"]],
                n_vars=UPDATE_NVAR, n_old=old.n_boxes, n_new=new.n_boxes,
                n_removed=len(set(ko) - set(kn)), n_added=len(added))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), recs=new.recs, dr_base=np.float64(new.dr_base),
                        added=added, vals=vals, meta=np.array(json.dumps(meta, sort_keys=True)))
    shutil.rmtree(work)
    print("wrote", name, {k: meta[k] for k in ("n_old", "n_new", "n_removed", "n_added")})


# Ghost cells and transfers (the XFERS cases): the solved tree of a golden
# case with variable 4 set from phi or rhs and changed by one routine of
# afivo (the harness's transfer_dumps).  Stored per routine: the packed
# positions of variable 4 that the routine changed and their values.
XFERS = {"xfer_g2d_nc4": "g2d_nc4", "xfer_g2d_nc8_guess": "g2d_nc8_guess", "xfer_g3d_nc4_ad": "g3d_nc4_ad",
         # cylindrical: the radius-weighted restriction and y-face fluxes
         "xfer_g2d_nc4_cyl": "g2d_nc4_cyl"}
XFER_OPS = {"gci": 1, "gcl": 1, "gc0": 1, "rst": 2, "pr1": 2, "pr0": 2, "pr2": 2}   # op: source variable


def make_xfer(name, base):
    from afivo_amd.treefile import read_tree_file
    g = json.load(open(os.path.join(HERE, base + ".json")))
    cfg = dict(g["cfg"], xfer=True)
    work = tempfile.mkdtemp(prefix="afg_golden_")
    prefix, _ = oracle.run_ref(g["dim"], cfg, workdir=work)
    with gzip.open(os.path.join(HERE, base + ".out.gz")) as f:
        assert f.read() == open(prefix + ".out", "rb").read(), "solved tree differs from the base fixture"
    out = read_tree_file(prefix + ".out")
    arrays = {}
    for op, src in XFER_OPS.items():
        st = read_tree_file(prefix + "." + op)
        assert np.array_equal(st.recs, out.recs)
        for iv in range(3):
            assert st.data[iv].tobytes() == out.data[iv].tobytes(), f"{op} changed variable {iv + 1}"
        before = out.data[src - 1].view(np.int64)
        after = st.data[3]
        pos = np.nonzero(after.view(np.int64) != before)[0].astype(np.int32)
        arrays[op + "_pos"] = pos
        arrays[op + "_val"] = after[pos]
        print(name, op, "changed", pos.size, "of", after.size)
    # face variable 1: the gradient of phi, then a$D_consistent_fluxes
    sys.path.insert(0, os.path.dirname(HERE))
    from helpers import face_gradient
    before = face_gradient(out, out.data[0])
    after = np.fromfile(prefix + ".flx", dtype="<f8")
    assert after.size == before.size
    pos = np.nonzero(after.view(np.int64) != before.view(np.int64))[0].astype(np.int32)
    arrays["flx_pos"] = pos
    arrays["flx_val"] = after[pos]
    print(name, "flx changed", pos.size, "of", after.size)
    # a$D_gc2_box of variable 4 after the gc0 step, every box (record order)
    arrays["gc2"] = np.fromfile(prefix + ".gc2", dtype="<f8")
    print(name, "gc2", arrays["gc2"].size, "values")
    meta = dict(base=base, dim=g["dim"], cfg=cfg, ops=XFER_OPS)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), meta=np.array(json.dumps(meta, sort_keys=True)),
                        **arrays)
    shutil.rmtree(work)


# name -> (dim, r0 of the domain [r0, r0 + 1]^dim, cylindrical)
VTKS = {
    "vtk_g2d": (2, 0.25, False),
    "vtk_g2d_cyl": (2, 0.0, True),
    "vtk_g3d": (3, -0.5, False),
}


def make_vtk(name, dim, r0, cyl):
    import hashlib
    import subprocess
    from afivo_amd.treefile import read_tree_file
    work = tempfile.mkdtemp(prefix="afg_golden_")
    out = os.path.join(work, name)
    exe = os.path.join(oracle.REF_DIR, f"vtk_ref_{dim}d")
    subprocess.run([exe, out, repr(r0)] + (["cyl"] if cyl else []), check=True, cwd=work)
    st = read_tree_file(out + ".tree")
    fc = np.fromfile(out + ".fc", dtype="<f8")
    vtu = open(out + ".vtu", "rb").read()
    head = vtu[:vtu.index(b"<AppendedData")].decode()
    meta = dict(dim=dim, nc=st.nc, min_lvl=st.min_lvl, highest_lvl=st.highest_lvl, dr_base=st.dr_base,
                r0=r0, cyl=cyl, n_cycle=7, time=0.125, cc_names=["phi", "rhs", "tmp"], fc_names=["flux"],
                ixs_fc=[1], vtu_len=len(vtu), vtu_sha256=hashlib.sha256(vtu).hexdigest(), vtu_head=head)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), meta=np.array(json.dumps(meta, sort_keys=True)),
                        recs=st.recs, data=np.stack(st.data), fc=fc)
    print(name, "vtu", len(vtu), "bytes")
    shutil.rmtree(work)


def main(names):
    for name, (dim, cfg) in CASES.items():
        if names and name not in names:
            continue
        work = tempfile.mkdtemp(prefix="afg_golden_")
        cfg = dict(cfg, ref_mode="gauss", out_file=name)
        prefix, _ = oracle.run_ref(dim, cfg, workdir=work)
        for ext in (".in", ".out"):
            with open(prefix + ext, "rb") as f, gzip.open(os.path.join(HERE, name + ext + ".gz"), "wb", 9) as g:
                g.write(f.read())
        shutil.copy(prefix + ".res", os.path.join(HERE, name + ".res"))
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump({"dim": dim, "cfg": cfg}, f, indent=1, sort_keys=True)
        shutil.rmtree(work)
        print("wrote", name)
    for name, (base, extra) in UPDATES.items():
        if not names or name in names:
            make_update(name, base, extra)
    for name, base in XFERS.items():
        if not names or name in names:
            make_xfer(name, base)
    for name, args in VTKS.items():
        if not names or name in names:
            make_vtk(name, *args)


if __name__ == "__main__":
    main(sys.argv[1:])
