"""Regenerate the golden fixtures of this directory from the reference.

Runs the reference solver (oracle/_ref/ref_harness_{2,3}d, built from the
reference sources by `make -C oracle ref`) on small problems and stores, per
case, the reference's own input dump (tree + initial cell data), output dump
(after FMG and V-cycles) and residual history.  Fixtures are data only.

Usage: python tests/golden/make_golden.py
"""
import gzip
import json
import os
import shutil
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

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
    # 3D, 4^3 boxes, constant coefficient
    "g3d_nc4": (3, dict(nc=4, l_uniform=1, l_max=3, thr=5e-3, n_base=4, n_fmg=1, n_vcycle=1, **G3)),
}


def main():
    for name, (dim, cfg) in CASES.items():
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


if __name__ == "__main__":
    main()
