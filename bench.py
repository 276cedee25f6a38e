"""Benchmark of the FAS multigrid hot path on one MI355X.

Workload: the 3D FP64 Poisson configuration -- unit cube, 8^3-cell boxes,
uniform refinement to level 4 and adaptive refinement to level 7 around four
Gaussians (sigma 0.03, centres uniform in [0.2, 0.8]^3 from seed 1), threshold
dr^2 max|rhs| > 1e-3, giving 2.63e6 leaf cells (3.0e6 cells on all levels);
Dirichlet zero, 7-point Laplacian.  One step is one whole solve: FMG followed
by 8 V-cycles with 2+2 GSRB sweeps and 50 sweeps on the coarsest level, the
residual stored after every cycle (as the reference examples do).

`value` is cell-updates/s over the timed steps (a cell-update is one cell
relaxed by one full red+black sweep; FMG's cycles count), with the tree and
its data resident in device memory before timing starts.  The roofline object
is for the dominant HBM kernel, the GSRB colour pass on the finest level,
timed with HIP events on the context's stream during the timed steps; its
traffic comes from the committed rocprofv3 counter summary.  The CPU baseline runs the
reference's own solver (built from its sources into oracle/_ref) on the same
tree on this host's cores.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="0: every core this process may use, at most OMP_NUM_THREADS when that is set")
    ap.add_argument("--no-timing", action="store_true",
                    help="skip the per-kernel event timing (roofline is then null)")
    ap.add_argument("--workload", choices=["c1", "c2", "c3", "c4", "c5", "update", "transfers"], default="c2",
                    help="c1: the 2D FP64 tree (poisson_basic_2d form); c2: the 3D FP64 tree (default); c3: batched 2D FP64 forest of --trees "
                         "trees; c4: 3D variable-eps tree, mixed precision against FP64; "
                         "c5: served request trace; update: refinement steps of the c2 tree "
                         "on the device; transfers: ghost cells, restriction, prolongation, "
                         "fluxes and arithmetic on the c2 tree")
    ap.add_argument("--requests", type=int, default=64, help="c5: number of requests")
    ap.add_argument("--rate", type=float, default=100.0, help="c5: arrival rate (requests/s)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed",
                    help="c5: arithmetic of the served forests")
    ap.add_argument("--big-cells", type=float, default=0,
                    help="c5: trees of at least this many leaf cells get forests of their own (0: no split)")
    ap.add_argument("--c4-solver", choices=["both", "fp64", "mixed"], default="both",
                    help="c4: solvers to run (one alone for profiling)")
    ap.add_argument("--mixed-cycles", type=int, default=0,
                    help="c4: FP32 V-cycles per correction (0: try 1 to 5)")
    ap.add_argument("--trees", type=int, default=256, help="c3: number of trees")
    ap.add_argument("--tol", type=float, default=1e-10, help="c3, c4: relative residual to reach")
    ap.add_argument("--tiles", default="",
                    help="c1, c2: tile shape per level from the lowest, comma-separated (with --pairs and "
                         "--bottom-cells: this configuration instead of the autotuner's)")
    ap.add_argument("--pairs", default="", help="c1, c2: fused pass pairs per level (0/1), comma-separated")
    ap.add_argument("--bottom-cells", type=int, default=-1, help="c1, c2: the bottom kernel's cell limit")
    ap.add_argument("--persist", default="",
                    help="c1, c2, c3: the persistent kernel as WORKGROUPS,CELLS (default: off)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy rhs to tmp over every box once, from cold caches (for the "
                         "HBM byte counters' scale factor), print its byte count and exit")
    ap.add_argument("--no-graphs", action="store_true",
                    help="c4: launch the V-cycles' kernels directly instead of replaying HIP graphs "
                         "(for kernel traces; the kernels are the same)")
    ap.add_argument("--copy-ceiling", action="store_true",
                    help="only time plain copies at the sizes of the colour passes (for the kernel trace): "
                         "rhs to tmp over every box and flat copies of the finest level's byte count, "
                         "each from cold caches and then three times warm; print the byte counts and exit")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_gsrb_c2_level7_traffic.md")


def measured_traffic(instances):
    """HBM bytes per launch of the finest-level colour passes: the committed
    rocprofv3 counter summary's figure (FETCH_SIZE and WRITE_SIZE, scaled by
    the factors a plain copy of known size gives) for each kernel instance
    that ran, `traffic[<instance>] = <bytes>`, averaged with the instances'
    launch counts; None when an instance has no figure there."""
    import re
    try:
        text = open(TRAFFIC_FILE).read()
    except OSError:
        return None
    table = {m.group(1).strip(): float(m.group(2))
             for m in re.finditer(r"traffic\[([^\]]+)\]\s*=\s*([0-9.eE+]+)", text)}
    tot = n = 0
    for name, count in instances.items():
        if name not in table:
            return None
        tot += table[name] * count
        n += count
    return tot / n if n else None


def copy_ceiling(dt, st):
    """Plain copies of the colour passes' sizes, each first from cold caches
    (after writing a 2 GiB buffer) and then three times warm: the library's
    box copy (rhs to tmp over every box, one launch per level) and the HIP
    runtime's device-to-device copy of flat buffers holding the finest
    level's box bytes and half of them.  The kernel trace of this run gives
    each copy's duration."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    top = st.lvl == st.lvl.max()
    fin = int(st.box_sizes()[top].sum()) * 8
    flush_bytes = ctypes.c_size_t(1 << 31)
    bufs = [ctypes.c_void_p() for _ in range(3)]
    sizes = [flush_bytes, ctypes.c_size_t(fin), ctypes.c_size_t(fin)]
    dt.sync()
    for b, n in zip(bufs, sizes):
        if hip.hipMalloc(ctypes.byref(b), n) != 0:
            raise RuntimeError("hipMalloc failed")
    flush, x, y = bufs

    def check(rc, what):
        if rc != 0:
            raise RuntimeError(what + " failed")

    check(hip.hipMemset(x, 1, ctypes.c_size_t(fin)), "hipMemset")
    for n in (fin, fin // 2):
        check(hip.hipMemset(flush, 1, flush_bytes), "cache flush")
        check(hip.hipDeviceSynchronize(), "sync")
        for _ in range(4):
            check(hip.hipMemcpy(y, x, ctypes.c_size_t(n), 3), "hipMemcpy")   # 3: device to device
        check(hip.hipDeviceSynchronize(), "sync")
    check(hip.hipMemset(flush, 2, flush_bytes), "cache flush")
    check(hip.hipDeviceSynchronize(), "sync")
    for _ in range(4):
        dt.copy_cc(2, 3)
    dt.sync()
    for b in bufs:
        hip.hipFree(b)
    print(json.dumps({"copy_ceiling": {"flat_copy_bytes": [fin, fin // 2],
                                       "box_copy": "rhs to tmp over every box, one launch per level"}}))
    dt.close()
    return 0


def pass_instances(dim, nc, shape, pairs, n_down, n_up, split=True):
    """The colour-pass kernel instances of a level per V-cycle and how often
    each runs: a down block whose first pass reads the stored ghosts, an up
    block that computes them all (after the correction).  Fused pairs take
    two passes per launch; on 3D levels of 8^3 or 16^3 boxes a block of two
    or more single passes runs on the colour-split copy of phi (k_gsrb_split:
    mode 0 the first pass, 1 the middle ones, 2 the last; 3 the last of an
    up-sweep block, which also fills the level's ghosts; phi stays in the
    copy from the down-sweep's block to the up-sweep's)."""
    out = {}

    def add(name, k):
        if k:
            out[name] = out.get(name, 0) + k

    if pairs:
        add("k_gsrb_pair<%d, %d, 0, double, %d>" % (dim, nc, shape), 1)
        add("k_gsrb_pair<%d, %d, 1, double, %d>" % (dim, nc, shape), (2 * n_down) // 2 - 1 + (2 * n_up) // 2)
        return out
    # with an up-sweep block to write it back, the down-sweep's block leaves
    # phi in the copy (its last pass a middle one) and the up-sweep's starts
    # from it (its first pass a middle one)
    keep = n_up > 0
    for n, up in ((2 * n_down, 0), (2 * n_up, 1)):
        if split and dim == 3 and nc in (8, 16) and n >= 2:
            if up and keep:
                add("k_gsrb_split<%d, %d, 1, 1, double, %d>" % (dim, nc, shape), 1)
            else:
                add("k_gsrb_split<%d, %d, 0, %d, double, %d>" % (dim, nc, up, shape), 1)
            add("k_gsrb_split<%d, %d, 1, 1, double, %d>" % (dim, nc, shape), n - 2)
            last = 3 if up else (1 if keep else 2)
            add("k_gsrb_split<%d, %d, %d, 1, double, %d>" % (dim, nc, last, shape), 1)
        else:
            add("k_gsrb_tile<%d, %d, %d, double, %d>" % (dim, nc, up, shape), 1)
            add("k_gsrb_tile<%d, %d, 1, double, %d>" % (dim, nc, shape), n - 1)
    return out


def host_cpu():
    """The host's CPU model, its logical CPU count and the CPUs this process
    may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def baseline_threads():
    """All cores this process may use, within the host's thread budget
    (OMP_NUM_THREADS when set: the CPU share of one GPU on a shared host)."""
    _, _, usable = host_cpu()
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(usable, int(env)) if env.isdigit() and int(env) > 0 else usable


def cpu_baseline(p, cell_updates_per_solve, threads, n_time=5):
    """The reference solver, built from the reference sources (oracle/_ref),
    on the same tree: median solve time of n_time solves after one warm-up
    solve, one OpenMP thread per core (OMP_PROC_BIND=close,
    OMP_PLACES=cores)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from afivo_amd import problems as P
    dim = int(p.state.dim)
    if not O.ref_available(dim):
        return None
    import tempfile
    work = tempfile.mkdtemp(prefix="afg_cpu_")
    cfg = P.harness_cfg(p, n_time=n_time, dump_input=False, dump_output=False)
    prefix, _ = O.run_ref(dim, cfg, workdir=work, threads=threads, timeout=900, bind=True)
    n_thr, t = O.read_time(prefix)
    med = float(np.median(t))
    model, n_cpu, usable = host_cpu()
    return {"value": cell_updates_per_solve / med, "unit": "cell-updates/s", "cores": int(n_thr),
            "kind": "reference",
            "sample": f"same tree and options, whole solve (FMG + {p.n_vcycle} V-cycles), "
                      f"median of {len(t)} solves after one warm-up, OpenMP threads={n_thr}, "
                      f"OMP_PROC_BIND=close OMP_PLACES=cores",
            "s_per_solve": med, "cpu_model": model, "host_cpus": n_cpu, "usable_cpus": usable}


def solve_to_tol(dt, mg, tol, rhs_norm, max_cycles=30):
    """FMG, then V-cycles until every tree's max|residual| <= tol * max|rhs|
    (per tree for a forest).  Returns the number of V-cycles."""
    from afivo_amd import mg as M
    M.mg_fas_fmg(dt, mg, True, False)
    for k in range(max_cycles + 1):
        if np.all(dt.forest_max_abs(3) <= tol * rhs_norm):
            return k
        M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
    raise RuntimeError("no convergence")


def solve_to_tol_together(items, mg, tol, max_cycles=30):
    """solve_to_tol for several contexts at once: each has its own HIP stream,
    so their launches overlap on the device while the host checks one
    context's norms after the other.  items: (DeviceTree, rhs norms);
    returns the V-cycles of each."""
    from afivo_amd import mg as M
    for dt, _ in items:
        M.mg_fas_fmg(dt, mg, True, False)
    cycles = [0] * len(items)
    running = list(range(len(items)))
    while running:
        for i in list(running):
            dt, rn = items[i]
            if np.all(dt.forest_max_abs(3) <= tol * rn):
                running.remove(i)
            elif cycles[i] >= max_cycles:
                raise RuntimeError("no convergence")
            else:
                M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
                cycles[i] += 1
    return cycles


def mixed_to_tol_together(items, mg, tol, cycles, max_corrections=40):
    """FP64 defect correction with FP32 FAS corrections (`cycles` V-cycles
    each, the first an FMG) for several forests at once, each on its own
    stream, until every tree of each meets `tol`; returns the corrections of
    each."""
    from afivo_amd import mg as M
    for dt, _ in items:
        M.mixed_begin(dt, mg)
    corr = [0] * len(items)
    running = list(range(len(items)))
    while running:
        for i in list(running):
            dt, rn = items[i]
            if np.all(dt.forest_max_abs(3) <= tol * rn):
                running.remove(i)
            elif corr[i] >= max_corrections:
                raise RuntimeError("no convergence")
            else:
                M.mixed_correct(dt, mg, cycles)
                corr[i] += 1
    return corr


def persist_arg(text):
    """WORKGROUPS,CELLS of --persist."""
    wgs, cells = (int(v) for v in text.split(","))
    return wgs, cells


def bench_c3(a):
    """Batched 2D FP64: --trees independent trees solved to --tol relative
    residual, as forests (one per n_cell) through afg_set_forest, the forests
    advancing together on their own streams, against the same trees solved
    one at a time, each in its own context."""
    from afivo_amd import forest as F
    from afivo_amd import mg as M

    t0 = time.time()
    trees = [F.c3_tree(a.seed, t, l_uniform=3, p_ref=C3_P_REF) for t in range(a.trees)]
    build_s = time.time() - t0
    mg = M.default_mg(n_base=50)
    leaf_cells = np.array([int((st.children[:, 0] == 0).sum()) * spec.nc ** 2 for st, spec in trees])

    def fresh(st):   # phi = tmp = 0, rhs from the tree
        return [np.zeros_like(st.data[0]), st.data[1], np.zeros_like(st.data[2])]

    # batched: one forest per n_cell
    groups = {}
    for t, (st, spec) in enumerate(trees):
        groups.setdefault(spec.nc, []).append(t)
    forests = []
    for nc, idx in sorted(groups.items()):
        sts = [trees[t][0] for t in idx]
        fdt = M.DeviceTree.from_forest(sts)
        if a.persist:
            fdt.set_persist(*persist_arg(a.persist))
        fst, _ = F.combine(sts)
        rhs_norm = np.array([np.abs(st.data[1][st.leaf_mask()]).max() for st in sts])
        forests.append((fdt, fst, rhs_norm))
    cycles_batch = []
    for rep in range(2):   # the first solve captures the graphs
        for fdt, fst, _ in forests:
            for iv, v in enumerate(fresh(fst), start=1):
                fdt.upload(iv, v)
            fdt.sync()
        tb0 = time.perf_counter()
        cycles_batch = solve_to_tol_together([(fdt, rn) for fdt, _, rn in forests], mg, a.tol)
        for fdt, _, _ in forests:
            fdt.sync()
        t_batch = time.perf_counter() - tb0
    phi64 = [fdt.download(1) for fdt, _, _ in forests]
    # the same forests by the mixed-precision batched solver: FP32 FAS
    # corrections of 2 or 3 V-cycles in FP64 defect correction (the faster is
    # reported), each solve after a warm-up solve
    mixed = {}
    for cyc in (2, 3, 4):
        for rep in range(2):
            for fdt, fst, _ in forests:
                for iv, v in enumerate(fresh(fst), start=1):
                    fdt.upload(iv, v)
                fdt.sync()
            tm0 = time.perf_counter()
            corr = mixed_to_tol_together([(fdt, rn) for fdt, _, rn in forests], mg, a.tol, cyc)
            for fdt, _, _ in forests:
                fdt.sync()
            t_mixed = time.perf_counter() - tm0
        dphi = max(float(np.abs(fdt.download(1) - p64).max() / np.abs(p64).max())
                   for (fdt, _, _), p64 in zip(forests, phi64))
        mixed[cyc] = {"ms": 1e3 * t_mixed, "solves_per_s": a.trees / t_mixed, "corrections": corr,
                      "max_phi_difference_vs_fp64": dphi}
    for fdt, _, _ in forests:
        fdt.close()
    best = min(mixed, key=lambda k: mixed[k]["ms"])

    # one tree at a time, each tree in its own context, warm (second solve)
    t_loop, cycles_loop = 0.0, []
    for st, spec in trees:
        dt = M.DeviceTree.from_state(st)
        rn = np.abs(st.data[1][st.leaf_mask()]).max()
        for rep in range(2):
            for iv, v in enumerate(fresh(st), start=1):
                dt.upload(iv, v)
            dt.sync()
            s0 = time.perf_counter()
            k = solve_to_tol(dt, mg, a.tol, rn)
            dt.sync()
            if rep == 1:
                t_loop += time.perf_counter() - s0
                cycles_loop.append(k)
        dt.close()

    out = {
        "metric": "solves/s (batched, relative residual <= %g)" % a.tol,
        "value": a.trees / t_batch,
        "unit": "solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * t_batch,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "batched 2D FP64 Poisson: %d trees from seed %d, n_cell 8 or 16, uniform to "
                        "level 3, then boxes below L ~ U{4..8} refined where a hash of (seed, tree, "
                        "level, index) < %g; rhs 1-3 Gaussians (sigma U[0.02,0.08]); Dirichlet 0; "
                        "FMG + V-cycles (2+2, 50 base) until max|res| <= %g max|rhs| per tree; "
                        "one forest per n_cell, the forests on their own streams"
                        % (a.trees, a.seed, C3_P_REF, a.tol),
            "leaf_cells_total": int(leaf_cells.sum()),
            "leaf_cells_per_tree": [int(leaf_cells.min()), int(np.median(leaf_cells)), int(leaf_cells.max())],
            "forests": {str(nc): len(idx) for nc, idx in sorted(groups.items())},
            "persist": list(persist_arg(a.persist)) if a.persist else [0, 0],
        },
        "one_at_a_time_solves_per_s": a.trees / t_loop,
        "speedup_vs_one_at_a_time": t_loop / t_batch,
        "mixed_batched": dict(mixed[best], vcycles_per_correction=best,
                              vs_fp64_batched=t_batch / (mixed[best]["ms"] / 1e3),
                              vs_fp64_one_at_a_time=t_loop / (mixed[best]["ms"] / 1e3),
                              by_vcycles_per_correction={str(k): v for k, v in mixed.items()}),
        "vcycles_batch": cycles_batch,
        "vcycles_one_at_a_time": [int(min(cycles_loop)), int(max(cycles_loop))],
        "host_tree_build_s": build_s,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


C3_P_REF = 0.3


def bench_c4(a):
    """Mixed precision: the 3D variable-eps tree solved to --tol relative
    residual by FP64 defect correction with FP32 multigrid corrections, and by
    the plain FP64 solver (FMG + V-cycles until the tolerance), both on the
    device; the error is the difference of the two solutions."""
    from afivo_amd import mg as M
    from afivo_amd import problems as P
    t0 = time.time()
    p = P.c4(seed=a.seed)
    st = p.state
    build_s = time.time() - t0
    mg = M.default_mg(eps=True, n_base=p.n_base)
    mask = st.leaf_mask()
    rhs_norm = float(np.abs(st.data[1][mask]).max())

    dt = M.DeviceTree.from_state(st)
    if a.no_graphs:
        dt.set_graphs(False)
    # the autotuner picks the tile shapes per level for both precisions
    # (on a copy of phi, before timing)
    dt.autotune(mg)
    dt.autotune(mg, fp32=True)
    tiles = {"fp64": [dt.tile(lvl) for lvl in range(dt.min_lvl, dt.highest_lvl + 1)],
             "fp32": [dt.tile(lvl, True) for lvl in range(dt.min_lvl, dt.highest_lvl + 1)],
             "bottom_cells": [dt.bottom_cells(), dt.bottom_cells(True)]}

    def reset():
        dt.upload(1, st.data[0])
        dt.upload(3, st.data[2])
        dt.upload(2, st.data[1])
        dt.sync()

    def fp64_solve():
        M.mg_fas_fmg(dt, mg, True, False)
        k = 0
        while dt.forest_max_abs(3)[0] > a.tol * rhs_norm and k < 40:
            M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
            k += 1
        return k

    times = {}
    cyc = [a.mixed_cycles] if a.mixed_cycles else [1, 2, 3, 4, 5]
    runs = ([("fp64", fp64_solve)] if a.c4_solver in ("both", "fp64") else []) + \
        ([("mixed%d" % k, lambda k=k: M.mixed_solve(dt, mg, a.tol, 40, cycles=k)) for k in cyc]
         if a.c4_solver in ("both", "mixed") else [])
    phi64 = None
    for name, fn in runs:
        best = None
        for rep in range(3):          # the first solve captures the graphs
            reset()
            s0 = time.perf_counter()
            info = fn()
            dt.sync()
            dt_s = time.perf_counter() - s0
            if rep > 0:
                best = dt_s if best is None else min(best, dt_s)
        times[name] = (best, info)
        if name == "fp64":
            phi64 = dt.download(1)
            scale = float(np.abs(phi64[mask]).max())
        elif phi64 is not None:
            phi_mixed = dt.download(1)
            times[name] += (float(np.abs(phi_mixed[mask] - phi64[mask]).max()) / scale,)
        else:
            times[name] += (None,)
    # the precision split of the mixed solver: the FP32 corrections' coarsest-grid
    # solve, their update_coarse (the FAS tau), and both, in FP64 arithmetic,
    # at the fastest cycles-per-correction count
    splits = {}
    mixed_names = [k for k in times if k.startswith("mixed")]
    if mixed_names:
        kbest = int(min(mixed_names, key=lambda k: times[k][0])[5:])
        for name, base, tau in (("coarsest_solve", True, False), ("tau", False, True), ("tau_and_coarsest_solve", True, True)):
            dt.mixed_fp64_base(base)
            dt.mixed_fp64_tau(tau)
            best_t, info = None, None
            for rep in range(3):
                reset()
                s0 = time.perf_counter()
                info = M.mixed_solve(dt, mg, a.tol, 40, cycles=kbest)
                dt.sync()
                if rep > 0:
                    t = time.perf_counter() - s0
                    best_t = t if best_t is None else min(best_t, t)
            diff = (float(np.abs(dt.download(1)[mask] - phi64[mask]).max()) / scale) if phi64 is not None else None
            splits[name] = {"ms": 1e3 * best_t, "corrections": info[0], "rel_residual": info[1],
                            "vcycles_per_correction": kbest, "phi_difference": diff}
        dt.mixed_fp64_base(False)
        dt.mixed_fp64_tau(False)
    dt.close()
    if a.c4_solver == "fp64":     # profiling run of the FP64 solver alone
        print(json.dumps({"c4_solver": "fp64", "ms": 1e3 * times["fp64"][0],
                          "vcycles_after_fmg": times["fp64"][1]}))
        return 0
    best = min((k for k in times if k.startswith("mixed")), key=lambda k: times[k][0])
    times["mixed"] = times[best]
    err = times["mixed"][2]
    it, rr = times["mixed"][1]
    leaf_cells = int(mask.sum())
    out = {
        "metric": "solves/s to relative residual <= %g (mixed precision)" % a.tol,
        "value": 1.0 / times["mixed"][0], "unit": "solves/s",
        "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * times["mixed"][0],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 in f64 defect correction", "data": "synthetic",
        "config": {
            "workload": "3D variable-coefficient Poisson div(eps grad phi) = rhs: 16^3 boxes, uniform "
                        "to level 3, adaptive to level 8 around 6 Gaussians (seed %d, threshold %g); "
                        "eps = 1 below z = 0.5 and %.3f above (harmonic face means, mg3_auto_*); "
                        "Dirichlet 0; mixed: FP32 FAS corrections (1-5 V-cycles each, the first an FMG; "
                        "the fastest reported) in FP64 defect correction; fp64: FMG + V-cycles (2+2, 50 base)"
                        % (a.seed, P.THR_C4, p.eps_hi),
            "boxes": int(st.n_boxes), "leaf_cells": leaf_cells,
            "levels": [int(st.min_lvl), int(st.highest_lvl)],
            "tile_shapes": tiles,
        },
        "mixed": {"ms": 1e3 * times["mixed"][0], "corrections": it, "rel_residual": rr,
                  "vcycles_per_correction": int(best[5:])},
        "mixed_by_vcycles_per_correction": {k[5:]: {"ms": 1e3 * v[0], "corrections": v[1][0],
                                                     "rel_residual": v[1][1], "phi_difference": v[2]}
                                            for k, v in times.items() if k.startswith("mixed") and k != "mixed"},
        "fp64": {"ms": 1e3 * times["fp64"][0], "vcycles_after_fmg": times["fp64"][1],
                 "solves_per_s": 1.0 / times["fp64"][0]} if "fp64" in times else None,
        "speedup_vs_fp64": times["fp64"][0] / times["mixed"][0] if "fp64" in times else None,
        "mixed_fp64_coarsest_solve": splits.get("coarsest_solve"),
        "mixed_fp64_tau": splits.get("tau"),
        "mixed_fp64_tau_and_coarsest_solve": splits.get("tau_and_coarsest_solve"),
        "max_phi_difference_vs_fp64_relative": err,
        "host_tree_build_s": build_s,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def bench_c5(a):
    """Serving: a Poisson-arrival request trace through the server
    (continuous batching at request granularity), the same requests offered
    at once (the server's capacity) and at a sweep of arrival rates (latency
    against load), and the same trace through the plain FP64 solver one
    request at a time."""
    from afivo_amd import serving as S
    t0 = time.time()
    trace = S.make_trace(a.seed, a.requests, a.rate)
    build_s = time.time() - t0
    cells = np.array([r.leaf_cells for r in trace])

    print("c5: %d requests built in %.1f s" % (len(trace), build_s), file=sys.stderr, flush=True)

    def run(rate, seed):
        print("c5: serving at %.4g requests/s" % rate, file=sys.stderr, flush=True)
        S.retime(trace, rate, seed)
        log = {}
        S.serve(trace, precision=a.precision, big_cells=a.big_cells, log=log)
        out = S.summary(trace)
        out["steps_per_request"] = [int(min(r.cycles for r in trace)), int(max(r.cycles for r in trace))]
        out["lanes"], out["compactions"], out["joined"] = (int(log.get(k, 0)) for k in ("lanes", "compactions", "joined"))
        return out

    S.serve(S.make_trace(a.seed + 100, 4, 1e4, cells=(1e4, 1e5)), precision=a.precision)   # warm-up
    served = run(a.rate, a.seed)   # retime with the trace's seed: its own arrival times
    saturated = run(1e6, a.seed)
    sweep = {}
    for f in (0.25, 0.5, 1.0, 2.0):
        rate = f * saturated["solves_per_s"]
        sweep["%.0f" % rate] = run(rate, a.seed + 1)
    print("c5: one at a time", file=sys.stderr, flush=True)
    S.retime(trace, a.rate, a.seed)
    S.one_at_a_time(trace, fmg=True)
    solo = S.summary(trace)
    out = {
        "metric": "solves/s and p50/p99 latency (served, %s)" % ("mixed precision" if a.precision == "mixed"
                                                                 else "FP64"),
        "value": served["solves_per_s"], "unit": "solves/s",
        "n_gpus": 1, "steps": a.requests, "warmup": 0,
        "ms_per_step": 1e3 / served["solves_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 in f64 defect correction" if a.precision == "mixed" else "f64", "data": "synthetic",
        "config": {
            "workload": "serving: %d requests, Poisson arrivals at %g/s (seed %d); 70%% 2D / 30%% 3D "
                        "trees, n_cell 8, leaf cells log-uniform in [1e4, 1e7] (uniform to the level "
                        "below the target, then two levels around 1-4 Gaussians); tolerance 1e-8 or "
                        "1e-10 relative residual; %s; continuous batching at request granularity: lanes "
                        "(forests of one class: dimension%s) of up to 64 trees, at most 6 lanes, admission "
                        "by HBM budget, set up by 3 host threads; a tree leaves its lane when it converges, "
                        "and a lane half of whose boxes have left is rebuilt with its live trees and the "
                        "waiting ones of its class; baseline: the plain FP64 solver (FMG + V-cycles), one "
                        "request at a time"
                        % (a.requests, a.rate, a.seed,
                           "mixed precision (FP64 defect correction, FP32 FAS corrections of 2 V-cycles)"
                           if a.precision == "mixed" else "FP64 V-cycles from zero",
                           ", and trees above %g leaf cells apart" % a.big_cells if a.big_cells else ""),
            "leaf_cells": [int(cells.min()), int(np.median(cells)), int(cells.max())],
        },
        "served": served,
        "saturated": saturated,
        "latency_vs_offered_rate": sweep,
        "one_at_a_time": solo,
        "host_tree_build_s": build_s,
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def bench_update(a):
    """Tree updates: the 3D benchmark tree (phi, rhs, tmp on the device)
    refined around one point and derefined elsewhere, and back, alternately
    (afg_update_tree with a$D_prolong2 of phi into the added boxes), against
    the round trip a program without it makes: download phi, rhs and tmp,
    set the new tree, upload the three variables (the host-side refinement
    and prolongation not counted)."""
    from afivo_amd import _lib
    from afivo_amd import mg as M
    from afivo_amd import problems as P
    from afivo_amd import tree as T
    t0 = time.time()
    p = P.c2(seed=a.seed)
    st = p.state
    recs_a = st.recs.copy()
    c0 = np.array([0.31, 0.62, 0.45])

    def ref(view, cur):
        out = np.array(cur, dtype=np.int64, copy=True) if np.ndim(cur) else np.full(view.ids.size, cur)
        mid = view.r_min + 0.5 * view.n_cell[:, None] * view.dr[:, None]
        d = np.sqrt(((mid - c0) ** 2).sum(axis=1))
        out[(d < 0.12) & (view.lvl < 8)] = T.REFINE
        out[(d > 0.3) & (view.lvl > 4)] = T.DEREFINE
        return out

    n_add = p.tree.adjust_refinement(ref)
    recs_b = p.tree.records()
    build_s = time.time() - t0
    ka = {tuple(r[1:5]) for r in recs_a}
    kb = {tuple(r[1:5]) for r in recs_b}
    dt = M.DeviceTree.from_state(st)
    stride = (st.nc + 2) ** 3
    seq = [recs_b, recs_a] * (a.warmup + a.steps)
    times = []
    for k, recs in enumerate(seq):
        dt.sync()
        s0 = time.perf_counter()
        dt.update_tree(recs, prolong_order=2, vars=(1,))
        dt.sync()
        if k >= 2 * a.warmup:
            times.append(time.perf_counter() - s0)
    ms_update = 1e3 * float(np.median(times))
    # the round trip without a device-side update
    L = _lib.load()
    trip = []
    for k, recs in enumerate(seq[:2 * (1 + min(a.steps, 5))]):
        dt.sync()
        s0 = time.perf_counter()
        data = [dt.download(iv) for iv in (1, 2, 3)]
        recs_c = np.ascontiguousarray(recs, dtype=np.int32)
        M.check(L.afg_set_tree(dt.h, recs_c.shape[0], M._i32(recs_c), float(st.dr_base)), dt.h)
        dt.records, dt.ids = recs_c, np.ascontiguousarray(recs_c[:, 0])
        dt._tree_info()
        new_vals = dt.n_values
        for iv in (1, 2, 3):
            # the host would upload the new tree's arrays; same byte count
            dt.upload(iv, np.resize(data[iv - 1], new_vals))
        dt.sync()
        if k >= 2:
            trip.append(time.perf_counter() - s0)
    ms_trip = 1e3 * float(np.median(trip))
    dt.close()
    n_b, n_a = len(recs_b), len(recs_a)
    kept = len(ka & kb)
    # algorithmic bytes of one update A -> B: every variable's pool gathered
    # (kept boxes read, all boxes written), plus phi of the added boxes
    gather_bytes = 3 * 8 * stride * (kept + n_b)
    out = {
        "metric": "tree updates/s (device-side refinement step)",
        "value": 1e3 / ms_update, "unit": "updates/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_update,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "tree update of the 3D benchmark tree (8^3 boxes, %d boxes): leaves within 0.12 "
                        "of (0.31, 0.62, 0.45) refined (below level 8), leaves beyond 0.3 above level 4 "
                        "derefined, %d boxes after (%d added, %d removed), and back; phi, rhs, tmp kept, "
                        "phi prolonged (a$D_prolong2) into the added boxes" % (n_a, n_b, n_add, n_a - kept),
            "boxes": [n_a, n_b], "kept": kept,
        },
        "update_ms": ms_update,
        "round_trip_ms": ms_trip,
        "speedup_vs_round_trip": ms_trip / ms_update,
        "gather_bytes_per_update": gather_bytes,
        "host_tree_build_s": build_s,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def bench_transfers(a):
    """Ghost cells and transfers of any variable on the 3D benchmark tree
    (random values): the time of one whole-tree call of each routine, and
    the HBM rate of its algorithmic bytes (one read and one write of each
    value the routine produces, plus the values it reads once more)."""
    from afivo_amd import mg as M
    from afivo_amd import problems as P
    p = P.c2(seed=a.seed)
    st = p.state
    rng = np.random.default_rng(3)
    st.data = [rng.standard_normal(st.n_values()) for _ in range(3)]
    dt = M.DeviceTree.from_state(st)
    dt.upload_fc(1, rng.standard_normal(int(st.face_offsets()[-1])))
    nc, dim = st.nc, 3
    n = st.ncell()
    ghost = int((2 * dim * n ** (dim - 1)).sum())
    fine = int((n ** dim)[st.lvl > st.min_lvl].sum())
    vals = st.n_values()
    # algorithmic bytes: ghost fill reads the source cell and writes the
    # ghost (16 B per ghost); restriction reads every fine cell, writes
    # 1/2^dim; prolongation reads 1/2^dim, writes every fine cell;
    # lincomb reads two values and writes one
    ops = {
        "gc_tree_mg_sides_rb": (lambda: dt.gc_tree(1, 0, 0), 16.0 * ghost),
        "gc_tree_interp": (lambda: dt.gc_tree(1, 2, 0), 16.0 * ghost),
        "gc_tree_interp_lim_neumann": (lambda: dt.gc_tree(1, 3, -2), 16.0 * ghost),
        "restrict_tree": (lambda: dt.restrict_tree(2, 3), 8.0 * fine * (1 + 2.0 ** -dim)),
        "prolong2_all_levels": (lambda: [dt.prolong_level(1, 3, l, 2) for l in range(st.min_lvl + 1, st.highest_lvl + 1)],
                                8.0 * fine * (1 + 2.0 ** -dim)),
        "consistent_fluxes": (lambda: dt.consistent_fluxes([1]), None),
        "lincomb_cc": (lambda: dt.lincomb_cc(0.5, 1, 2.0, 3), 24.0 * vals),
    }
    out = {}
    for name, (fn, nbytes) in ops.items():
        for _ in range(max(a.warmup, 1)):
            fn()
        dt.sync()
        reps = max(a.steps, 5)
        s0 = time.perf_counter()
        for _ in range(reps):
            fn()
        dt.sync()
        ms = 1e3 * (time.perf_counter() - s0) / reps
        out[name] = {"ms": ms, "GB_per_s": None if nbytes is None else nbytes / ms / 1e6}
    dt.close()
    head = out["gc_tree_interp"]
    res = {
        "metric": "whole-tree ghost fills/s (a$D_gc_tree with a$D_gc_interp)",
        "value": 1e3 / head["ms"], "unit": "fills/s",
        "n_gpus": 1, "steps": max(a.steps, 5), "warmup": max(a.warmup, 1), "ms_per_step": head["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "ghost cells and transfers on the 3D benchmark tree (%d boxes of 8^3, levels "
                               "%d..%d), random values; host-timed whole-tree calls (every level's launch "
                               "included)" % (st.n_boxes, st.min_lvl, st.highest_lvl),
                   "ghost_cells": ghost, "fine_cells": fine},
        "routines": out,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(res))
    return 0


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {a.gpus} is not supported "
              "(sharding over several GPUs is designed, not built)", file=sys.stderr)
        return 2
    if a.workload == "c3":
        return bench_c3(a)
    if a.workload == "c5":
        return bench_c5(a)
    if a.workload == "c4":
        return bench_c4(a)
    if a.workload == "update":
        return bench_update(a)
    if a.workload == "transfers":
        return bench_transfers(a)

    from afivo_amd import mg as M
    from afivo_amd import problems as P

    t0 = time.time()
    p = P.c1() if a.workload == "c1" else P.c2(seed=a.seed)
    st = p.state
    dim = int(st.dim)
    build_s = time.time() - t0
    dt = M.DeviceTree.from_state(st, 3)
    mg = M.default_mg(n_base=p.n_base)
    if a.copy_calibration:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        buf = ctypes.c_void_p()
        flush_bytes = ctypes.c_size_t(1 << 31)            # 2 GiB written: evicts L2 and MALL
        dt.sync()
        if hip.hipMalloc(ctypes.byref(buf), flush_bytes) != 0:
            raise RuntimeError("hipMalloc of the cache-flush buffer failed")
        if hip.hipMemset(buf, 1, flush_bytes) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("cache flush failed")
        hip.hipFree(buf)
        dt.copy_cc(2, 3)
        dt.sync()
        nbytes = int(((st.ncell().astype(np.int64) + 2) ** 3).sum()) * 8
        print(json.dumps({"copy_calibration": {"kernel": "k_copy_boxes", "bytes_read": nbytes,
                                               "bytes_written": nbytes}}))
        dt.close()
        return 0
    if a.copy_ceiling:
        return copy_ceiling(dt, st)

    # the autotuner picks the tile shape and pass mode of the colour passes
    # per level and the bottom kernel's limit (on a copy of phi, before
    # timing), unless they are given
    lvls = list(range(dt.min_lvl, dt.highest_lvl + 1))
    if a.tiles or a.pairs or a.bottom_cells >= 0:
        if a.tiles:
            for lvl, v in zip(lvls, a.tiles.split(",")):
                dt.set_tile(lvl, int(v))
        if a.pairs:
            for lvl, v in zip(lvls, a.pairs.split(",")):
                dt.set_level_pass_pairs(lvl, int(v) != 0)
        if a.bottom_cells >= 0:
            dt.set_bottom_cells(a.bottom_cells)
        tuned = False
    else:
        dt.autotune(mg)
        tuned = True
    if a.persist:
        dt.set_persist(*persist_arg(a.persist))
    tiles = [dt.tile(lvl) for lvl in lvls]
    pairs = [int(dt.pass_pairs(lvl)) for lvl in lvls]
    bottom = dt.bottom_cells()
    # cell-updates of one solve: FMG runs a V-cycle up to every level
    cu_solve = sum(dt.vcycle_cell_updates(mg, lvl) for lvl in range(dt.min_lvl, dt.highest_lvl + 1))
    cu_solve += p.n_vcycle * dt.vcycle_cell_updates(mg)

    def solve():
        M.mg_fas_fmg(dt, mg, True, False)
        for _ in range(p.n_vcycle):
            M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)

    for _ in range(a.warmup):
        solve()
    dt.sync()
    rel_res = M.max_abs_cc(dt, 3) / max(abs(M.tree_min_cc(dt, 2)), abs(M.tree_max_cc(dt, 2)))

    if not a.no_timing:
        dt.set_timing(True, dt.highest_lvl)   # the finest level's colour passes
    dt.sync()
    t1 = time.perf_counter()
    for _ in range(a.steps):
        solve()
    dt.sync()
    t2 = time.perf_counter()
    elapsed = t2 - t1
    n_launch, k_ms, k_bytes = dt.timing()
    dt.set_timing(False)

    value = cu_solve * a.steps / elapsed
    if a.workload == "c1":
        workload = ("2D FP64 Poisson (poisson_basic_2d form), 8^2 boxes, uniform to level 5, refined to level 7 "
                    "where dr^2 max|rhs| > %g, rhs = Laplacian of two Gaussians (sigma 0.04); step = FMG + %d "
                    "V-cycles, 2+2 GSRB sweeps, %d base sweeps" % (p.thr, p.n_vcycle, p.n_base))
    else:
        workload = ("3D FP64 Poisson (poisson_basic_3d form), 8^3 boxes, uniform to level 4, "
                    "adaptive to level 7 around 4 Gaussians (sigma 0.03, seed %d), "
                    "threshold dr^2 max|rhs| > %g; step = FMG + %d V-cycles, 2+2 GSRB sweeps, "
                    "%d base sweeps" % (a.seed, p.thr, p.n_vcycle, p.n_base))
    out = {
        "metric": "cell-updates/s",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "boxes": int(st.n_boxes),
            "leaf_cells": int(p.tree.n_leaf_cells()),
            "cells_all_levels": int((st.ncell().astype(np.int64) ** dim).sum()),
            "levels": [int(st.min_lvl), int(st.highest_lvl)],
            "cell_updates_per_step": cu_solve,
            "tile_shapes": tiles,
            "pass_pairs": pairs,
            "bottom_cells": bottom,
            "persist": list(dt.persist()),
            "autotuned": tuned,
        },
        "solves_per_s": a.steps / elapsed,
        "rel_residual_after_step": rel_res,
        "host_tree_build_s": build_s,
    }
    if n_launch and k_ms > 0:
        achieved = k_bytes / (k_ms * 1e-3) / 1e9
        n_top = int((st.lvl == st.highest_lvl).sum())
        inst = pass_instances(dim, st.nc, tiles[-1], pairs[-1], mg.n_cycle_down, mg.n_cycle_up)
        out["roofline"] = {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            # counter traffic is profiled for the 3D tree's finest level only
            "traffic": measured_traffic(inst) if a.workload == "c2" else None,
            "kernel": "%s (GSRB colour passes) on the finest level (%d, %d boxes)"
                      % (" and ".join(inst), st.highest_lvl, n_top),
            "instances_per_vcycle": inst,
            # the whole step at the pass's algorithmic bytes per cell-update,
            # (3 + 2D/nc) w: every level's relaxation counted, nothing else
            "whole_step_alg_GBs": cu_solve * (3 + 2 * dim / st.nc) * 8 * a.steps / elapsed / 1e9,
            "whole_step_frac": cu_solve * (3 + 2 * dim / st.nc) * 8 * a.steps / elapsed / 1e9 / HBM_PEAK_GBS,
            "launches": n_launch, "avg_launch_us": 1e3 * k_ms / n_launch,
            "alg_bytes_per_launch": k_bytes / n_launch,
            "share_of_step_time": (k_ms * 1e-3) / elapsed,
        }
    else:
        out["roofline"] = None
    if not a.no_cpu_baseline:
        thr = a.cpu_threads or baseline_threads()
        out["cpu_baseline"] = cpu_baseline(p, cu_solve, thr)
        one = cpu_baseline(p, cu_solve, 1, n_time=2) if out["cpu_baseline"] and thr > 1 else None
        if one:   # the same on one core, for the per-core figure
            out["cpu_baseline"]["one_thread"] = {"value": one["value"], "s_per_solve": one["s_per_solve"],
                                                 "sample": one["sample"]}
    else:
        out["cpu_baseline"] = None
    dt.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
