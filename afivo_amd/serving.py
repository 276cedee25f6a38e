"""Serving: Poisson-solve requests arriving over time, solved on one GPU with
continuous batching.

A request is one tree (2D or 3D) with its right-hand side and a tolerance on
the relative residual max|rhs - A phi| / max|rhs| over leaf interiors.  The
server keeps a set of *lanes*; a lane is one forest context (afg_set_forest:
its own device pools, tables and HIP stream) whose trees run through the
same FMG / V-cycle stages together, every stage one launch over the boxes of
a level in all the lane's trees.  Requests that arrive while lanes are busy
wait; whenever a lane is free and the device memory of the running lanes
plus the new forest fits the HBM budget, the waiting requests of one
(dimension, n_cell) class -- oldest first, up to `max_trees` -- are admitted
as a new lane.  Lanes advance in turn, one V-cycle each, so the work of
several lanes overlaps on the device; a lane retires when each of its trees
meets its tolerance, and its results are copied back.

Lanes solve in mixed precision by default: FP64 defect correction with FP32
FAS multigrid corrections (afg_mixed_begin / afg_mixed_correct), one
correction of `cycles` FP32 V-cycles per step; with precision="fp64" a step is
one FP64 V-cycle after an FP64 FMG.

A request's result does not depend on the mix it was solved in beyond its
tolerance: a tree in a forest goes through exactly the stages it would go
through alone, and only the number of steps (the lane stops when its last
tree converges) can differ.

`one_at_a_time` replays the same trace one request after the other, each in
its own context: through the plain FP64 solver, or through the same
mixed-precision steps as a lane of one tree.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import forest as F
from . import mg as M
from . import problems as P
from . import tree as treemod


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the trace starts
    state: object             # TreeState with phi = 0, rhs, tmp = 0
    tol: float
    rhs_norm: float
    leaf_cells: int
    # filled by the server
    start: float = -1.0
    finish: float = -1.0
    cycles: int = -1
    residual: float = float("nan")
    phi: np.ndarray | None = None


def request_tree(seed, k, dim, target_cells, nc=8):
    """A tree of about `target_cells` leaf cells: uniform to the level whose
    cell count is nearest below the target, then two levels of refinement
    around 1-4 Gaussians (from hashes of (seed, k))."""
    h = lambda *a: F._hash(seed, k, *a)   # noqa: E731
    n_g = 1 + int(h(1) * 4)
    amp = np.array([1.0 if h(10, g) < 0.5 else -1.0 for g in range(n_g)])
    sig = np.array([0.02 + 0.06 * h(11, g) for g in range(n_g)])
    r0 = np.array([[0.2 + 0.6 * h(12, g, d) for d in range(dim)] for g in range(n_g)])
    gs = P.Gaussians(amp, sig, r0, exact=False)
    l_u = 1
    while (nc * 2 ** l_u) ** dim <= target_cells:   # cells of level l_u + 1
        l_u += 1
    ref = P.gauss_refinement(gs, l_u, l_u + 2, 1e-3 if dim == 2 else 3e-3)
    tr = treemod.build(dim, nc, ref)
    st = tr.state()
    rhs = P.set_rhs(tr, st, gs)
    zeros = np.zeros_like(rhs)
    st.data = [zeros.copy(), rhs, zeros.copy()]
    st.n_var = 3
    return st


def make_trace(seed, n_requests, rate, frac_2d=0.7, cells=(1e4, 1e7), tols=(1e-8, 1e-10)):
    """Requests with Poisson arrivals at `rate` per second: 2D with
    probability frac_2d (else 3D), leaf cells log-uniform in `cells`,
    tolerance drawn from `tols`."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    out = []
    for k in range(n_requests):
        dim = 2 if rng.random() < frac_2d else 3
        target = float(np.exp(rng.uniform(np.log(cells[0]), np.log(cells[1]))))
        tol = float(tols[int(rng.integers(len(tols)))])
        st = request_tree(seed, k, dim, target)
        leaf = st.leaf_mask()
        out.append(Request(k, float(t[k]), st, tol, float(np.abs(st.data[1][leaf]).max()),
                           int(leaf.sum())))
    return out


MAX_STEPS = 40


class Lane:
    """One forest of admitted requests in flight."""

    def __init__(self, reqs, mg, now, precision="mixed", cycles=2):
        self.reqs = reqs
        self.states = [r.state for r in reqs]
        self.dt = M.DeviceTree.from_forest(self.states)
        self.mg = mg
        self.mixed = precision == "mixed"
        self.n_cyc = cycles
        self.tol = np.array([r.tol * r.rhs_norm for r in reqs])
        self.cycles = 0
        for r in reqs:
            r.start = now
        if self.mixed:
            M.mixed_begin(self.dt, mg)
        else:
            M.mg_fas_fmg(self.dt, mg, True, False)

    def step(self):
        """Check convergence of the last step; launch the next one if needed
        (a mixed-precision correction, or an FP64 V-cycle).  Returns True when
        every tree meets its tolerance."""
        res = self.dt.forest_max_abs(3)          # waits for this lane's work
        if np.all(res <= self.tol) or self.cycles >= MAX_STEPS:
            self.res = res
            return True
        if self.mixed:
            M.mixed_correct(self.dt, self.mg, self.n_cyc)
        else:
            M.mg_fas_vcycle(self.dt, self.mg, self.dt.highest_lvl, True)
        self.cycles += 1
        return False

    def retire(self, now):
        phis = F.split(self.states, self.dt.download(1))
        for r, phi, res in zip(self.reqs, phis, self.res):
            r.finish, r.cycles, r.residual, r.phi = now, self.cycles, float(res) / r.rhs_norm, phi

    def device_bytes(self):
        return self.dt.device_bytes()


def forest_bytes(reqs, n_var=3, precision="mixed"):
    """Device bytes a forest of these requests takes (pools and tables; the
    FP32 copies of phi, rhs and tmp in mixed precision)."""
    boxes = sum(r.state.n_boxes for r in reqs)
    st = reqs[0].state
    stride = -(-(st.nc + 2) ** st.dim // 16) * 16
    per_box = n_var * stride * 8 + (3 * stride * 4 if precision == "mixed" else 0)
    return boxes * (per_box + 64)


def serve(trace, hbm_budget=64 << 30, max_lanes=4, max_trees=64, n_base=50, lane_log=None,
          builders=3, precision="mixed", cycles=2):
    """Run the trace through the continuously batching server; returns the
    requests with start/finish times (seconds after the trace start).

    Lanes are set up (forest tables, uploads, the first launches) by
    `builders` host threads while the main thread drives the running lanes,
    so host work overlaps device work.  If given, `lane_log` receives one
    (admitted, setup_s, retired, n_trees, steps) tuple per lane."""
    import concurrent.futures as cf

    mg = M.default_mg(n_base=n_base)
    fbytes = lambda g: forest_bytes(g, precision=precision)   # noqa: E731
    pending = collections.deque(sorted(trace, key=lambda r: r.arrival))
    waiting, lanes, building = [], [], []
    # Retired lanes release their device memory only while no lane runs (or
    # when memory is needed): freeing device memory waits for all the
    # device's work, which would stall the running lanes.
    retired = []

    def release():
        for lane in retired:
            lane.dt.close()
        retired.clear()

    t0 = time.perf_counter()

    def build(group, t_adm):
        lane = Lane(group, mg, t_adm, precision, cycles)
        lane.t_adm, lane.setup = t_adm, time.perf_counter() - t0 - t_adm
        return lane

    with cf.ThreadPoolExecutor(max_workers=builders) as pool:
        while pending or waiting or lanes or building:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            if not lanes and not building:
                release()
            # admission: one class at a time, oldest waiting request first
            while waiting and len(lanes) + len(building) < max_lanes:
                cls = (waiting[0].state.dim, waiting[0].state.nc)
                group = [r for r in waiting if (r.state.dim, r.state.nc) == cls][:max_trees]
                used = (sum(l.device_bytes() for l in lanes + retired) +
                        sum(fbytes(g) for g, _ in building))
                if retired and used + fbytes(group[:1]) > hbm_budget:
                    if building:
                        break                           # release once the builders are done
                    release()
                    used = sum(l.device_bytes() for l in lanes)
                while len(group) > 1 and used + fbytes(group) > hbm_budget:
                    group = group[:-1]
                if used + fbytes(group) > hbm_budget and (lanes or building):
                    break                               # wait for memory to free up
                for r in group:
                    waiting.remove(r)
                building.append((group, pool.submit(build, group, time.perf_counter() - t0)))
            for item in list(building):
                if item[1].done():
                    building.remove(item)
                    lanes.append(item[1].result())
            if not lanes:
                if building:
                    cf.wait([f for _, f in building], timeout=0.002, return_when=cf.FIRST_COMPLETED)
                elif pending:
                    time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
                continue
            for lane in list(lanes):
                if lane.step():
                    t_ret = time.perf_counter() - t0
                    lane.retire(t_ret)
                    lanes.remove(lane)
                    retired.append(lane)
                    if lane_log is not None:
                        lane_log.append((lane.t_adm, lane.setup, t_ret, len(lane.reqs), lane.cycles))
    release()
    return trace


def one_at_a_time(trace, n_base=50, precision="fp64", cycles=2):
    """The same trace one request after the other in arrival order, each in
    its own context: the plain FP64 solver (FMG, then V-cycles), or the
    mixed-precision steps of a lane."""
    mg = M.default_mg(n_base=n_base)
    t0 = time.perf_counter()
    for r in sorted(trace, key=lambda q: q.arrival):
        now = time.perf_counter() - t0
        if now < r.arrival:
            time.sleep(r.arrival - now)
        r.start = time.perf_counter() - t0
        dt = M.DeviceTree.from_state(r.state)
        if precision == "mixed":
            M.mixed_begin(dt, mg)
        else:
            M.mg_fas_fmg(dt, mg, True, False)
        k = 0
        while True:
            res = float(dt.forest_max_abs(3)[0])
            if res <= r.tol * r.rhs_norm or k >= MAX_STEPS:
                break
            if precision == "mixed":
                M.mixed_correct(dt, mg, cycles)
            else:
                M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
            k += 1
        r.cycles, r.residual, r.phi = k, res / r.rhs_norm, dt.download(1)
        dt.close()
        r.finish = time.perf_counter() - t0
    return trace


def summary(trace):
    lat = np.array([r.finish - r.arrival for r in trace])
    span = max(r.finish for r in trace) - min(r.arrival for r in trace)
    return {"solves_per_s": len(trace) / span, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)), "mean_ms": 1e3 * float(lat.mean()),
            "max_rel_residual_over_tol": float(max(r.residual / r.tol for r in trace))}
