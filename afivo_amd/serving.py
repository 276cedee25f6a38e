"""Serving: Poisson-solve requests arriving over time, solved on one GPU with
continuous batching at request granularity.

A request is one tree (2D or 3D) with its right-hand side and a tolerance on
the relative residual max|rhs - A phi| / max|rhs| over leaf interiors.

The server keeps a set of *lanes*.  A lane is one forest context
(afg_set_forest: its own device pools, tables and HIP stream) whose member
trees, all of one request class (dimension, n_cell and, with `big_cells`,
small or large), run through the same multigrid stages together -- every
stage one launch over the boxes of a level in all member trees.

* Admission: while fewer than `max_lanes` lanes run and the device bytes of
  the lanes (the exact allocations of each context) plus the new forest fit
  the HBM budget, the waiting requests of the oldest request's class are
  admitted together as a new lane, set up on host threads (forest tables,
  uploads, the first residual) while the main thread steps the running
  lanes.
* Steps: lanes advance in turn, one step each: one mixed-precision
  correction (FP64 defect correction, `cycles` FP32 FAS V-cycles from
  e = 0; afg_mixed_first_fmg(0), so that every correction a tree receives is
  the same whenever it joined) or one FP64 V-cycle.  Launches are
  asynchronous, so one lane's work runs while the host checks another's
  per-tree residual norms (which waits for that lane's stream only).
* Release: a member that meets its tolerance leaves at once -- its phi is
  copied back and its latency ends there, whatever its lane-mates still need.
* Compaction: when the members that left hold at least `compact_at` of a
  lane's boxes, the forest is rebuilt in place with afg_update_forest: the
  live members keep phi, rhs and tmp on the device, and waiting requests of
  the class that fit join it (their rhs uploaded; in mixed precision
  afg_mixed_begin with have_guess forms their residual and leaves the live
  members' state bit for bit as it was).

A tree goes through exactly the operations it would go through alone and
stops at the same residual test, so each request's phi is bit-identical to
the same request solved alone (`solve_alone`): a tree in a forest gets
exactly its solo results (links never cross trees, and each tree's stages
stop at its own highest level).

FP64 lanes run plain FAS V-cycles from phi = 0 (mg$D_fas_vcycle,
m_aX_mg.f90:208-255): a tree can join a running forest at any step, where an
FMG would have to run for every member at once.
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


def retime(trace, rate, seed):
    """The same requests with new Poisson arrival times at `rate` per second
    (results cleared), for a load sweep over one set of trees."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=len(trace)))
    for r, a in zip(sorted(trace, key=lambda q: q.rid), t):
        r.arrival, r.start, r.finish, r.cycles, r.residual, r.phi = float(a), -1.0, -1.0, -1, float("nan"), None
    return trace


MAX_STEPS = 60


def forest_bytes(states, precision="mixed", n_var=3):
    """Device bytes a forest of these trees takes: every pool the solver
    allocates per box (the variables, the second phi pool of the fused pass
    pairs, in 3D the colour-split rhs and the colour-split working copy of
    phi; in mixed precision the FP32 copies of phi, rhs and tmp and their
    second phi pool and colour-split rhs and phi), the topology tables,
    per-tree and fixed parts for norms and the bottom kernel's tables, and
    the allocator's rounding of each of the context's fourteen large
    allocations to 2 MiB.  An upper bound of afg_device_bytes and of the
    device memory the forest takes."""
    st = states[0]
    dim, nc = st.dim, st.nc
    stride = -(-(nc + 2) ** dim // 16) * 16
    interior = nc ** dim
    per_box = (n_var + 1) * stride * 8 + (2 * interior * 8 if dim == 3 else 0)
    if precision == "mixed":
        per_box += 4 * stride * 4 + (2 * interior * 4 if dim == 3 else 0)
    per_box += (2 * dim + 8) * 4 + 2          # neighbours, parent, child offset, lists, tags
    boxes = sum(s.n_boxes for s in states)
    return boxes * per_box + len(states) * 4096 + (128 << 10) + 14 * (2 << 20)


def combine_records(states):
    """The records of a forest of `states` (afg_set_forest), tree_of per
    record and each tree's ids in the forest, in record order."""
    D = states[0].dim
    c_par, c_ch, c_nb = 2 + D, 3 + D, 3 + D + (1 << D)
    recs, tree_of, ids, off = [], [], [], 0
    for t, st in enumerate(states):
        r = st.recs.astype(np.int64).copy()
        r[:, 0] += off
        for cols in (slice(c_par, c_par + 1), slice(c_ch, c_nb), slice(c_nb, c_nb + 2 * D)):
            block = r[:, cols]
            block[block > 0] += off
            r[:, cols] = block
        recs.append(r)
        tree_of.append(np.full(st.n_boxes, t, dtype=np.int32))
        ids.append(np.ascontiguousarray(r[:, 0], dtype=np.int32))
        off += int(st.ids.max())
    if off >= 2 ** 31:
        raise ValueError("too many boxes for int32 ids")
    return np.concatenate(recs).astype(np.int32), np.concatenate(tree_of), ids


class Lane:
    """One forest context of a request class: the member trees run every
    step together; members leave as soon as they converge, and the forest
    is compacted (and takes in waiting requests) when enough of it has
    left."""

    def __init__(self, key, mg, precision, cycles, use_graphs):
        self.key, self.mg, self.mixed, self.n_cyc = key, mg, precision == "mixed", cycles
        self.precision, self.use_graphs = precision, use_graphs
        self.dt = None
        self.members = []       # Requests, in tree order
        self.live = []          # member still being solved
        self.steps = []         # steps done per member
        self.ids = []           # each member's ids in the forest
        self.boxes = []         # each member's box count
        self.rebuilds = 0

    def device_bytes(self):
        return self.dt.device_bytes() if self.dt is not None else 0

    def live_members(self):
        return [r for r, a in zip(self.members, self.live) if a]

    def dead_fraction(self):
        tot = sum(self.boxes)
        return sum(b for b, a in zip(self.boxes, self.live) if not a) / tot if tot else 0.0

    def rebuild(self, keep, add):
        """The forest of members `keep` (live members, in their current
        order: they keep their state on the device) and `add` (new: their
        rhs is uploaded); with neither, the context is closed."""
        old_index = {id(r): k for k, r in enumerate(self.members)}
        old_steps = {id(r): s for r, s in zip(self.members, self.steps)}
        members = keep + add
        if not members:
            if self.dt is not None:
                self.dt.close()
            self.dt, self.members, self.live, self.steps, self.ids, self.boxes = None, [], [], [], [], []
            return
        states = [r.state for r in members]
        recs, tree_of, ids = combine_records(states)
        st0 = states[0]
        if self.dt is None:
            self.dt = M.DeviceTree(st0.dim, st0.nc, 3, recs, st0.dr_base, tree_of)
            self.dt.set_graphs(self.use_graphs)
            self.dt.mixed_first_fmg(False)
        else:
            old = [old_index[id(r)] for r in keep] + [-1] * len(add)
            self.dt.update_forest(recs, tree_of, old, st0.dr_base)
        self.rebuilds += 1
        for r, tids in zip(add, ids[len(keep):]):
            self.dt.upload(2, r.state.data[1], tids)
        if self.mixed:
            M.mixed_begin(self.dt, self.mg, have_guess=True)
        self.members, self.ids = members, ids
        self.live = [True] * len(members)
        self.steps = [old_steps.get(id(r), 0) for r in members]
        self.boxes = [r.state.n_boxes for r in members]

    def check(self):
        """Per-member residual norms of the last step (waits for this lane's
        stream); the live members whose residual is valid and meets their
        tolerance, or that ran out of steps."""
        if not any(self.live):
            return []
        res = self.dt.forest_max_abs(3)
        done = []
        for k, r in enumerate(self.members):
            valid = self.mixed or self.steps[k] > 0
            if self.live[k] and valid and (res[k] <= r.tol * r.rhs_norm or self.steps[k] >= MAX_STEPS):
                r.residual = float(res[k]) / r.rhs_norm
                done.append(k)
        return done

    def release(self, done, now):
        """Members `done` leave now: phi is copied back and their latency
        ends, whatever their lane-mates still need."""
        for k in done:
            r = self.members[k]
            r.phi = self.dt.download(1, self.ids[k])
            r.finish, r.cycles = now, self.steps[k]
            self.live[k] = False

    def step(self):
        """Launch the next step of every member (asynchronous)."""
        if self.mixed:
            M.mixed_correct(self.dt, self.mg, self.n_cyc)
        else:
            M.mg_fas_vcycle(self.dt, self.mg, self.dt.highest_lvl, True)
        self.steps = [s + 1 for s in self.steps]


def request_class(r, big_cells):
    return (r.state.dim, r.state.nc, bool(big_cells) and r.leaf_cells >= big_cells)


def serve(trace, hbm_budget=64 << 30, max_lanes=6, max_trees=64, n_base=50, precision="mixed", cycles=2,
          big_cells=0, use_graphs=True, compact_at=0.5, builders=3, log=None):
    """Run the trace through the server; returns the requests with
    start/finish times (seconds after the trace start), their step counts,
    residuals and phi.

    Waiting requests of one class (the oldest request's) are admitted
    together as a new lane while fewer than `max_lanes` run and the HBM
    budget allows; `builders` host threads set lanes up (forest tables,
    uploads, the first residual) while the main thread steps the running
    lanes in turn.  A member leaves its lane as soon as it converges.  When
    the members that left hold at least `compact_at` of a lane's boxes, the
    lane is rebuilt in place (afg_update_forest) with its live members and
    the waiting requests of its class that fit.  `log`, if given, receives
    counts of lanes, rebuilds and steps."""
    import concurrent.futures as cf

    mg = M.default_mg(n_base=n_base)
    pending = collections.deque(sorted(trace, key=lambda r: r.arrival))
    waiting, lanes, building, retired = [], [], [], []
    counts = collections.Counter()
    t0 = time.perf_counter()

    def used_bytes():
        return (sum(l.device_bytes() for l in lanes + retired) +
                sum(forest_bytes([r.state for r in g], precision) for g, _ in building))

    def release_retired():
        # freeing device memory waits for all the device's work: done only
        # while nothing runs, or when the memory is needed
        for lane in retired:
            lane.rebuild([], [])
        retired.clear()

    def build(group):
        lane = Lane(request_class(group[0], big_cells), mg, precision, cycles, use_graphs)
        lane.rebuild([], group)
        return lane

    def fit(cands, base_states, room):
        out = []
        for r in cands:
            if len(base_states) + len(out) >= max_trees:
                break
            if forest_bytes(base_states + [q.state for q in out] + [r.state], precision) > room:
                break
            out.append(r)
        return out

    with cf.ThreadPoolExecutor(max_workers=builders) as pool:
        while pending or waiting or lanes or building:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            if not lanes and not building:
                release_retired()
            # admission of new lanes, oldest waiting request first
            while waiting and len(lanes) + len(building) < max_lanes:
                key = request_class(waiting[0], big_cells)
                cands = [r for r in waiting if request_class(r, big_cells) == key]
                group = fit(cands, [], hbm_budget - used_bytes())
                if not group and retired:
                    if building:
                        break                       # release once the builders are done
                    release_retired()
                    group = fit(cands, [], hbm_budget - used_bytes())
                if not group:
                    if not lanes and not building:
                        raise RuntimeError("request %d does not fit the HBM budget" % waiting[0].rid)
                    break                           # wait for memory to free up
                for r in group:
                    waiting.remove(r)
                    r.start = time.perf_counter() - t0
                building.append((group, pool.submit(build, group)))
                counts["lanes"] += 1
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
                done = lane.check()
                lane.release(done, time.perf_counter() - t0)
                if not any(lane.live):
                    lanes.remove(lane)
                    retired.append(lane)
                    continue
                if done and lane.dead_fraction() >= compact_at:
                    keep = lane.live_members()
                    cands = [r for r in waiting if request_class(r, big_cells) == lane.key]
                    room = hbm_budget - used_bytes()     # the old forest stays allocated meanwhile
                    add = fit(cands, [r.state for r in keep], room)
                    for r in add:
                        waiting.remove(r)
                        r.start = time.perf_counter() - t0
                    lane.rebuild(keep, add)
                    counts["compactions"] += 1
                    counts["joined"] += len(add)
                lane.step()
                counts["steps"] += 1
    release_retired()
    if log is not None:
        log.update(counts)
        log["rebuilds"] = counts["lanes"] + counts["compactions"]
    return trace


def solve_alone(r, mg, precision="fp64", cycles=2, use_graphs=True):
    """One request in its own context through the operations a pool gives
    it: mixed-precision corrections (afg_mixed_begin, then V-cycle
    corrections until the residual test passes) or FP64 V-cycles from
    phi = 0.  Sets r.cycles, r.residual and r.phi."""
    dt = M.DeviceTree.from_state(r.state)
    dt.set_graphs(use_graphs)
    mixed = precision == "mixed"
    if mixed:
        dt.mixed_first_fmg(False)
        M.mixed_begin(dt, mg, have_guess=True)
    k = 0
    while True:
        if mixed or k > 0:
            res = float(dt.forest_max_abs(3)[0])
            if res <= r.tol * r.rhs_norm or k >= MAX_STEPS:
                break
        if mixed:
            M.mixed_correct(dt, mg, cycles)
        else:
            M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
        k += 1
    r.cycles, r.residual, r.phi = k, res / r.rhs_norm, dt.download(1)
    dt.close()
    return r


def one_at_a_time(trace, n_base=50, precision="fp64", cycles=2, fmg=None):
    """The same trace one request after the other in arrival order, each in
    its own context.  fmg=False (the default for the served precisions):
    the operations a pool gives the request (`solve_alone`); fmg=True, FP64
    only: the plain FP64 solver, FMG and then V-cycles to the tolerance."""
    mg = M.default_mg(n_base=n_base)
    fmg = False if fmg is None else fmg
    t0 = time.perf_counter()
    for r in sorted(trace, key=lambda q: q.arrival):
        now = time.perf_counter() - t0
        if now < r.arrival:
            time.sleep(r.arrival - now)
        r.start = time.perf_counter() - t0
        if fmg:
            dt = M.DeviceTree.from_state(r.state)
            M.mg_fas_fmg(dt, mg, True, False)
            k = 0
            while True:
                res = float(dt.forest_max_abs(3)[0])
                if res <= r.tol * r.rhs_norm or k >= MAX_STEPS:
                    break
                M.mg_fas_vcycle(dt, mg, dt.highest_lvl, True)
                k += 1
            r.cycles, r.residual, r.phi = k, res / r.rhs_norm, dt.download(1)
            dt.close()
        else:
            solve_alone(r, mg, precision, cycles)
        r.finish = time.perf_counter() - t0
    return trace


def summary(trace):
    lat = np.array([r.finish - r.arrival for r in trace])
    span = max(r.finish for r in trace) - min(r.arrival for r in trace)
    return {"solves_per_s": len(trace) / span, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)), "mean_ms": 1e3 * float(lat.mean()),
            "max_rel_residual_over_tol": float(max(r.residual / r.tol for r in trace))}
