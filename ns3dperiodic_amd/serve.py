"""Continuous-batching request server for mixed-size solves on one GPU.

A request is a whole solve on its own grid: ``integrate`` advances a seeded
Kolmogorov-forced field to its horizon T with RK3_SSP
(Source/RK_time_step.cu:310-327); ``jv`` carries a perturbation along the
trajectory with the tangent map RK3_SSP_UV (Source/RK_time_step.cu:342-355),
i.e. v(T) = dPhi_T(u0) v0, the product the reference's Krylov drivers build
from finite differences of the time-T map; ``return_map`` steps until the
velocity at a probe point returns through the plane through its initial
value, normal to its initial rate of change (find_intersection,
Source/return_map.cu:425-480), or until T.  Return-map requests probe after
every step, so they advance one step per loop turn with a host
synchronisation; the other kinds stream chunks.

The server keeps every admitted request's state resident in HBM and admits
new requests while others run, as soon as a slot and the HBM budget allow.
Each slot owns a context and a stream; its steps are issued in chunks of K
RK3 steps replayed from a captured HIP graph, at most two chunks in flight
per slot, so the host loop keeps every active stream fed and the device
overlaps the kernels of all active requests.  A request's arithmetic is the
same kernels in the same order whether it runs alone or in a mix, so its
result is bitwise the same either way.

Requests run in the physical mode (``ipq=False``: NS3D_FLAG_NO_IPQ and
NS3D_FLAG_MASK_NYQUIST).  Zeroing Re(Q^) (Source/advection_2_3.cu:35-70) is
exact only for fields with purely imaginary coefficients (odd symmetry, the
reference's Kolmogorov states), and the reference's mask, normalised by
(Nx, Ny, Mz) (Source/math_support.cu:471-519), leaves the x and y axes
untruncated; on the stream's generic random fields at Re up to 500 on grids
down to 32^3 either one lets aliasing or the IPQ rule feed energy in and
solves diverge.  With both off the discrete nonlinear term conserves energy.

Scheduling (``Ledger``, ``poisson_stream``, ``choose_dt``) is plain Python
without GPU calls; ``Server`` and ``one_at_a_time`` drive the HIP library.
"""
import collections
import dataclasses
import math
import time

import numpy as np

from . import fields

GRID_SIZES = (32, 48, 64, 96, 128, 192)


@dataclasses.dataclass
class Request:
    rid: int
    kind: str          # "integrate" | "jv" | "return_map"
    N: int             # cubic grid N^3 on [0, 2 pi]^3
    Re: float
    T: float           # horizon
    seed: int
    arrival: float     # seconds after the stream starts
    nsteps: int = 0
    dt: float = 0.0


@dataclasses.dataclass
class Result:
    rid: int
    admitted: float
    done: float
    latency: float
    u: list            # final state, host complex128 [N, N, N/2+1] x 3
    v: list = None     # final perturbation (jv)
    info: dict = None  # return_map: found, time, point, steps, bisections


def choose_dt(u_hat, T, N, cfl=0.2, L=2 * math.pi, fmax=0.0):
    """dt = cfl * dx / U, shrunk so that T/dt is an integer; returns (nsteps, dt).
    U = max|u0| + T*fmax bounds the velocity over the horizon (forcing of
    amplitude fmax can raise it by at most T*fmax), so the CFL number stays
    <= cfl for the whole solve.  max|u0| is the largest velocity component in
    physical space (as calc_dt uses, Source/cuda_supp.cu:170-238)."""
    umax = max(float(np.max(np.abs(np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))))) for c in u_hat)
    U = umax + T * fmax
    dt0 = cfl * (L / N) / max(U, 1e-30)
    n = max(1, int(math.ceil(T / dt0 - 1e-12)))
    return n, T / n


def poisson_stream(n, rate, seed=0, sizes=GRID_SIZES, re_range=(50.0, 500.0), t_range=(0.5, 5.0),
                   jv_fraction=0.25, jv_T=1.0, rm_fraction=0.0):
    """Seeded request stream: exponential inter-arrival times at `rate` per
    second, N uniform over `sizes`, Re ~ U(re_range), T ~ U(t_range) for
    integrate requests and T = jv_T for J.v and return-map requests (a
    return-map request stops at its crossing or at T)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate)) if rate > 0 else 0.0
        x = rng.random()
        kind = "jv" if x < jv_fraction else ("return_map" if x < jv_fraction + rm_fraction else "integrate")
        N = int(sizes[int(rng.integers(len(sizes)))])
        Re = float(rng.uniform(*re_range))
        T = float(rng.uniform(*t_range)) if kind == "integrate" else float(jv_T)
        out.append(Request(i, kind, N, Re, T, seed=1000 * seed + i, arrival=t))
    return out


def request_inputs(req, energy=0.5):
    """Host inputs of a request: a seeded solenoidal field
    (E(k) ~ k^4 exp(-2 (k/4)^2), energy 0.5), the forcing f_x = sin 4y, and for
    J.v a unit-norm solenoidal perturbation.  Fills req.nsteps / req.dt."""
    N = req.N
    u = fields.random_solenoidal(N, seed=req.seed, energy=energy)
    fx = fields.kolmogorov_forcing_sin(N, n=4)[0]
    v = None
    if req.kind == "jv":
        v = fields.random_solenoidal(N, seed=req.seed + 7919, energy=1.0)
        nrm = math.sqrt(sum(float(np.vdot(c, c).real) for c in v))
        v = [c / nrm for c in v]
    req.nsteps, req.dt = choose_dt(u, req.T, N, fmax=1.0)
    return u, fx, v


def _rows(N, a):
    """Modes of one field whose rows of Mz = N/2+1 are padded to a multiple of a."""
    return N * N * ((N // 2 + 1 + a - 1) // a * a)


def field_bytes(N, precision):
    """HBM held by an integrate request on an N^3 grid, in the library's layout:
    3 state fields and the forcing x-component (FP64, reference layout); 9 FP64
    RK stage fields (rows padded to 8 modes); 15 scratch fields in the bulk
    precision (rows padded to 8 modes); for "fp32s" the FP32 state and
    stages of the RK3 loop (9 fields, rows padded to 8) replace the FP64
    stage fields."""
    nc = N * N * (N // 2 + 1)
    bulk, pad = (16, 8) if precision == "fp64" else (8, 8)
    b = 16 * 4 * nc + bulk * 15 * _rows(N, pad)
    # the FP32-state loop keeps its stages in FP32 (no FP64 stage buffer)
    b += 8 * 9 * _rows(N, 8) if precision == "fp32s" else 16 * 9 * _rows(N, 8)
    return b


def request_bytes(req, precision):
    """field_bytes plus, for J.v, the perturbation (3 FP64 fields), the
    tangent map's 6 FP64 stage fields (padded rows) and 15 more scratch
    fields (B(u,v)); for a return map, the saved state (3 FP64 fields)."""
    nc = req.N * req.N * (req.N // 2 + 1)
    b = field_bytes(req.N, precision)
    if req.kind == "return_map":
        b += nc * 16 * 3              # the state saved before each step (bisection restart)
    if req.kind == "jv":
        bulk, pad = (16, 8) if precision == "fp64" else (8, 8)
        b += 16 * 3 * nc + 16 * 6 * _rows(req.N, 8) + bulk * 15 * _rows(req.N, pad)
    return b


class Ledger:
    """HBM budget and slot accounting for admission (no GPU calls)."""

    def __init__(self, budget_bytes, max_active):
        self.budget = int(budget_bytes)
        self.max_active = int(max_active)
        self.used = 0
        self.active = 0
        self.peak = 0
        self.peak_active = 0

    def fits(self, nbytes):
        return self.active < self.max_active and self.used + nbytes <= self.budget

    def admit(self, nbytes):
        if not self.fits(nbytes):
            raise RuntimeError("admission beyond budget or slot limit")
        self.used += nbytes
        self.active += 1
        self.peak = max(self.peak, self.used)
        self.peak_active = max(self.peak_active, self.active)

    def release(self, nbytes):
        self.used -= nbytes
        self.active -= 1
        assert self.used >= 0 and self.active >= 0

    def reserve(self, nbytes):
        """Memory held without a request of its own (a batching group's slots)."""
        self.used += nbytes
        self.peak = max(self.peak, self.used)

    def unreserve(self, nbytes):
        self.used -= nbytes
        assert self.used >= 0


def admit_order(queue, ledger, cost, max_overtakes=2):
    """Requests of the FIFO `queue` to admit now.  Requests are taken in
    order while slots and budget allow; a request that does not fit may be
    overtaken by smaller ones behind it, at most `max_overtakes` times, after
    which nothing behind it is admitted before it (no starvation)."""
    picked = []
    waiting = []                  # earlier requests that did not fit
    used, active = ledger.used, ledger.active
    for r in list(queue):
        if active >= ledger.max_active:
            break
        c = cost(r)
        if used + c > ledger.budget:
            waiting.append(r)
            continue
        if any(getattr(w, "overtaken", 0) >= max_overtakes for w in waiting):
            break
        for w in waiting:
            w.overtaken = getattr(w, "overtaken", 0) + 1
        picked.append(r)
        used += c
        active += 1
    return picked


def summarize(results, t_first_arrival=0.0):
    lat = np.array([r.latency for r in results])
    span = max(r.done for r in results) - t_first_arrival
    return {"solves": len(results), "span_s": span, "solves_per_s": len(results) / span,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99))}


def probe_point(N):
    """Grid point whose velocity a return-map request follows."""
    return (N // 8, N // 4, N // 2)


def _return_map_search(g, u, f, dt):
    from .return_map import CrossingSearch, NSSystem
    sysm = NSSystem(g, u, f, point=probe_point(g.Nx))
    x0 = sysm.probe()
    v = sysm.probe_rhs()
    nrm = float(np.linalg.norm(v))
    normal = v / nrm if nrm > 0 else np.array([1.0, 0.0, 0.0])
    return CrossingSearch(sysm, x0, normal, dt)


# ---------------------------------------------------------------------------- GPU side
class _Slot:
    """A context, a stream and resident buffers for one request at a time;
    reused by later requests of the same size and kind."""

    def __init__(self, N, kind, precision, device, chunk, use_graphs, ipq):
        import torch
        from .solver import NS3D
        self.N, self.kind = N, kind
        self.bytes = request_bytes(Request(-1, kind, N, 1.0, 1.0, 0, 0.0), precision)
        self.stream = torch.cuda.Stream(device=device)
        self.g = NS3D(N, precision=precision, stream=self.stream, device=device, ipq=ipq, mask_nyquist=not ipq)
        cs = (N, N, N // 2 + 1)
        dev = torch.device("cuda", device)
        self.u = [torch.empty(cs, dtype=torch.complex128, device=dev) for _ in range(3)]
        self.v = [torch.empty(cs, dtype=torch.complex128, device=dev) for _ in range(3)] if kind == "jv" else None
        self.f = [torch.empty(cs, dtype=torch.complex128, device=dev), None, None]
        self.chunk = chunk
        self.use_graphs = use_graphs
        self.req = None
        self.done_ev = None

    def load(self, req, io):
        import torch
        self.req, self.io = req, io
        self.g.set_params(req.dt, req.Re)
        with torch.cuda.stream(self.stream):
            for d, h in zip(self.u, io["u"]):
                d.copy_(h, non_blocking=True)
            self.f[0].copy_(io["f"], non_blocking=True)
            if req.kind == "jv":
                for d, h in zip(self.v, io["v"]):
                    d.copy_(h, non_blocking=True)
        self.left = req.nsteps
        self.inflight = collections.deque()
        self.search = None
        self.info = None

    def issue(self):
        """Enqueue the next chunk of at most K steps; False once all are issued."""
        import torch
        if self.left <= 0:
            return False
        k = min(self.left, self.chunk)
        if self.req.kind == "return_map":
            with torch.cuda.stream(self.stream):
                if self.search is None:
                    self.search = _return_map_search(self.g, self.u, self.f, self.req.dt)
                k = 1
                if self.search.advance():
                    self.info = {k2: v for k2, v in self.search.result.items() if k2 != "step_sizes"}
                    k = self.left
                elif self.left == 1:
                    self.info = {k2: v for k2, v in self.search.not_found().items() if k2 != "step_sizes"}
        elif self.req.kind == "jv":
            for _ in range(k):
                self.g.RK3_SSP_UV(self.u, self.v, self.f)
        elif self.use_graphs and k == self.chunk:
            self.g.RK3_SSP_graph(self.u, self.f, steps=k)
        else:
            self.g.RK3_SSP(self.u, self.f, steps=k)
        self.left -= k
        ev = torch.cuda.Event()
        ev.record(self.stream)
        self.inflight.append(ev)
        return True

    def start_download(self):
        import torch
        with torch.cuda.stream(self.stream):
            for h, d in zip(self.io["out_u"], self.u):
                h.copy_(d, non_blocking=True)
            if self.req.kind == "jv":
                for h, d in zip(self.io["out_v"], self.v):
                    h.copy_(d, non_blocking=True)
        self.done_ev = torch.cuda.Event()
        self.done_ev.record(self.stream)


def group_member_bytes(N, precision):
    """HBM of one member slot of a batching group (its share of the batch):
    3 FP64 state fields (reference layout), 6 FP64 RK stage fields (rows padded
    to 8 modes) and 15 scratch fields in the bulk precision (padded rows)."""
    nc = N * N * (N // 2 + 1)
    bulk, pad = (16, 8) if precision == "fp64" else (8, 8)
    return 16 * 3 * nc + 16 * 6 * _rows(N, 8) + bulk * 15 * _rows(N, pad)


class _Group:
    """Resident integrate requests of one grid size advanced together: a batch
    of `cap` member slots on one context and stream, one launch per pass for
    every member (ns3d_rk3_ssp_batch_members, per-member dt and Re).  Each
    chunk of K steps gives every member min(K, steps left) steps; idle slots
    and members that have finished sit the chunk out.  A member's result is
    bitwise what its steps give it alone."""

    def __init__(self, N, cap, precision, device, chunk, use_graphs, ipq, inflight):
        import torch
        from .solver import NS3D
        self.N, self.cap, self.chunk, self.use_graphs = N, cap, chunk, use_graphs
        self.member_bytes = group_member_bytes(N, precision)
        self.bytes = cap * self.member_bytes + 16 * N * N * (N // 2 + 1)     # + the shared forcing
        self.stream = torch.cuda.Stream(device=device)
        self.g = NS3D(N, precision=precision, stream=self.stream, device=device, ipq=ipq, mask_nyquist=not ipq)
        dev = torch.device("cuda", device)
        cs = (N, N, N // 2 + 1)
        self.ub = torch.zeros((cap, 3) + cs, dtype=torch.complex128, device=dev)
        self.f = [torch.zeros(cs, dtype=torch.complex128, device=dev), None, None]
        self.f_loaded = False
        self.dt_d = torch.zeros(cap, dtype=torch.float64, device=dev)
        self.re_d = torch.ones(cap, dtype=torch.float64, device=dev)
        self.left_d = torch.zeros(cap, dtype=torch.int32, device=dev)
        # host copies of the member parameters, uploaded with every chunk from a ring of pinned buffers
        self.dt_h = np.zeros(cap)
        self.re_h = np.ones(cap)
        self.ring = [(torch.zeros(cap, dtype=torch.float64).pin_memory(), torch.zeros(cap, dtype=torch.float64).pin_memory(),
                      torch.zeros(cap, dtype=torch.int32).pin_memory(), None) for _ in range(inflight + 1)]
        self.ring_i = 0
        self.members = [None] * cap           # slot -> dict(req, io, rem, admitted, done_ev)
        self.inflight = collections.deque()

    def free_slot(self):
        return next((m for m in range(self.cap) if self.members[m] is None), None)

    def active(self):
        return sum(1 for m in self.members if m is not None)

    def load(self, m, req, io, admitted):
        import torch
        with torch.cuda.stream(self.stream):
            for b in range(3):
                self.ub[m, b].copy_(io["u"][b], non_blocking=True)
            if not self.f_loaded:             # every request of a grid has the forcing f_x = sin 4y
                self.f[0].copy_(io["f"], non_blocking=True)
                self.f_loaded = True
        self.dt_h[m], self.re_h[m] = req.dt, req.Re
        self.members[m] = {"req": req, "io": io, "rem": req.nsteps, "admitted": admitted, "done_ev": None}

    def issue(self):
        """Enqueue one chunk for every member with steps left; False if none has."""
        import torch
        left = np.zeros(self.cap, dtype=np.int32)
        for m, mem in enumerate(self.members):
            if mem is not None and mem["rem"] > 0:
                left[m] = min(mem["rem"], self.chunk)
        if not left.any():
            return False
        dt_p, re_p, left_p, ev = self.ring[self.ring_i]
        if ev is not None:
            ev.synchronize()                  # the upload that last used this host buffer is done
        dt_p.numpy()[:] = self.dt_h
        re_p.numpy()[:] = self.re_h
        left_p.numpy()[:] = left
        with torch.cuda.stream(self.stream):
            self.dt_d.copy_(dt_p, non_blocking=True)
            self.re_d.copy_(re_p, non_blocking=True)
            self.left_d.copy_(left_p, non_blocking=True)
            up = torch.cuda.Event()
            up.record(self.stream)
            self.ring[self.ring_i] = (dt_p, re_p, left_p, up)
            self.ring_i = (self.ring_i + 1) % len(self.ring)
            self.g.RK3_SSP_batch_members(self.ub, self.f, self.dt_d, self.re_d, self.left_d, self.chunk,
                                         graph=self.use_graphs)
            for m, mem in enumerate(self.members):
                if left[m]:
                    mem["rem"] -= int(left[m])
                    if mem["rem"] == 0:       # its last chunk: download right behind it on this stream
                        for h, b in zip(mem["io"]["out_u"], range(3)):
                            h.copy_(self.ub[m, b], non_blocking=True)
                        mem["done_ev"] = torch.cuda.Event()
                        mem["done_ev"].record(self.stream)
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.inflight.append(ev)
        return True

    def close(self):
        self.g.close()


class Server:
    """Continuous-batching server: up to `max_active` resident requests,
    admitted by HBM budget.  With `batch` (default), integrate requests of one
    grid size share a batching group (one launch per pass for all of them);
    J.v and return-map requests, and every request without `batch`, run on
    slots of their own (context + stream)."""

    def __init__(self, precision="fp32", max_active=8, budget_bytes=64 << 30, chunk=16, use_graphs=True,
                 device=0, inflight=2, ipq=False, batch=True):
        self.precision = precision
        self.ipq = ipq
        self.ledger = Ledger(budget_bytes, max_active)
        self.chunk = chunk
        self.use_graphs = use_graphs
        self.device = device
        self.inflight = inflight
        self.batch = batch and precision in ("fp64", "fp32")
        self.idle = collections.defaultdict(list)     # (N, kind) -> idle slots (context, graph, buffers kept)
        self.idle_bytes = 0
        self.groups = {}                              # N -> _Group

    def _slot(self, r):
        key = (r.N, r.kind)
        if self.idle[key]:
            s = self.idle[key].pop()
            self.idle_bytes -= s.bytes
            return s
        return _Slot(r.N, r.kind, self.precision, self.device, self.chunk, self.use_graphs, self.ipq)

    def _make_room(self, need):
        # drop idle slots (largest first), then empty groups, while the budget is short
        for key in sorted(self.idle, key=lambda k: -k[0]):
            while self.idle[key] and self.ledger.used + self.idle_bytes + need > self.ledger.budget:
                s = self.idle[key].pop()
                s.g.close()
                self.idle_bytes -= s.bytes
        for N in sorted(self.groups, reverse=True):
            g = self.groups[N]
            if g.active() == 0 and self.ledger.used + self.idle_bytes + need > self.ledger.budget:
                g.close()
                self.idle_bytes -= g.bytes
                del self.groups[N]

    def _batched(self, r):
        return self.batch and r.kind == "integrate"

    def _cost(self, r):
        """HBM a request adds when admitted: its own bytes on a slot; on a
        batching group, nothing if the group has members and a free member
        slot, the group's memory if it is idle, else the memory of a new
        one-member group.  A group holds its members' memory from its
        creation: charged to the ledger as a reservation while it has members,
        counted with the idle slots (reclaimable) while it has none."""
        if not self._batched(r):
            return request_bytes(r, self.precision)
        g = self.groups.get(r.N)
        if g is not None and g.free_slot() is not None:
            return 0 if g.active() else g.bytes
        return group_member_bytes(r.N, self.precision) + 16 * r.N * r.N * (r.N // 2 + 1)

    def _admit_batched(self, r, now):
        """Put r in its grid's group (created with as many member slots as
        the budget and the slot limit allow); False if that is not possible now."""
        g = self.groups.get(r.N)
        if g is None:
            mb = group_member_bytes(r.N, self.precision)
            fb = 16 * r.N * r.N * (r.N // 2 + 1)
            self._make_room(mb + fb)
            room = self.ledger.budget - self.ledger.used - self.idle_bytes - fb
            if room < mb:
                return False
            # at most half of the free memory beyond the first member, so other grids keep room
            cap = min(self.ledger.max_active, max(1, (room // 2) // mb))
            g = self.groups[r.N] = _Group(r.N, int(cap), self.precision, self.device, self.chunk, self.use_graphs,
                                          self.ipq, self.inflight)
            self.idle_bytes += g.bytes
        m = g.free_slot()
        if m is None:
            return False
        if g.active() == 0:                   # the group's memory moves from idle to reserved
            self.idle_bytes -= g.bytes
            self.ledger.reserve(g.bytes)
        self.ledger.admit(0)
        g.load(m, r, self._io[r.rid], now)
        return True

    def run(self, requests, io, clock=time.perf_counter):
        """Serve `requests`; io[rid] holds the request's pinned host buffers
        (see pin_io).  Returns {rid: Result}; results stay in io[rid]."""
        import torch
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        active = []
        results = {}
        cost = self._cost
        self._io = io
        for r in requests:
            if request_bytes(r, self.precision) > self.ledger.budget:
                raise ValueError(f"request {r.rid} needs {cost(r)} bytes, more than the whole budget")
            if r.nsteps < 1:
                raise ValueError(f"request {r.rid} has no steps (request_inputs sets them)")
        t0 = clock()
        while pending or queue or active or any(g.active() for g in self.groups.values()):
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            for r in admit_order(queue, self.ledger, cost):
                if self._batched(r):
                    if self._admit_batched(r, clock() - t0):
                        queue.remove(r)
                    continue                  # not possible now (group full or no room): stays queued
                self._make_room(cost(r))
                if not self.ledger.fits(cost(r)):
                    continue                  # a group created in this pass took the room: stays queued
                queue.remove(r)
                self.ledger.admit(cost(r))
                s = self._slot(r)
                s.load(r, io[r.rid])
                s.admitted = clock() - t0
                active.append(s)
            progressed = False
            for g in self.groups.values():
                while g.inflight and g.inflight[0].query():
                    g.inflight.popleft()
                while len(g.inflight) < self.inflight and g.issue():
                    progressed = True
                for m, mem in enumerate(g.members):
                    if mem is not None and mem["done_ev"] is not None and mem["done_ev"].query():
                        t = clock() - t0
                        r = mem["req"]
                        results[r.rid] = Result(r.rid, mem["admitted"], t, t - r.arrival,
                                                [c.numpy() for c in mem["io"]["out_u"]], None, None)
                        g.members[m] = None
                        self.ledger.release(0)
                        if g.active() == 0:   # idle group: its memory becomes reclaimable
                            self.ledger.unreserve(g.bytes)
                            self.idle_bytes += g.bytes
                        progressed = True
            for s in list(active):
                while s.inflight and s.inflight[0].query():
                    s.inflight.popleft()
                while len(s.inflight) < self.inflight and s.issue():
                    progressed = True
                if s.left == 0 and not s.inflight and s.done_ev is None:
                    s.start_download()
                if s.done_ev is not None and s.done_ev.query():
                    t = clock() - t0
                    r = s.req
                    out_v = [c.numpy() for c in s.io["out_v"]] if r.kind == "jv" else None
                    results[r.rid] = Result(r.rid, s.admitted, t, t - r.arrival,
                                            [c.numpy() for c in s.io["out_u"]], out_v, s.info)
                    s.done_ev = None
                    s.req = s.io = None
                    active.remove(s)
                    self.ledger.release(request_bytes(r, self.precision))
                    self.idle[(r.N, r.kind)].append(s)
                    self.idle_bytes += s.bytes
                    progressed = True
            if not progressed:
                time.sleep(20e-6)
        torch.cuda.synchronize(self.device)
        return results

    def close(self):
        for lst in self.idle.values():
            for s in lst:
                s.g.close()
        self.idle.clear()
        for g in self.groups.values():
            g.close()
            if g.active():
                self.ledger.unreserve(g.bytes)
            else:
                self.idle_bytes -= g.bytes
        self.groups.clear()


def pin_io(u, fx, v):
    """Pinned host buffers of one request: its inputs (what a client hands
    over) and room for its outputs."""
    import torch
    pin = lambda a: torch.from_numpy(np.ascontiguousarray(a)).pin_memory()   # noqa: E731
    io = {"u": [pin(c) for c in u], "f": pin(fx), "v": [pin(c) for c in v] if v is not None else None}
    io["out_u"] = [torch.empty_like(c).pin_memory() for c in io["u"]]
    io["out_v"] = [torch.empty_like(c).pin_memory() for c in io["v"]] if v is not None else None
    return io


def one_at_a_time(requests, io, precision="fp64", device=0, clock=time.perf_counter, ctxs=None, ipq=False):
    """Baseline: requests in arrival order, one at a time, eager RK3 steps on
    one context per grid size (contexts reused; pass `ctxs` to keep them
    across calls), upload and download included."""
    import torch
    from .solver import NS3D
    own = ctxs is None
    ctxs = {} if own else ctxs
    results = {}
    t0 = clock()
    dev = torch.device("cuda", device)
    for r in sorted(requests, key=lambda q: q.arrival):
        while clock() - t0 < r.arrival:
            time.sleep(50e-6)
        adm = clock() - t0
        if r.N not in ctxs:
            ctxs[r.N] = NS3D(r.N, precision=precision, device=device, ipq=ipq, mask_nyquist=not ipq)
        g = ctxs[r.N]
        g.set_params(r.dt, r.Re)
        inp = io[r.rid]
        u = [h.to(dev) for h in inp["u"]]
        f = [inp["f"].to(dev), None, None]
        info = None
        if r.kind == "jv":
            v = [h.to(dev) for h in inp["v"]]
            for _ in range(r.nsteps):
                g.RK3_SSP_UV(u, v, f)
        elif r.kind == "return_map":
            v = None
            search = _return_map_search(g, u, f, r.dt)
            found = False
            for _ in range(r.nsteps):
                if search.advance():
                    found = True
                    break
            res = search.result if found else search.not_found()
            info = {k: val for k, val in res.items() if k != "step_sizes"}
        else:
            v = None
            g.RK3_SSP(u, f, steps=r.nsteps)
        out_u = [c.cpu().numpy() for c in u]
        out_v = [c.cpu().numpy() for c in v] if v is not None else None
        t = clock() - t0
        results[r.rid] = Result(r.rid, adm, t, t - r.arrival, out_u, out_v, info)
    if own:
        for g in ctxs.values():
            g.close()
    return results
