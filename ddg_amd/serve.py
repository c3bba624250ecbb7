"""Request loop for Poisson/Helmholtz solves on one GPU (continuous batching).

The reference has no serving layer: each ``solve_sipg_helmholtz_2d`` call
(src/ddg/dg2d/sipg.py:536-612) rebuilds its closures and runs alone.  Here a
stream of requests -- each a Helmholtz problem ``-Lap u + k^2 u = f`` on a
``2 n x n`` mesh of order ``N`` with its own right-hand side -- is served by
a fixed set of worker slots that run solves side by side on the device:

* Every (n, N) problem class keeps its mesh data, float64 and float32
  operators and a float32 hp-multigrid preconditioner resident in HBM
  (:class:`ProblemClass`), built once.
* A solve needs private Krylov state: a ``ddgx_solver`` context (work
  vectors, multigrid cycle buffers), taken from a per-class pool and
  returned afterwards.  Its device bytes plus the request's vectors are the
  request's HBM footprint.
* A dispatcher admits requests in arrival order as soon as a slot is idle
  and the footprint fits the HBM budget (:class:`Admission`); requests that
  arrive while others are on the GPU join immediately, so their kernels
  overlap on the device (each slot has its own HIP stream).
* Requests of one class that have arrived by the time a slot frees up run
  as one batch in that slot (up to ``max_batch``): one kernel launch per
  step serves all of them (``ddgx_batch``, :class:`ddg_amd.operator.
  BatchSolver`), so small problems stop paying one launch latency each.
* Each solve is float64 CG preconditioned by the class's float32 V-cycle
  ("pcg32", batched), or mixed-precision refinement ("refine": float32
  MG-PCG steps in a float64 residual loop, ``ddgx_solver_ir``, one request
  per slot), to a relative residual ``tol``.

A solve reads only its own vectors and the read-only class data, and every
reduction has a fixed order per request, so a request's result does not
depend on what else is in flight, nor on the batch it runs in: alone and in
a mix it is bitwise the same.

The scheduling logic (:class:`Admission`, :func:`poisson_arrivals`) is pure
Python and needs no GPU.
"""

from __future__ import annotations

import ctypes as C
import sys
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib


# ---------------------------------------------------------------------------
# requests and arrivals
# ---------------------------------------------------------------------------
@dataclass
class Request:
    rid: int
    n: int              # mesh: 2 n^2 triangles of the unit square
    N: int              # polynomial order
    k2: float           # Helmholtz coefficient k^2 (alpha of alpha M + A)
    seed: int           # seed of the Fourier right-hand side
    arrival: float = 0.0  # seconds after the start of the stream


@dataclass
class Result:
    rid: int
    arrival: float
    start: float        # admitted (seconds after the start of the stream)
    end: float          # solution ready on the host side
    outer: int
    iterations: int
    converged: bool     # float64 residual met the tolerance
    residual2: float
    atol2: float
    at_floor: bool = False  # stopped at the float64 residual's round-off floor
    x: object = None    # device solution (kept only when asked for)

    @property
    def latency(self) -> float:
        return self.end - self.arrival


def make_requests(count, seed=0, sizes=(16, 32, 64, 128), orders=(2, 3, 4, 5, 6),
                  k2_range=(0.0, 50.0)):
    """``count`` requests with n, N uniform over ``sizes`` x ``orders`` and
    k^2 uniform in ``k2_range`` (seeded), arrival times unset."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        out.append(Request(rid=i, n=int(rng.choice(sizes)), N=int(rng.choice(orders)),
                           k2=float(rng.uniform(*k2_range)), seed=int(rng.integers(2 ** 31))))
    return out


def poisson_arrivals(requests, rate, seed=0):
    """Set arrival times of a Poisson process of ``rate`` requests per second
    (exponential gaps, seeded); ``rate=None`` puts them all at t = 0."""
    rng = np.random.default_rng(seed)
    t = 0.0
    for r in requests:
        if rate is not None:
            t += float(rng.exponential(1.0 / rate))
        r.arrival = t
    return requests


# ---------------------------------------------------------------------------
# admission control
# ---------------------------------------------------------------------------
class Admission:
    """Slots and HBM budget.  ``try_admit(bytes)`` succeeds when a slot is
    idle and the bytes fit; ``release(bytes)`` returns both.  Thread-safe."""

    def __init__(self, slots: int, budget_bytes: int):
        if slots < 1:
            raise ValueError("need at least one slot")
        if budget_bytes <= 0:
            raise ValueError("HBM budget must be positive")
        self.slots = int(slots)
        self.budget = int(budget_bytes)
        self.busy = 0
        self.used = 0
        self.peak_used = 0
        self.peak_busy = 0
        self._lock = threading.Lock()

    def fits_ever(self, nbytes: int) -> bool:
        return nbytes <= self.budget

    def try_admit(self, nbytes: int) -> bool:
        with self._lock:
            if self.busy >= self.slots or self.used + nbytes > self.budget:
                return False
            self.busy += 1
            self.used += nbytes
            self.peak_used = max(self.peak_used, self.used)
            self.peak_busy = max(self.peak_busy, self.busy)
            return True

    def release(self, nbytes: int) -> None:
        with self._lock:
            if self.busy < 1 or nbytes > self.used:
                raise RuntimeError("release without a matching admission")
            self.busy -= 1
            self.used -= nbytes


def dispatch_groups(requests, admission: Admission, footprint, start_fn, wait_fn, clock,
                    key=None, max_group=1):
    """Admission loop shared by the GPU server and its CPU tests.

    The earliest arrived request not yet started (the head) is admitted as
    soon as ``clock() >= arrival``, a slot is idle and its group's bytes fit:
    the group is the head and, up to ``max_group`` in all, the next arrived
    requests with the head's ``key(req)``, in arrival order (a group that
    would not fit the budget is cut down); ``max_group`` is a number or a
    function of the key.  ``footprint(group)`` gives the
    bytes, ``start_fn(group, nbytes, t_admit)`` hands the group to a worker
    (which calls ``admission.release(nbytes)`` when done), and
    ``wait_fn(timeout)`` blocks until a worker finishes or ``timeout``
    seconds pass.  Returns when every request has been started."""
    key = (lambda r: r.rid) if key is None else key
    future = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
    for r in future:
        if not admission.fits_ever(footprint([r])):
            raise ValueError(f"request {r.rid} needs {footprint([r])} bytes, more than the "
                             f"HBM budget of {admission.budget}")
    arrived = deque()   # arrived and not started, in arrival order
    by_key = {}         # key -> the same, per key
    started = set()
    while future or arrived:
        now = clock()
        while future and future[0].arrival <= now:
            r = future.popleft()
            arrived.append(r)
            by_key.setdefault(key(r), deque()).append(r)
        while arrived and arrived[0].rid in started:
            arrived.popleft()
        if not arrived:
            if not future:
                break
            wait_fn(future[0].arrival - now)
            continue
        k = key(arrived[0])
        q = by_key[k]      # q[0] is the head
        mg = max_group(k) if callable(max_group) else max_group
        group = [q[i] for i in range(min(max(1, mg), len(q)))]
        while len(group) > 1 and not admission.fits_ever(footprint(group)):
            group.pop()
        nb = footprint(group)
        if admission.try_admit(nb):
            for r in group:
                q.popleft()
                started.add(r.rid)
            start_fn(group, nb, now)
        else:
            wait_fn(0.05)


def dispatch(requests, admission: Admission, footprint, start_fn, wait_fn, clock):
    """:func:`dispatch_groups` one request at a time: ``footprint(req)``,
    ``start_fn(req, nbytes, t_admit)``."""
    dispatch_groups(requests, admission, lambda g: footprint(g[0]),
                    lambda g, nb, t: start_fn(g[0], nb, t), wait_fn, clock)


# ---------------------------------------------------------------------------
# right-hand sides
# ---------------------------------------------------------------------------
def fourier_modes(seed, modes=64, kmax=8):
    """Seeded random Fourier series: wave vectors k (integer, |k_i| <= kmax,
    k != 0), amplitudes a ~ N(0,1) |k|^-2, phases ~ U(0, 2 pi)."""
    rng = np.random.default_rng(seed)
    k = np.zeros((modes, 2))
    for m in range(modes):
        while not k[m].any():
            k[m] = rng.integers(-kmax, kmax + 1, size=2)
    a = rng.standard_normal(modes) / np.sum(k * k, axis=1)
    phi = rng.uniform(0.0, 2.0 * np.pi, modes)
    return k, a, phi


def fourier_field(x, y, seed, modes=64, chunk=8):
    """``sum_m a_m cos(2 pi k_m . (x, y) + phi_m)`` at device nodes x, y."""
    import torch
    k, a, phi = fourier_modes(seed, modes)
    f = torch.zeros_like(x)
    for c in range(0, modes, chunk):
        kx = torch.as_tensor(k[c:c + chunk, 0], device=x.device, dtype=x.dtype)[:, None]
        ky = torch.as_tensor(k[c:c + chunk, 1], device=x.device, dtype=x.dtype)[:, None]
        aa = torch.as_tensor(a[c:c + chunk], device=x.device, dtype=x.dtype)[:, None]
        pp = torch.as_tensor(phi[c:c + chunk], device=x.device, dtype=x.dtype)[:, None]
        f += (aa * torch.cos(2.0 * np.pi * (kx * x[None] + ky * y[None]) + pp)).sum(0)
    return f


# ---------------------------------------------------------------------------
# resident problem classes
# ---------------------------------------------------------------------------
def mg_degrees(N):
    """p-levels of the hp-multigrid: N, N/2, 1 (every degree for N < 4)."""
    return [N, N // 2, 1] if N >= 4 else list(range(N, 0, -1))


class ProblemClass:
    """Everything of an (n, N) class that requests share: the forest-ordered
    2 n^2 mesh with homogeneous Dirichlet walls, float64/float32 operators,
    a float32 hp-multigrid (h-levels down to 2 x n_base^2), node coordinates
    on the device, and pools of idle solve contexts: batches of each size
    (method "pcg32") or solver contexts (method "refine")."""

    def __init__(self, n, N, n_base=16, tau_scale=200.0, method="pcg32"):
        import torch
        if method not in ("refine", "pcg32"):
            raise ValueError(f"unknown solve method {method!r} (refine, pcg32)")
        self.method = method
        from . import mesh as pm
        from . import multigrid as MG
        from .sipg import get_operator
        if n & (n - 1) or n < 2:
            raise ValueError("n must be a power of two")
        self.n, self.N = int(n), int(N)
        nb = min(n_base, n)
        lb = int(round(np.log2(nb))) - 1
        lv = int(round(np.log2(n))) - 1
        VX, VY, E = pm.forest_rectangle_hierarchy(2, lb)[-1]
        base = pm.build_mesh_2d(VX, VY, E, N)
        self.pc, self.mesh = MG.make_sipg_hpmg_preconditioner(
            base, lv - lb, bc_fn=pm.all_dirichlet, tau_scale=tau_scale,
            degrees=mg_degrees(N), dtype=torch.float32)
        self.bc = pm.all_dirichlet(self.mesh)
        self.op64 = get_operator(self.mesh, self.bc, tau_scale)
        self.op32 = get_operator(self.mesh, self.bc, tau_scale, dtype=torch.float32)
        self.K, self.Np = self.mesh.K, self.mesh.Np
        dev = self.op64.device
        self.x = torch.as_tensor(np.ascontiguousarray(self.mesh.x.T).reshape(-1), device=dev)
        self.y = torch.as_tensor(np.ascontiguousarray(self.mesh.y.T).reshape(-1), device=dev)
        self._idle = []      # refinement solver contexts
        self._batches = {}   # batch size -> idle BatchSolvers
        self._lock = threading.Lock()
        self._all = []
        if method == "pcg32":
            probe = self._new_batch(1)
            self._batches[1] = [probe]
        else:
            probe = self._new_solver()
            self._idle.append(probe)
        self.solver_bytes = probe.nbytes
        # per request: its solve context (a batch's device copies of the
        # operator and multigrid and its Krylov state grow by this much per
        # request) plus f, b and x (float64)
        self.footprint = self.solver_bytes + 3 * self.K * self.Np * 8

    def group_footprint(self, count):
        return count * self.footprint

    def _new_solver(self):
        s = SolverContext(self.op64, self.op32, self.pc)
        with self._lock:
            self._all.append(s)
        return s

    def _new_batch(self, count):
        from .operator import BatchSolver
        b = BatchSolver(self.op64, self.pc, count)
        with self._lock:
            self._all.append(b)
        return b

    def acquire(self):
        with self._lock:
            if self._idle:
                return self._idle.pop()
        return self._new_solver()

    def release(self, s):
        with self._lock:
            self._idle.append(s)

    def acquire_batch(self, count):
        with self._lock:
            pool = self._batches.setdefault(count, [])
            if pool:
                return pool.pop()
        return self._new_batch(count)

    def release_batch(self, b):
        with self._lock:
            self._batches.setdefault(b.nseg, []).append(b)

    def close(self):
        for s in self._all:
            s.close()
        self._all, self._idle = [], []


class _CGAsIR:
    """A CG solve's info under the refinement info's field names (one
    "outer" step; a float64 CG has no round-off floor rule)."""

    def __init__(self, info):
        self.outer = 1
        self.inner_total = info.iterations
        self.converged = info.converged
        self.residual2 = info.residual2
        self.atol2 = info.atol2
        self.at_floor = 0


class SolverContext:
    """A ``ddgx_solver`` context: private Krylov state for solves with one
    operator and preconditioner.  ``op32`` given: a refinement context
    (float64 ``op``, float32 twin, float32 ``pc``), used by :meth:`solve`;
    else a PCG context on ``op`` (:meth:`pcg`).  On a non-default stream the
    context replays captured chunks of CG iterations as HIP graphs."""

    def __init__(self, op, op32=None, pc=None):
        L = _lib.lib()
        h = C.c_void_p()
        _lib.check(L.ddgx_solver_create(op.handle, None if op32 is None else op32.handle,
                                        None if pc is None else pc.handle, C.byref(h)),
                   "ddgx_solver_create")
        self._h = h
        nb = C.c_int64()
        _lib.check(L.ddgx_solver_bytes(h, C.byref(nb)), "ddgx_solver_bytes")
        self.nbytes = int(nb.value)
        self._keep = (op, op32, pc)

    def solve(self, b, x, alpha, cfg, stream):
        """Mixed-precision refinement (``ddgx_solver_ir``); returns IRInfo."""
        info = _lib.IRInfo()
        _lib.check(_lib.lib().ddgx_solver_ir(self._h, float(alpha), 0.0, C.c_void_p(b.data_ptr()),
                                             C.c_void_p(x.data_ptr()), C.byref(cfg),
                                             C.byref(info), C.c_void_p(stream)), "ddgx_solver_ir")
        return info

    def pcg(self, b, x, alpha, cfg, stream, shift=0.0):
        """Preconditioned CG (``ddgx_solver_pcg``); returns CGInfo."""
        info = _lib.CGInfo()
        _lib.check(_lib.lib().ddgx_solver_pcg(self._h, float(alpha), float(shift),
                                              C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                              C.byref(cfg), C.byref(info), C.c_void_p(stream)),
                   "ddgx_solver_pcg")
        return info

    def close(self):
        if self._h:
            _lib.lib().ddgx_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# the server
# ---------------------------------------------------------------------------
@dataclass
class ServeConfig:
    # "pcg32": float64 CG preconditioned by the float32 V-cycle (the faster:
    # 1.4x the solves/s on the mixed stream, every request converged);
    # "refine": mixed-precision iterative refinement (float32 MG-PCG steps in
    # a float64 residual loop).  Both stop at tol on the float64 residual.
    method: str = "pcg32"
    # solves in flight at once (one HIP stream each).  The stream's small
    # kernels leave most of the GPU idle, so more solves side by side raise
    # the rate: on one MI355X (four hardware queues per process) the mixed
    # stream's closed-loop capacity is about 71/s with 4 slots, 87/s with
    # 8 and 84/s with 12
    slots: int = 8
    # "pcg32": requests of one class that have arrived when a slot frees up
    # run together in that slot, up to max_batch and while the batch holds at
    # most batch_dofs unknowns (a batch of small problems costs about what
    # one does -- each launch is latency-bound -- but a batch of large ones
    # costs the sum of their bandwidth; a class above batch_dofs runs alone)
    max_batch: int = 8
    batch_dofs: int = 1 << 19
    hbm_budget: int = 64 << 30      # bytes of per-request state in flight
    tol: float = 1e-10              # outer (float64) relative residual
    inner_tol: float = 1e-4         # float32 MG-PCG relative residual per step
    inner_maxiter: int = 500
    max_outer: int = 20
    check_every: int = 8
    keep_solutions: bool = False


@dataclass
class ServeReport:
    results: list = field(default_factory=list)
    wall: float = 0.0
    tol: float = 1e-10
    peak_slots: int = 0
    peak_bytes: int = 0
    batch_sizes: list = field(default_factory=list)

    def stats(self):
        lat = np.array([r.latency for r in self.results])
        svc = np.array([r.end - r.start for r in self.results])
        return {"requests": len(self.results),
                "solves_per_s": round(len(self.results) / self.wall, 2),
                "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                "service_p50_ms": round(float(np.percentile(svc, 50)) * 1e3, 3),
                "service_p99_ms": round(float(np.percentile(svc, 99)) * 1e3, 3),
                "converged": int(sum(r.converged for r in self.results)),
                "at_residual_floor": int(sum(r.at_floor and not r.converged for r in self.results)),
                "all_solved": bool(all(r.converged or r.at_floor for r in self.results)),
                "max_rel_residual": float(max(np.sqrt(r.residual2 / r.atol2)
                                              for r in self.results)) * self.tol,
                "peak_slots_busy": self.peak_slots,
                "peak_request_bytes": self.peak_bytes,
                "batches": len(self.batch_sizes),
                "mean_batch": round(float(np.mean(self.batch_sizes)), 2)
                if self.batch_sizes else 0.0,
                "max_batch": int(max(self.batch_sizes)) if self.batch_sizes else 0}


class SolveServer:
    """Continuous-batching request loop over resident problem classes."""

    def __init__(self, cfg: ServeConfig | None = None, classes=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("SolveServer needs a HIP device (no CPU fallback)")
        self.cfg = cfg or ServeConfig()
        self.classes = {} if classes is None else classes
        self.streams = [torch.cuda.Stream() for _ in range(self.cfg.slots)]

    def problem_class(self, n, N) -> ProblemClass:
        key = (int(n), int(N), self.cfg.method)
        pc = self.classes.get(key)
        if pc is None:
            pc = ProblemClass(n, N, method=self.cfg.method)
            self.classes[key] = pc
        return pc

    def prepare(self, requests):
        """Build (once) every problem class the requests use."""
        for r in requests:
            self.problem_class(r.n, r.N)

    def _ir_cfg(self):
        c = self.cfg
        return _lib.IRConfig(c.tol, 0.0, c.max_outer, c.inner_tol, c.inner_maxiter,
                             c.check_every)

    def _max_group(self, key):
        """Largest batch of class ``key = (n, N)``."""
        pc = self.problem_class(*key)
        c = self.cfg
        return max(1, min(c.max_batch, c.batch_dofs // (pc.K * pc.Np)))

    def solve_group(self, reqs, stream):
        """Run requests of one class as one batch on ``stream`` (blocking;
        method "pcg32"); returns [(x, info)] in the order of ``reqs``."""
        import torch
        from types import SimpleNamespace
        pc = self.problem_class(reqs[0].n, reqs[0].N)
        if any((r.n, r.N) != (pc.n, pc.N) for r in reqs):
            raise ValueError("a batch holds requests of one problem class")
        n1 = pc.K * pc.Np
        bs = pc.acquire_batch(len(reqs))
        try:
            with torch.cuda.stream(stream):
                b = torch.empty(len(reqs) * n1, dtype=torch.float64, device=pc.op64.device)
                for i, r in enumerate(reqs):
                    f = fourier_field(pc.x, pc.y, r.seed)
                    pc.op64.rhs(f, out=b[i * n1:(i + 1) * n1])
                c = self.cfg
                x, infos = bs.pcg(b, [r.k2 for r in reqs], tol=c.tol,
                                  maxiter=c.inner_maxiter * c.max_outer,
                                  check_every=c.check_every, stream=stream.cuda_stream)
        finally:
            pc.release_batch(bs)
        return [(x[i * n1:(i + 1) * n1], _CGAsIR(SimpleNamespace(**infos[i])))
                for i in range(len(reqs))]

    def solve_one(self, req: Request, stream):
        """Run one request on ``stream`` (blocking); returns (x, info)."""
        import torch
        if self.cfg.method == "pcg32":
            return self.solve_group([req], stream)[0]
        pc = self.problem_class(req.n, req.N)
        ctx = pc.acquire()
        try:
            with torch.cuda.stream(stream):
                f = fourier_field(pc.x, pc.y, req.seed)
                b = pc.op64.rhs(f)
                x = torch.zeros_like(b)
                info = ctx.solve(b, x, req.k2, self._ir_cfg(), stream.cuda_stream)
        finally:
            pc.release(ctx)
        return x, info

    def run(self, requests, progress_every: int = 0) -> ServeReport:
        """Serve ``requests`` by their arrival times; blocks until all are done.
        ``progress_every > 0`` prints a line to stderr after every that many
        completed requests."""
        import queue
        import torch
        self.prepare(requests)
        torch.cuda.synchronize()
        cfg = self.cfg
        batched = cfg.method == "pcg32"
        adm = Admission(cfg.slots, cfg.hbm_budget)
        report = ServeReport(tol=cfg.tol)
        work = queue.Queue()
        cond = threading.Condition()
        errors = []
        t0 = time.perf_counter()

        def clock():
            return time.perf_counter() - t0

        def worker(slot):
            # one persistent worker per slot, each with its own stream; the
            # admission bound keeps at most `slots` requests in the queue
            while True:
                item = work.get()
                if item is None:
                    return
                group, nb, t_admit = item
                res = []
                try:
                    if batched:
                        outs = self.solve_group(group, self.streams[slot])
                    else:
                        outs = [self.solve_one(group[0], self.streams[slot])]
                    t_end = clock()
                    for req, (x, info) in zip(group, outs):
                        res.append(Result(req.rid, req.arrival, t_admit, t_end, info.outer,
                                          info.inner_total, bool(info.converged),
                                          info.residual2, info.atol2, bool(info.at_floor),
                                          x if cfg.keep_solutions else None))
                except Exception as e:  # surfaced after the loop
                    errors.append(e)
                with cond:
                    if res:
                        before = len(report.results)
                        report.results.extend(res)
                        report.batch_sizes.append(len(res))
                        done = len(report.results)
                        if progress_every > 0 and done // progress_every > before // progress_every:
                            print(f"[serve] {done}/{len(requests)} done at {clock():.1f} s",
                                  file=sys.stderr, flush=True)
                    adm.release(nb)
                    cond.notify_all()

        threads = [threading.Thread(target=worker, args=(i,), daemon=True)
                   for i in range(cfg.slots)]
        for th in threads:
            th.start()

        def start(group, nb, t_admit):
            work.put((group, nb, t_admit))

        def wait(timeout):
            with cond:
                cond.wait(timeout=max(0.0, min(timeout, 0.05)))

        try:
            dispatch_groups(
                requests, adm,
                lambda g: self.problem_class(g[0].n, g[0].N).group_footprint(len(g)),
                start, wait, clock, key=lambda r: (r.n, r.N),
                max_group=self._max_group if batched else 1)
        finally:
            for _ in threads:
                work.put(None)
            for th in threads:
                th.join()
        report.wall = clock()
        report.peak_slots = adm.peak_busy
        report.peak_bytes = adm.peak_used
        if errors:
            raise errors[0]
        report.results.sort(key=lambda r: r.rid)
        return report

    def close(self):
        for pc in self.classes.values():
            pc.close()
