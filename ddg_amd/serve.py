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
* Each solve is the mixed-precision refinement (float32 MG-PCG inside a
  float64 residual loop, ``ddgx_solver_ir``) to a relative residual ``tol``.

A solve reads only its own vectors and the read-only class data, and every
reduction has a fixed order, so a request's result does not depend on what
else is in flight: alone and in a mix it is bitwise the same.

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


def dispatch(requests, admission: Admission, footprint, start_fn, wait_fn, clock):
    """Admission loop shared by the GPU server and its CPU tests.

    Requests (sorted by arrival) are admitted first-come first-served as
    soon as ``clock() >= arrival``, a slot is idle and ``footprint(req)``
    bytes fit.  ``start_fn(req, nbytes, t_admit)`` hands a request to a
    worker (which calls ``admission.release(nbytes)`` when done);
    ``wait_fn(timeout)`` blocks until a worker finishes or ``timeout``
    seconds pass.  Returns when every request has been started."""
    pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
    for r in pending:
        if not admission.fits_ever(footprint(r)):
            raise ValueError(f"request {r.rid} needs {footprint(r)} bytes, more than the "
                             f"HBM budget of {admission.budget}")
    while pending:
        now = clock()
        head = pending[0]
        if head.arrival > now:
            wait_fn(head.arrival - now)
            continue
        nb = footprint(head)
        if admission.try_admit(nb):
            pending.popleft()
            start_fn(head, nb, now)
        else:
            wait_fn(0.05)


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
    on the device, and a pool of idle solver contexts."""

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
        self._idle = []
        self._lock = threading.Lock()
        self._all = []
        probe = self._new_solver()
        self.solver_bytes = probe.nbytes
        self._idle.append(probe)
        # per request: the solver context plus f, b and x (float64)
        self.footprint = self.solver_bytes + 3 * self.K * self.Np * 8

    def _new_solver(self):
        if self.method == "pcg32":
            s = SolverContext(self.op64, None, self.pc)
        else:
            s = SolverContext(self.op64, self.op32, self.pc)
        self._all.append(s)
        return s

    def acquire(self):
        with self._lock:
            if self._idle:
                return self._idle.pop()
            return self._new_solver()

    def release(self, s):
        with self._lock:
            self._idle.append(s)

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
                "peak_request_bytes": self.peak_bytes}


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

    def _cg_cfg(self):
        c = self.cfg
        return _lib.CGConfig(c.tol, 0.0, c.inner_maxiter * c.max_outer, c.check_every)

    def solve_one(self, req: Request, stream):
        """Run one request on ``stream`` (blocking); returns (x, info)."""
        import torch
        pc = self.problem_class(req.n, req.N)
        ctx = pc.acquire()
        try:
            with torch.cuda.stream(stream):
                f = fourier_field(pc.x, pc.y, req.seed)
                b = pc.op64.rhs(f)
                x = torch.zeros_like(b)
                if pc.method == "pcg32":
                    info = _CGAsIR(ctx.pcg(b, x, req.k2, self._cg_cfg(), stream.cuda_stream))
                else:
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
                req, nb, t_admit = item
                res = None
                try:
                    x, info = self.solve_one(req, self.streams[slot])
                    res = Result(req.rid, req.arrival, t_admit, clock(), info.outer,
                                 info.inner_total, bool(info.converged), info.residual2,
                                 info.atol2, bool(info.at_floor),
                                 x if cfg.keep_solutions else None)
                except Exception as e:  # surfaced after the loop
                    errors.append(e)
                with cond:
                    if res is not None:
                        report.results.append(res)
                        done = len(report.results)
                        if progress_every > 0 and done % progress_every == 0:
                            print(f"[serve] {done}/{len(requests)} done at {clock():.1f} s",
                                  file=sys.stderr, flush=True)
                    adm.release(nb)
                    cond.notify_all()

        threads = [threading.Thread(target=worker, args=(i,), daemon=True)
                   for i in range(cfg.slots)]
        for th in threads:
            th.start()

        def start(req, nb, t_admit):
            work.put((req, nb, t_admit))

        def wait(timeout):
            with cond:
                cond.wait(timeout=max(0.0, min(timeout, 0.05)))

        try:
            dispatch(requests, adm, lambda r: self.problem_class(r.n, r.N).footprint,
                     start, wait, clock)
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
