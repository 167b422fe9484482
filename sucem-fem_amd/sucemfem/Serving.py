"""Request loop for cavity and driven solves of uneven size on one GPU.

Requests arrive over time (a seeded Poisson trace).  A dispatcher admits them in
arrival order whenever a slot is free and the request's estimated HBM footprint
fits in the remaining budget (continuous admission: a request that arrives while
others run starts as soon as a slot and memory free up).  Each slot is a host
thread with its own library context on its own stream, so admitted requests
overlap on the device.  A request that can never fit the budget is rejected;
pending requests can be cancelled.  Results carry arrival / start / finish times
for latency percentiles.

The executor (what runs a request) is injectable: the device executor builds the
box mesh, space and system on the device and runs ``sucem_eigs`` or the driven
COCG solve; tests drive the admission logic with a CPU stand-in executor.
"""
import heapq
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .Consts import c0


@dataclass
class Request:
    id: int
    kind: str                 # 'eigen' or 'driven'
    dims: tuple               # box size (m)
    n: int                    # boxes per axis
    order: int
    arrival: float            # seconds after the start of the trace
    nev: int = 8
    sigma: float = 0.0
    freq: float = 0.0


@dataclass
class Result:
    id: int
    kind: str
    status: str               # 'done', 'rejected', 'cancelled', 'failed'
    arrival: float
    start: float = float("nan")
    finish: float = float("nan")
    value: object = None
    stats: dict = field(default_factory=dict)
    error: str = ""

    @property
    def latency(self):
        return self.finish - self.arrival


def analytic_lowest_k2(dims):
    a, b, c = sorted(dims, reverse=True)
    return (np.pi / a) ** 2 + (np.pi / b) ** 2          # TE101 of the largest face


def make_trace(count, rate=200.0, seed=0, n_range=(6, 48), dims_range=(0.2, 1.5), p_eigen=0.7,
               p_order2=0.5):
    """Seeded Poisson trace: dims ~ U[dims_range]^3, n ~ U{n_range}, order 1/2, eigen/driven."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        dims = tuple(float(v) for v in rng.uniform(dims_range[0], dims_range[1], size=3))
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        order = 2 if rng.random() < p_order2 else 1
        if rng.random() < p_eigen:
            out.append(Request(i, 'eigen', dims, n, order, t, nev=8,
                               sigma=0.5 * analytic_lowest_k2(dims)))
        else:
            h = max(dims) / n
            out.append(Request(i, 'driven', dims, n, order, t, freq=c0 / (10 * h)))
    return out


def box_sizes(n, order):
    """Exact DOF and nonzero counts of the Kuhn box mesh (n^3 boxes)."""
    nv = (n + 1) ** 3
    ne = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3      # axis + face diagonals + body
    nf = 6 * n * n * (n + 1) + 6 * n ** 3                          # 2 per square face + 6 per box
    ndofs = ne if order == 1 else 2 * ne + 2 * nf
    nnz = int(ndofs * (16.4 if order == 1 else 43.7))
    return nv, ne, nf, ndofs, nnz


def footprint_bytes(req):
    """Upper estimate of the device memory a request holds at its peak."""
    nc = 6 * req.n ** 3
    nloc = 6 if req.order == 1 else 20
    _, _, _, N, nnz = box_sizes(req.n, req.order)
    contrib = nc * nloc * nloc
    build = contrib * (8 + 8 + 4 + 4) + contrib * 4               # pattern sort peak
    resident = nnz * (4 + 16 + 8) + contrib * 4 + nnz * 4 + N * 64
    if req.kind == 'eigen':
        # block-2 Lanczos: basis of max(2 nev + 2, nev + 20) plus the two residual vectors
        # (V and T V), the restart copy, five work vectors, and the inner solve's workspace
        m = max(2 * req.nev + 2, req.nev + 20)
        solver = (3 * m + 10) * N * 16 + 12 * N * 16
    else:
        solver = 14 * N * 16
    return int(max(build, resident) + solver + (64 << 20))


def device_free_bytes(device=None):
    """Free device memory (hipMemGetInfo through the library), or None without a device."""
    try:
        import torch
        if not torch.cuda.is_available():
            return None
    except ImportError:
        return None
    import ctypes
    from . import _hip
    from .runtime import get_context
    ctx = get_context(device)
    free, total = ctypes.c_int64(), ctypes.c_int64()
    _hip.call("sucem_ctx_mem_info", ctx.handle, ctypes.byref(free), ctypes.byref(total))
    return int(free.value)


class Server(object):
    """Admission + slots.  ``executor(request) -> (value, stats)`` runs one request.

    ``hbm_budget``: bytes the admitted requests' footprints may sum to; ``"device"`` (the
    default) takes ``reserve`` times the device's free memory when ``run`` starts (no
    budget when there is no device), None admits by slots only."""

    def __init__(self, executor, slots=4, hbm_budget="device", footprint=footprint_bytes,
                 reserve=0.9):
        self.executor = executor
        self.slots = int(slots)
        self.hbm_budget = hbm_budget
        self.reserve = float(reserve)
        self.footprint = footprint
        self._lock = threading.Condition()
        self._in_use = 0
        self._running = 0
        self._cancelled = set()

    def cancel(self, request_id):
        with self._lock:
            self._cancelled.add(request_id)

    def budget(self):
        """The admission budget in bytes (None: none)."""
        if self.hbm_budget == "device":
            free = device_free_bytes()
            return None if free is None else int(self.reserve * free)
        return self.hbm_budget

    def run(self, trace, realtime=True, thread_init=None):
        """Replay ``trace``; returns the results in request-id order and the makespan."""
        budget = self.budget()
        results = {}
        pending = []                                   # (arrival, id, request) heap
        for r in trace:
            heapq.heappush(pending, (r.arrival, r.id, r))
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0           # noqa: E731
        admitted = queue.Queue()

        def slot_loop():
            # one persistent thread per slot: its stream and context are created once
            if thread_init:
                thread_init()
            while True:
                item = admitted.get()
                if item is None:
                    return
                req, need = item
                res = results[req.id]
                res.start = now()
                try:
                    res.value, res.stats = self.executor(req)
                    res.status = 'done'
                except Exception as e:  # noqa: BLE001 - reported per request
                    res.status, res.error = 'failed', repr(e)
                res.finish = now()
                with self._lock:
                    self._in_use -= need
                    self._running -= 1
                    self._lock.notify_all()

        workers = [threading.Thread(target=slot_loop, daemon=True) for _ in range(self.slots)]
        for th in workers:
            th.start()

        while pending:
            arrival, rid, req = pending[0]
            if realtime and arrival > now():
                time.sleep(min(arrival - now(), 0.01))
                continue
            need = self.footprint(req)
            with self._lock:
                if rid in self._cancelled:
                    heapq.heappop(pending)
                    results[rid] = Result(rid, req.kind, 'cancelled', arrival)
                    continue
                if budget is not None and need > budget:
                    heapq.heappop(pending)
                    results[rid] = Result(rid, req.kind, 'rejected', arrival,
                                          error="needs %d bytes, budget %d" % (need, budget))
                    continue
                while self._running >= self.slots or (
                        budget is not None and self._in_use + need > budget):
                    self._lock.wait(0.05)
                heapq.heappop(pending)
                self._in_use += need
                self._running += 1
            results[rid] = Result(rid, req.kind, 'running', arrival)
            admitted.put((req, need))
        for _ in workers:
            admitted.put(None)
        for th in workers:
            th.join()
        makespan = now()
        return [results[k] for k in sorted(results)], makespan


def summarize(results, makespan):
    done = [r for r in results if r.status == 'done']
    lat = np.array([r.latency for r in done]) if done else np.array([np.nan])
    return {"requests": len(results), "done": len(done),
            "rejected": sum(r.status == 'rejected' for r in results),
            "failed": sum(r.status == 'failed' for r in results),
            "solves_per_s": len(done) / makespan if makespan > 0 else 0.0,
            "latency_ms_p50": float(np.percentile(lat, 50) * 1e3),
            "latency_ms_p99": float(np.percentile(lat, 99) * 1e3),
            "makespan_s": makespan}


# ---------------------------------------------------------------------------------------
# device executor
# ---------------------------------------------------------------------------------------
def device_executor(tol=1e-8, max_iter=50000, precision=64, eig_tol=1e-8, eig_inner_tol=1e-9,
                    eig_precision=None):
    """Executor running each request on the device on the calling thread's own stream.
    ``precision``: 64 (FP64 throughout) or 32 (FP32 operator and Krylov vectors with FP64
    residual replacement) for the driven solves and, unless ``eig_precision`` says
    otherwise, for the eigen solver's inner solves.  The eigen requests stop at a relative
    Ritz residual of ``eig_tol`` with inner solves to ``eig_inner_tol``: the returned k^2
    are FP64 Rayleigh quotients, whose error is of the order of the square of the modes'
    residual."""
    if precision not in (32, 64):
        raise ValueError("precision must be 32 or 64")
    eig_precision = precision if eig_precision is None else int(eig_precision)
    import torch
    from .Utilities.Workloads import DeviceProblem
    from . import _hip

    local = threading.local()

    def ctx():
        # a stream and a library context per slot thread; torch hands out streams from a
        # fixed pool, so two threads can be given the same stream: the context (and the
        # solver workspace it holds) is still the thread's own, never shared
        if not hasattr(local, "stream"):
            from .runtime import Context
            local.stream = torch.cuda.Stream()             # default priority
            local.context = Context(stream=local.stream)
        return local.stream, local.context

    def run(req):
        stream, context = ctx()
        with torch.cuda.stream(stream):
            if req.kind == 'eigen':
                prob = DeviceProblem(req.n, req.dims, req.order, _hip.MAT_S | _hip.MAT_T, ctx=context)
                k2, resid, st = prob.eigs(req.sigma, req.nev, tol=eig_tol,
                                          inner_tol=eig_inner_tol,
                                          inner_precision=eig_precision)
                stream.synchronize()
                return k2, dict(st, ndofs=prob.ndofs)
            prob = DeviceProblem(req.n, req.dims, req.order,
                                 _hip.MAT_S | _hip.MAT_T | _hip.MAT_S0, ctx=context)
            k0 = 2 * np.pi * req.freq / c0
            centre = np.asarray(req.dims) / 2 + 1e-9
            b = prob.rhs_dipoles(centre[None, :], np.array([[0, 0, c0 / req.freq / 1000]]), k0)
            x = torch.zeros_like(b)
            st = prob.solve(k0, b, x, precision=precision, tol=tol, max_iter=max_iter)
            stream.synchronize()
            return x.cpu().numpy(), dict(st, ndofs=prob.ndofs)

    return run
