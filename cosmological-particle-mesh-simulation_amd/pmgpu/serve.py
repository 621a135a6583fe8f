"""Serving loop: simulation requests admitted by HBM footprint and advanced with continuous
batching.

A request names one realisation of the reference's run (pmesh.py:18-77): seed, box, N_PARTS,
N_CELLS and STEPS, plus the cosmology.  The server keeps every admitted simulation's particles
and grids resident on the device and advances all of them together through one
``PMEnsemble`` (one launch per stage for all members), a chunk of steps at a time.  Between
chunks it retires finished simulations (their memory is freed at once) and admits the
requests that have arrived, in arrival order, as long as their footprint fits the HBM budget
and an ensemble slot is free; a request that does not fit waits, later smaller ones may pass
it.  Host-side preparation of a request (the initial power spectrum table) runs in a worker
thread while the device steps the others.

Results carry the arrival, admission and completion times (time to completion =
completion - arrival), the executed step count, the final scale factor, the matter power
spectrum of the final density and, on request, the final particle state.
"""
from __future__ import annotations

import concurrent.futures as cf
import re
import threading
import time
from collections import deque
from dataclasses import dataclass, field

import ctypes as C
import numpy as np

from . import _lib
from .ensemble import MAX_MEMBERS, PMEnsemble
from .sim import PMSim


@dataclass
class SimRequest:
    id: int
    seed: int
    box: float
    n_cells: int
    n_parts: int
    steps: int                  # STEPS of configure_me.py: da = (A_END - A_INIT) / STEPS
    arrival: float = 0.0        # seconds after the start of serving
    a_init: float = 0.02
    a_end: float = 1.0
    om: float = 0.3
    ob: float = 0.045
    ol: float = 0.7
    ok: float = 0.0
    h: float = 0.7
    ns: float = 0.96
    sigma8: float = 0.8

    @property
    def da(self) -> float:
        return (self.a_end - self.a_init) / self.steps

    @property
    def loop_steps(self) -> int:
        """Steps pmesh.py:56-63 executes: `while a < A_END - da: ...; a += da` in float64."""
        da = self.da
        a, k = self.a_init, 0
        while a < self.a_end - da:
            a += da
            k += 1
        return k

    def ic_params(self) -> dict:
        return dict(n_parts=self.n_parts, n_cells=self.n_cells, box=self.box, seed=self.seed,
                    om=self.om, ob=self.ob, ol=self.ol, ok=self.ok, h=self.h, ns=self.ns,
                    sigma8=self.sigma8, a_init=self.a_init)

    def footprint(self) -> int:
        return PMSim.footprint(self.n_cells, self.n_parts, 32)


@dataclass
class SimResult:
    id: int
    arrival: float
    admitted: float
    finished: float
    steps: int
    a_final: float
    k: np.ndarray
    pk: np.ndarray
    positions: np.ndarray | None = None
    velocities: np.ndarray | None = None
    error: str | None = None     # set when the request could not be served

    @property
    def time_to_completion(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Active:
    req: SimRequest
    sim: PMSim
    a: float
    remaining: int
    admitted: float
    fp: int


def device_memory() -> tuple[int, int]:
    """(free, total) bytes of device memory."""
    _lib.init()
    free, total = C.c_size_t(), C.c_size_t()
    _lib.call("pm_mem_info", C.byref(free), C.byref(total))
    return free.value, total.value


def _prepare(req: SimRequest):
    """Host-side inputs of the device initial conditions (power spectrum table, factors)."""
    import initial_conditions as ic
    p = req.ic_params()
    kf = 2 * np.pi / p["box"]
    scale = (p["n_cells"] / p["box"]) / kf / float(p["n_parts"]) ** 3
    return ic.amplitude_table(p), scale, ic.velocity_factor(p)


def _start(req: SimRequest, prepared) -> PMSim:
    amp, scale, vfac = prepared
    sim = PMSim(req.n_cells, req.n_parts, 32, omega_m0=req.om, omega_lambda0=req.ol,
                omega_k0=req.ok, h0=req.h)
    try:
        sim.zeldovich(amp, scale, vfac, seed=req.seed)
        sim.deposit()
    except Exception:
        sim.close()
        raise
    return sim


class PMServer:
    """Continuous-batching server on one GPU.

    hbm_budget: bytes the admitted simulations may hold together (default: 90 % of the
    device memory free at construction).  chunk_steps: most steps per batched call; a chunk
    also ends when the first active simulation finishes, so completions and admissions happen
    at chunk boundaries.

    Admission (device allocation, initial conditions, first deposit) and retirement (power
    spectrum, state copy, release) run in worker threads on each simulation's own stream while
    the ensemble's chunk runs on its stream, so the device does not idle between chunks.  A
    request whose start fails, or whose state turns non-finite, ends with ``error`` set; the
    others are served normally.

    ``cancel(id)`` (from any thread, or from ``on_result``) withdraws a request: one still
    waiting ends at once, one being started or stepped at the next chunk boundary; its memory
    and ensemble slot are released and its result carries ``error == "cancelled"`` and the
    steps it took.  A request already retiring is served as usual.
    """

    def __init__(self, hbm_budget: int | None = None, max_members: int = MAX_MEMBERS,
                 chunk_steps: int = 16, keep_state: bool = False, prepare_workers: int = 2,
                 start=None, ensemble=None, prepare=None):
        if hbm_budget is None:
            hbm_budget = int(device_memory()[0] * 0.9)
        self.hbm_budget = int(hbm_budget)
        self.max_members = min(int(max_members), MAX_MEMBERS)
        self.chunk_steps = max(1, int(chunk_steps))
        self.keep_state = keep_state
        self.prepare_workers = prepare_workers
        # replaceable pieces (tests of the scheduling and failure paths use stand-ins)
        self._start = start or _start
        self._ensemble = ensemble or PMEnsemble
        self._prepare = prepare or _prepare
        self.stats = {}
        self._cancel_lock = threading.Lock()
        self._cancel_ids: set[int] = set()

    def cancel(self, request_id: int) -> None:
        """Withdraw request `request_id` (served requests and unknown ids are ignored)."""
        with self._cancel_lock:
            self._cancel_ids.add(int(request_id))

    def _take_cancels(self) -> set[int]:
        with self._cancel_lock:
            ids, self._cancel_ids = self._cancel_ids, set()
        return ids

    def serve(self, requests, realtime: bool = True, clock=time.perf_counter,
              on_result=None) -> list[SimResult]:
        """Serve `requests` (arrival times in seconds from now).  realtime=False treats the
        arrival times as an order only: every request is available from the start."""
        reqs = sorted(requests, key=lambda r: (r.arrival, r.id))
        for r in reqs:
            if r.footprint() > self.hbm_budget:
                raise ValueError(f"request {r.id} needs {r.footprint()} bytes, more than the "
                                 f"HBM budget {self.hbm_budget}")
        pending = deque(reqs)
        queue: list[SimRequest] = []
        prepared: dict[int, cf.Future] = {}
        starting: list[tuple[SimRequest, cf.Future, float, int]] = []
        retiring: list[tuple[_Active, cf.Future]] = []
        active: list[_Active] = []
        results: list[SimResult] = []
        used = 0
        busy = admit_s = retire_s = 0.0
        chunks = 0
        peak_members = 0
        failed = cancelled = 0
        withdrawn: set[int] = set()   # cancelled while being started: dropped when they join
        ens = self._ensemble()
        pool = cf.ThreadPoolExecutor(max_workers=self.prepare_workers)
        starter = cf.ThreadPoolExecutor(max_workers=1)
        retirer = cf.ThreadPoolExecutor(max_workers=1)
        keep = self.keep_state
        t0 = clock()

        def now():
            return clock() - t0

        def arrival(req):
            return req.arrival if realtime else 0.0

        def finish(res):
            results.append(res)
            if on_result:
                on_result(res)

        def fail_result(req, admitted, exc):
            return SimResult(req.id, arrival(req), admitted, now(), 0, req.a_init,
                             np.zeros(0), np.zeros(0), error=f"{type(exc).__name__}: {exc}")

        def cancel_result(req, admitted, steps=0, a=None):
            nonlocal cancelled
            cancelled += 1
            finish(SimResult(req.id, arrival(req), admitted, now(), steps,
                             req.a_init if a is None else a, np.zeros(0), np.zeros(0),
                             error="cancelled"))

        def apply_cancels():
            nonlocal used
            ids = self._take_cancels()
            if not ids:
                return
            for r in [r for r in pending if r.id in ids]:
                pending.remove(r)
                cancel_result(r, now())
            for r in [r for r in queue if r.id in ids]:
                queue.remove(r)
                fut = prepared.pop(r.id, None)
                if fut is not None:
                    fut.cancel()
                cancel_result(r, now())
            withdrawn.update(r.id for r, _, _, _ in starting if r.id in ids)
            for x in [x for x in active if x.req.id in ids]:
                ens.remove(x.sim)
                x.sim.close()
                used -= x.fp
                active.remove(x)
                cancel_result(x.req, x.admitted, x.req.loop_steps - x.remaining, x.a)

        def start_job(req, fut):
            ta = clock()
            sim = self._start(req, fut.result())
            return sim, clock() - ta

        def retire_job(x):
            tr = clock()
            try:
                k, pk, _ = x.sim.power_spectrum(x.req.box)
                pos = vel = None
                if keep:
                    pos, vel = x.sim.get_state()
            finally:
                x.sim.close()
            return k, pk, pos, vel, clock() - tr

        def collect_retired(wait):
            nonlocal used, retire_s
            for x, fut in list(retiring):
                if not (wait or fut.done()):
                    continue
                retiring.remove((x, fut))
                used -= x.fp
                try:
                    k, pk, pos, vel, dt = fut.result()
                    retire_s += dt
                    finish(SimResult(x.req.id, arrival(x.req), x.admitted, now(), x.req.loop_steps,
                                     x.a, k, pk, pos, vel))
                except Exception as exc:  # noqa: BLE001 - reported in the result
                    finish(fail_result(x.req, x.admitted, exc))

        def join_started(wait):
            nonlocal used, admit_s, failed
            for item in list(starting):
                r, fut, t_adm, fp = item
                if not (wait or fut.done()):
                    continue
                starting.remove(item)
                try:
                    sim, dt = fut.result()
                except Exception as exc:  # noqa: BLE001 - reported in the result
                    used -= fp
                    if r.id in withdrawn:
                        withdrawn.discard(r.id)
                        cancel_result(r, t_adm)
                    else:
                        failed += 1
                        finish(fail_result(r, t_adm, exc))
                    continue
                if r.id in withdrawn:
                    withdrawn.discard(r.id)
                    sim.close()
                    used -= fp
                    cancel_result(r, t_adm)
                    continue
                try:
                    ens.add(sim)
                except Exception as exc:  # noqa: BLE001
                    sim.close()
                    used -= fp
                    failed += 1
                    finish(fail_result(r, t_adm, exc))
                    continue
                admit_s += dt
                active.append(_Active(r, sim, r.a_init, r.loop_steps, t_adm, fp))

        try:
            while pending or queue or active or starting or retiring:
                apply_cancels()
                t = now()
                while pending and (not realtime or pending[0].arrival <= t):
                    r = pending.popleft()
                    queue.append(r)
                    prepared[r.id] = pool.submit(self._prepare, r)
                # admission: arrival order, skipping requests that do not fit yet; the sims are
                # built in the starter thread and join the ensemble at the next chunk boundary
                for r in list(queue):
                    if len(active) + len(starting) >= self.max_members:
                        break
                    fp = r.footprint()
                    if used + fp > self.hbm_budget:
                        continue
                    used += fp
                    queue.remove(r)
                    starting.append((r, starter.submit(start_job, r, prepared.pop(r.id)), now(), fp))
                join_started(wait=not active)
                collect_retired(wait=not active and not starting)
                peak_members = max(peak_members, len(active))
                if not active:
                    if pending and realtime and not starting:
                        time.sleep(max(0.0, min(pending[0].arrival - now(), 0.05)))
                    continue
                chunk = min(self.chunk_steps, min(x.remaining for x in active))
                if chunk > 0:
                    tb = clock()
                    sick = None
                    try:
                        a_new = ens.run([x.a for x in active], [x.req.da for x in active], chunk)
                    except RuntimeError as exc:
                        a_new = getattr(exc, "a_after", None)
                        m = re.search(r"ensemble member (\d+) \(after (\d+) steps\)", str(exc))
                        if a_new is None or m is None:
                            raise
                        sick, chunk = (int(m.group(1)), exc), int(m.group(2))
                    busy += clock() - tb
                    chunks += 1
                    for x, a in zip(active, a_new):
                        x.a = a
                        x.remaining -= chunk
                    if sick is not None:
                        x = active[sick[0]]
                        ens.remove(x.sim)
                        x.sim.close()
                        used -= x.fp
                        active.remove(x)
                        failed += 1
                        finish(fail_result(x.req, x.admitted, sick[1]))
                for x in [x for x in active if x.remaining <= 0]:
                    ens.remove(x.sim)
                    active.remove(x)
                    retiring.append((x, retirer.submit(retire_job, x)))
        finally:
            for x in active:
                x.sim.close()
            for _, fut, _, _ in starting:
                try:
                    fut.result()[0].close()
                except Exception:  # noqa: BLE001 - closing what did start
                    pass
            starter.shutdown(wait=True)
            retirer.shutdown(wait=True)
            ens.close()
            pool.shutdown(wait=True)
        wall = now()
        self._take_cancels()  # ids of requests that had already finished
        self.stats = {"wall_s": wall, "device_busy_s": busy, "admit_s": admit_s,
                      "retire_s": retire_s, "chunks": chunks, "peak_members": peak_members,
                      "requests": len(results), "failed": failed, "cancelled": cancelled}
        return sorted(results, key=lambda r: r.id)


def run_alone(req: SimRequest, keep_state: bool = False, tiles=None) -> SimResult:
    """One request by itself: its own sim, device initial conditions, the graph-replayed
    step loop and the final power spectrum (the one-at-a-time way of serving).  ``tiles``
    fixes the tile shape (the ensemble's 16^3 makes the run use the served kernels)."""
    t0 = time.perf_counter()
    sim = _start(req, _prepare(req))
    try:
        if tiles:
            sim.set_tiles(*tiles)
        n = req.loop_steps
        a = sim.run(req.a_init, req.da, n)
        k, pk, _ = sim.power_spectrum(req.box)
        pos = vel = None
        if keep_state:
            pos, vel = sim.get_state()
    finally:
        sim.close()
    return SimResult(req.id, 0.0, 0.0, time.perf_counter() - t0, n, a, k, pk, pos, vel)


def poisson_stream(rate: float, duration: float, seed: int = 0, first_id: int = 0,
                   sizes=(64, 128, 192, 256, 384), weights=(0.3, 0.3, 0.2, 0.15, 0.05),
                   box=(50.0, 400.0), steps=(50, 500), time_scale: float = 1.0):
    """Requests arriving as a Poisson process of `rate` per second over `duration` seconds:
    N_CELLS drawn from `sizes` with `weights`, N_PARTS = N_CELLS / 2, box uniform, STEPS
    uniform over the inclusive range, sequential seeds.  time_scale < 1 compresses the arrival
    times (the same requests, arriving faster)."""
    rng = np.random.default_rng(seed)
    t, out, i = 0.0, [], first_id
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        n = int(rng.choice(sizes, p=np.asarray(weights) / np.sum(weights)))
        out.append(SimRequest(id=i, seed=1000 + i, box=float(rng.uniform(*box)), n_cells=n,
                              n_parts=n // 2, steps=int(rng.integers(steps[0], steps[1] + 1)),
                              arrival=t * time_scale))
        i += 1
    return out


def latency_summary(results) -> dict:
    ttc = np.array([r.time_to_completion for r in results])
    return {"p50_s": float(np.percentile(ttc, 50)), "p99_s": float(np.percentile(ttc, 99)),
            "mean_s": float(ttc.mean()), "max_s": float(ttc.max())}
