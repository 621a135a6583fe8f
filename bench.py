#!/usr/bin/env python3
"""Benchmark of the PM step on one MI355X.

Workload: 256^3 particles on a 512^3 grid (the reference README's shape), box 250 Mpc/h,
Omega_m 0.3, Omega_L 0.7, h 0.7, Eisenstein-Hu no-wiggle P(k), n_s 0.96, sigma_8 0.8, seed 7,
Zel'dovich ICs at a = 0.02, 1000 steps to a = 1 (pmesh.py loop).  One "step" is one PM step
(deposit, FFT Poisson solve, gather + kick + drift) of the device-resident simulation.

The timed region is K consecutive steps starting from the initial conditions (K = 1000 by
default: the whole run), FP32; the same run is then timed in FP64.  `value` is
particle-updates/s = N_p * K / wall time, inputs resident in HBM.  A second, shorter FP32 pass
records HIP events around every stage on the simulation's stream for the per-kernel roofline.
The CPU baseline is the C/OpenMP port of the reference loops plus scipy complex128 FFTs,
timed on this host for a bounded number of steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "cosmological-particle-mesh-simulation_amd"
sys.path.insert(0, str(PKG))

METRIC = "PM particle-updates/s at 256³/512³ FP32 vs FP64 GPU; serve p50/p99"
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(n_cells=512, n_parts=256, box=250.0, seed=7, a_init=0.02, a_end=1.0, steps=1000,
                om=0.3, ol=0.7, ok=0.0, h=0.7, ob=0.045, ns=0.96, sigma8=0.8)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_bytes(n, n_p, prec):
    """Algorithmic HBM bytes per launch of each stage (one read + one write of what it must
    touch).  Half spectrum: N^2 (N/2+1) complex values."""
    r = prec // 8
    cells = n ** 3
    half = n * n * (n // 2 + 1) * 2 * r
    return {
        "bin": 3 * r * n_p + 4 * n_p,                   # read displacements, write >= 1 entry each
        "deposit": 3 * r * n_p + 4 * n_p + r * cells,   # read entries + displacements, write rho
        "r2c_x": r * cells + half,                      # read rho, write half spectrum
        "c2c_y_fwd": 2 * half,
        "z_green": 2 * half,
        "c2c_y_inv": 2 * half,
        "c2r_x": half + r * cells,                      # read half spectrum, write phi
        "gather_kick_drift": 12 * r * n_p + 4 * n_p + r * cells,  # entries, x, v in/out, phi
        "deposit_r2c": 3 * r * n_p + 4 * n_p + half,    # entries + displacements in, spectrum out
    }


def step_bytes(n, n_p, prec):
    """Per-step algorithmic bytes as defined for this metric: 60 N_p + 32 N^3 + 16 N^2 (N/2+1)
    at FP32 (doubled at FP64)."""
    return (60 * n_p + 32 * n ** 3 + 16 * n * n * (n // 2 + 1)) * (prec // 32)


def _ic_params(w):
    return dict(n_parts=w["n_parts"], n_cells=w["n_cells"], box=w["box"], seed=w["seed"], om=w["om"],
                ob=w["ob"], ol=w["ol"], ok=w["ok"], h=w["h"], ns=w["ns"], sigma8=w["sigma8"],
                a_init=w["a_init"])


def make_ics(w):
    import contextlib
    import initial_conditions as ic
    p = _ic_params(w)
    with contextlib.redirect_stdout(sys.stderr):   # stdout carries only the JSON line
        delta = ic.gaussian_random_field(p)
        return ic.zeldovich_numpy(delta, p)


def run_gpu(prec, pos, vel, w, steps, warmup, tiles=None):
    from pmgpu.sim import PMSim
    sim = PMSim(w["n_cells"], w["n_parts"], prec, omega_m0=w["om"], omega_lambda0=w["ol"],
                omega_k0=w["ok"], h0=w["h"])
    da = (w["a_end"] - w["a_init"]) / w["steps"]
    sim.set_state(pos, vel)
    if tiles:                        # a fixed shape (profiles that must see one kernel variant)
        sim.set_tiles(*tiles)
        sim.tuned = (sim.tiles(), None)
    else:
        sim.tuned = sim.autotune(3)  # tile shape for these particles (set-up, not timed)
    if warmup:
        sim.run(w["a_init"], da, warmup)
    sim.set_state(pos, vel)          # the timed region starts from the initial conditions
    sim.deposit()
    sim.sync()
    t0 = time.perf_counter()
    a_end = sim.run(w["a_init"], da, steps)
    sim.sync()
    dt = time.perf_counter() - t0
    return sim, dt, a_end


def stage_timing(pos, vel, w, steps, tiles=(0, 0, 0)):
    from pmgpu.sim import PMSim
    sim = PMSim(w["n_cells"], w["n_parts"], 32, omega_m0=w["om"], omega_lambda0=w["ol"],
                omega_k0=w["ok"], h0=w["h"])
    da = (w["a_end"] - w["a_init"]) / w["steps"]
    sim.set_tiles(*tiles)
    sim.set_state(pos, vel)
    sim.run(w["a_init"], da, 2)
    sim.set_timing(True)
    sim.run(w["a_init"] + 2 * da, da, steps)
    t = {k: v for k, v in sim.timing().items() if v[1] > 0}   # stages that ran
    sim.close()
    return {k: ms / c for k, (ms, c) in t.items()}, {k: c for k, (ms, c) in t.items()}


def plain_fp64_bench(pos, vel, w, steps):
    """The reference's algorithm as plain GPU code: the reference-layout entry points
    (density(): global float atomics into a new grid; advance_time(): complex128 FFT Poisson
    solve, float32 potential, reference integrate with pre-step cells for all axes), absolute
    float32 positions, called once per step like pmesh.py:60-61."""
    from pmgpu import ops
    from pmgpu.device import DeviceArray
    n, npart = w["n_cells"], w["n_parts"]
    mass = (n / npart) ** 3
    da = (w["a_end"] - w["a_init"]) / w["steps"]
    x = DeviceArray.from_numpy(pos)
    v = DeviceArray.from_numpy(vel)
    g = ops.fourier_grid(n, host=False)
    cf = (w["h"], w["ol"], w["ok"])        # integrate.py:12 passes [H0, OL, OK]
    a = w["a_init"]
    rho = ops.density(x, mass, n)          # warm-up: workspaces, tables
    ops.advance_time(rho, x, v, g, a, da, cf, w["om"])
    a += da
    t0 = time.perf_counter()
    for _ in range(steps):
        rho = ops.density(x, mass, n)
        ops.advance_time(rho, x, v, g, a, da, cf, w["om"])   # synchronous, like the reference
        a += da
    dt = time.perf_counter() - t0
    return {"value": npart ** 3 * steps / dt, "ms_per_step": dt / steps * 1e3, "steps": steps,
            "path": "reference-layout entry points pm_density + pm_advance_time (float32 "
                    "storage, complex128 FFT, global atomics, one call per reference function)"}


# Batched ensemble workload: 48 realisations, seeds 100-147, N_grid uniform over 128/192/256
# with (N_grid/2)^3 particles, box U(50, 300) Mpc/h, sigma8 U(0.7, 0.9), 200 steps a 0.02 -> 1.
ENSEMBLE = dict(members=48, seed0=100, sizes=(128, 192, 256), box=(50.0, 300.0),
                sigma8=(0.7, 0.9), steps=200, a_init=0.02, a_end=1.0, meta_seed=2024)


def loop_steps(a_init, a_end, steps):
    """Steps the reference's host loop executes (pmesh.py:56-63, float64 accumulation of a)."""
    da = (a_end - a_init) / steps
    a, k = a_init, 0
    while a < a_end - da:
        a += da
        k += 1
    return k


def ensemble_members(e):
    """The ensemble's per-member draws from a fixed meta generator (documented seed)."""
    rng = np.random.default_rng(e["meta_seed"])
    out = []
    for i in range(e["members"]):
        n = int(rng.choice(e["sizes"]))
        out.append(dict(WORKLOAD, n_cells=n, n_parts=n // 2, seed=e["seed0"] + i,
                        box=float(rng.uniform(*e["box"])), sigma8=float(rng.uniform(*e["sigma8"]))))
    return out


def ensemble_bench(e, steps=None):
    """All members in one ensemble (one launch per stage, graph-replayed) against the same
    members run one at a time, each with its own graph."""
    from pmgpu.ensemble import PMEnsemble
    from pmgpu.sim import PMSim
    members = ensemble_members(e)
    n_steps = steps or loop_steps(e["a_init"], e["a_end"], e["steps"])
    da = (e["a_end"] - e["a_init"]) / e["steps"]
    import initial_conditions as ic
    log(f"ensemble: {len(members)} members, initial conditions on the device (Philox seeds)")
    sims = []
    for m in members:
        s = PMSim(m["n_cells"], m["n_parts"], 32, omega_m0=m["om"], omega_lambda0=m["ol"],
                  omega_k0=m["ok"], h0=m["h"])
        ic.zeldovich_device(s, _ic_params(m))
        s.deposit()
        sims.append(s)
    # the members' starting state, restored before each timed run
    ics = [s.get_state() for s in sims]
    updates = sum(m["n_parts"] ** 3 for m in members) * n_steps
    bytes_step = sum(step_bytes(m["n_cells"], m["n_parts"] ** 3, 32) for m in members)
    ens = PMEnsemble(sims)
    ens.run([e["a_init"]] * len(sims), [da] * len(sims), 8)        # warm-up: capture + replay
    for s, (pos, vel) in zip(sims, ics):
        s.set_state(pos, vel)
        s.deposit()
    log(f"ensemble: {n_steps} batched steps")
    t0 = time.perf_counter()
    ens.run([e["a_init"]] * len(sims), [da] * len(sims), n_steps)
    ens.sync()
    t_ens = time.perf_counter() - t0
    ens.set_timing(True)                              # per-stage HIP events, 10 more steps
    ens.run([e["a_end"] - 11 * da] * len(sims), [da] * len(sims), 10)
    stage = {k: round(ms / max(c, 1), 4) for k, (ms, c) in ens.timing().items()}
    ens.close()
    log(f"ensemble: {t_ens:.3f} s; one at a time")
    t_solo = 0.0
    for s, (pos, vel) in zip(sims, ics):
        s.run(e["a_init"], da, 8)                                   # capture + warm-up
        s.set_state(pos, vel)
        s.deposit()
        s.sync()
        t0 = time.perf_counter()
        s.run(e["a_init"], da, n_steps)
        s.sync()
        t_solo += time.perf_counter() - t0
        s.close()
    log(f"one at a time: {t_solo:.3f} s")
    counts = {n: sum(1 for m in members if m["n_cells"] == n) for n in e["sizes"]}
    return {"workload": f"ensemble_{len(members)}: seeds {e['seed0']}-{e['seed0'] + len(members) - 1}, "
                        f"N_grid in {list(e['sizes'])} (counts {counts}), (N_grid/2)^3 particles, "
                        f"box U{e['box']}, sigma8 U{e['sigma8']}, a {e['a_init']} -> {e['a_end']}",
            "steps": n_steps, "particle_updates": updates,
            "value": updates / t_ens, "unit": "particle-updates/s", "seconds": t_ens,
            "ms_per_step": t_ens / n_steps * 1e3,
            "roofline_frac": bytes_step * n_steps / t_ens / 1e9 / HBM_PEAK_GBS,
            "stages_ms": stage,
            "one_at_a_time": {"value": updates / t_solo, "seconds": t_solo},
            "speedup_vs_one_at_a_time": t_solo / t_ens}


# Largest single run: 512^3 particles on a 1024^3 grid, box 500 Mpc/h, seed 42, a 0.02 -> 1 in
# 500 steps; device initial conditions (the host generator would need tens of GB of RAM).
LARGEST = dict(WORKLOAD, n_cells=1024, n_parts=512, box=500.0, seed=42, steps=500)


def largest_bench(w, steps):
    """The largest run, FP32, graph-replayed, timed like the main workload: all of its steps
    (steps < 0) or the first `steps`."""
    total = loop_steps(w["a_init"], w["a_end"], w["steps"])
    steps = total if steps < 0 else min(steps, total)
    import initial_conditions as ic
    from pmgpu.sim import PMSim
    sim = PMSim(w["n_cells"], w["n_parts"], 32, omega_m0=w["om"], omega_lambda0=w["ol"],
                omega_k0=w["ok"], h0=w["h"])
    ic.zeldovich_device(sim, _ic_params(w))
    tiles, _ = sim.autotune(2)
    pos, vel = sim.get_state()
    da = (w["a_end"] - w["a_init"]) / w["steps"]
    sim.run(w["a_init"], da, 8)                       # capture + warm-up
    sim.set_state(pos, vel)
    sim.deposit()
    sim.sync()
    t0 = time.perf_counter()
    sim.run(w["a_init"], da, steps)
    sim.sync()
    dt = time.perf_counter() - t0
    n, n_p = w["n_cells"], w["n_parts"] ** 3
    sim.close()
    ms = dt / steps * 1e3
    return {"workload": "pm_step_512p_1024g: 512^3 particles, 1024^3 grid, box 500 Mpc/h, seed 42, "
                        + (f"all {total} steps" if steps == total else f"first {steps} of {total} steps")
                        + f" from a = {w['a_init']}, device ICs",
            "value": n_p * steps / dt, "unit": "particle-updates/s", "ms_per_step": ms,
            "steps": steps, "tiles": list(tiles),
            "step_roofline_frac": step_bytes(n, n_p, 32) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


# Serving stream: Poisson arrivals at 2 requests/s for 600 s, N_grid in {64..384} with weights
# 0.3/0.3/0.2/0.15/0.05, N_parts = N_grid/2, box U(50, 400), STEPS U{50..500}, sequential seeds.
SERVE = dict(rate=2.0, duration=600.0, seed=4, sizes=(64, 128, 192, 256, 384),
             weights=(0.3, 0.3, 0.2, 0.15, 0.05), box=(50.0, 400.0), steps=(50, 500))


def serve_bench(sv, window, chunk_steps=16):
    """The first `window` seconds of the serving stream, three ways: served in real time with
    continuous batching (time to completion = completion - arrival); served with every request
    available at once (capacity, sims/hour); and one request at a time back to back (its
    capacity, and its time to completion under the same arrivals by FIFO replay of the
    measured run times)."""
    from pmgpu.serve import PMServer, SimRequest, latency_summary, poisson_stream, run_alone
    reqs = poisson_stream(sv["rate"], window, seed=sv["seed"], sizes=sv["sizes"],
                          weights=sv["weights"], box=sv["box"], steps=sv["steps"])
    warm = [SimRequest(id=-1 - i, seed=1, box=100.0, n_cells=n, n_parts=n // 2, steps=20)
            for i, n in enumerate(sv["sizes"])]
    PMServer().serve(warm, realtime=False)
    for r in warm:
        run_alone(r)
    updates = sum(r.n_parts ** 3 * r.loop_steps for r in reqs)
    log(f"serving: {len(reqs)} requests over {window:.0f} s, real time")
    server = PMServer(chunk_steps=chunk_steps)
    res = server.serve(reqs, realtime=True)
    live = dict(latency_summary(res), **server.stats)
    log(f"serving: p50 {live['p50_s']:.3f} s p99 {live['p99_s']:.3f} s; all at once")
    server = PMServer(chunk_steps=chunk_steps)
    t0 = time.perf_counter()
    server.serve(reqs, realtime=False)
    t_batch = time.perf_counter() - t0
    log(f"serving all at once: {t_batch:.2f} s; one at a time")
    service = []
    for r in reqs:
        service.append(run_alone(r).finished)
    t_solo = float(np.sum(service))
    done, ttc = 0.0, []
    for r, t in zip(reqs, service):                    # requests are in arrival order
        start = max(r.arrival, done)
        done = start + t
        ttc.append(done - r.arrival)
    log(f"one at a time: {t_solo:.2f} s; the pmesh.py way on a sample")
    drv = driver_sample(reqs, per_size=2)
    t_drv = sum(drv["mean_s"][r.n_cells] for r in reqs)
    counts = {n: sum(1 for r in reqs if r.n_cells == n) for n in sv["sizes"]}
    return {"workload": f"serve: Poisson {sv['rate']}/s over the first {window:.0f} s of the "
                        f"stream (seed {sv['seed']}), N_grid counts {counts}, N_parts = N_grid/2, "
                        f"box U{sv['box']}, STEPS U{sv['steps']}, device ICs, FP32",
            "requests": len(reqs), "particle_updates": updates,
            "realtime": {"p50_s": live["p50_s"], "p99_s": live["p99_s"], "mean_s": live["mean_s"],
                         "wall_s": live["wall_s"], "device_busy_s": live["device_busy_s"],
                         "peak_members": live["peak_members"]},
            "capacity_sims_per_hour": len(reqs) / t_batch * 3600,
            "capacity_particle_updates_per_s": updates / t_batch,
            "one_at_a_time": {"sims_per_hour": len(reqs) / t_solo * 3600,
                              "p50_s": float(np.percentile(ttc, 50)),
                              "p99_s": float(np.percentile(ttc, 99))},
            "capacity_vs_one_at_a_time": t_solo / t_batch,
            "one_at_a_time_pmesh_driver": {
                "sims_per_hour": len(reqs) / t_drv * 3600,
                "how": "each request as its own pmesh.py-style run: host initial conditions "
                       "(NumPy), resident sim stepped once per loop iteration from Python, "
                       "device P(k); time per N_grid from the first "
                       f"{drv['per_size']} requests of each size, summed over the stream",
                "mean_s_by_n_grid": drv["mean_s"]},
            "capacity_vs_pmesh_driver": t_drv / t_batch}


def driver_sample(reqs, per_size):
    """Run time of requests run the way pmesh.py runs one simulation (host ICs, one Python
    call per step with the state resident), on the first `per_size` requests of each size."""
    import contextlib
    import initial_conditions as ic
    from pmgpu.sim import PMSim
    times = {}
    for r in reqs:
        if len(times.setdefault(r.n_cells, [])) >= per_size:
            continue
        t0 = time.perf_counter()
        p = r.ic_params()
        with contextlib.redirect_stdout(sys.stderr):
            pos, vel = ic.zeldovich_numpy(ic.gaussian_random_field(p), p)
        sim = PMSim(r.n_cells, r.n_parts, 32, omega_m0=r.om, omega_lambda0=r.ol, omega_k0=r.ok,
                    h0=r.h)
        sim.set_state(pos, vel)
        a, da = r.a_init, r.da
        while a < r.a_end - da:                   # pmesh.py:56-63
            sim.step(a, da)
            a += da
        sim.power_spectrum(r.box)
        sim.close()
        times[r.n_cells].append(time.perf_counter() - t0)
    return {"per_size": per_size, "mean_s": {n: float(np.mean(t)) for n, t in times.items()}}


def pk_error(sim32, sim64, box):
    """max |P32/P64 - 1| over k < k_Nyquist/2, both spectra from the device estimator."""
    _, p32, _ = sim32.power_spectrum(box)
    _, p64, _ = sim64.power_spectrum(box)
    n = sim32.n_cells
    sel = np.arange(1, len(p32) + 1) <= n // 4          # k < k_Nyquist / 2
    return float(np.max(np.abs(p32[sel] / p64[sel] - 1)))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(pos, vel, w, min_steps, budget_s):
    """The CPU port on this host's cores: one untimed warm-up step, then at least `min_steps`
    timed steps (more while the budget lasts); the median step time and the per-stage medians."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pm_oracle as O
    import pm_port
    lib = pm_port.load()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or pm_port.cores()
    threads = max(1, min(threads, pm_port.cores()))
    lib.port_set_threads(threads)
    n = w["n_cells"]
    fgrid = O.fourier_grid(n, ref_dtypes=True)
    mass = (n / w["n_parts"]) ** 3
    da = (w["a_end"] - w["a_init"]) / w["steps"]
    x, v = pos.copy(), vel.copy()
    a = w["a_init"]
    t0 = time.perf_counter()
    pm_port.step(x, v, fgrid, mass, a, da, w["om"], w["ol"], w["ok"], w["h"], threads)  # warm-up
    a += da
    log(f"cpu baseline warm-up step: {time.perf_counter() - t0:.1f} s")
    steps, stages = [], []
    t0 = time.perf_counter()
    while len(steps) < min_steps or time.perf_counter() - t0 < budget_s * 0.5:
        st = {}
        ts = time.perf_counter()
        pm_port.step(x, v, fgrid, mass, a, da, w["om"], w["ol"], w["ok"], w["h"], threads, st)
        steps.append(time.perf_counter() - ts)
        stages.append(st)
        a += da
        log(f"cpu baseline step {len(steps)}: {steps[-1]:.2f} s")
        if time.perf_counter() - t0 > budget_s:
            break
    med = float(np.median(steps))
    n_p = w["n_parts"] ** 3
    return {"value": n_p / med, "unit": "particle-updates/s", "cores": threads,
            "kind": "port", "cpu": cpu_model(),
            "sample": f"median of {len(steps)} PM steps after 1 warm-up step, the same "
                      f"256^3/512^3 workload from the same initial conditions (C/OpenMP port of "
                      f"density.py/integrate.py + scipy.fft complex128 c2c standing in for FFTW)",
            "median_s_per_step": med, "min_s_per_step": float(np.min(steps)),
            "stage_median_s": {k: float(np.median([st[k] for st in stages])) for k in stages[0]}}


def copy_ceiling(gib=2, reps=5):
    """Device-to-device copy of `gib` GiB (hipMemcpy, a copy kernel on this device): the
    read + write bandwidth a streaming kernel can reach, measured in this run."""
    from pmgpu import _lib
    from pmgpu.device import DeviceArray
    nbytes = gib << 30
    src = DeviceArray.zeros((nbytes // 4,), np.float32)
    dst = DeviceArray.empty((nbytes // 4,), np.float32)
    _lib.call("pm_memcpy_dtod", dst.ptr, src.ptr, nbytes)     # warm-up
    _lib.call("pm_synchronize")
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        _lib.call("pm_memcpy_dtod", dst.ptr, src.ptr, nbytes)
        _lib.call("pm_synchronize")          # a device-to-device copy returns before it ends
        best = min(best, time.perf_counter() - t0)
    del src, dst
    gbs = 2 * nbytes / best / 1e9
    return {"bytes_per_copy": 2 * nbytes, "GBs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS,
            "how": f"hipMemcpy device-to-device of {gib} GiB, best of {reps}, read + write bytes"}


def late_time(pos, vel, w, steps, tiles):
    """The same number of FP32 steps timed at the end of the run (a -> 1), where clustering is
    strongest: the run is advanced untimed to `steps` before its end, then timed."""
    from pmgpu.sim import PMSim
    sim = PMSim(w["n_cells"], w["n_parts"], 32, omega_m0=w["om"], omega_lambda0=w["ol"],
                omega_k0=w["ok"], h0=w["h"])
    sim.set_tiles(*tiles)
    sim.set_state(pos, vel)
    da = (w["a_end"] - w["a_init"]) / w["steps"]
    total = loop_steps(w["a_init"], w["a_end"], w["steps"])
    a = sim.run(w["a_init"], da, total - steps)
    sim.sync()
    t0 = time.perf_counter()
    a_end = sim.run(a, da, steps)
    sim.sync()
    dt = time.perf_counter() - t0
    sim.close()
    ms = dt / steps * 1e3
    return {"steps": steps, "a_start": a, "a_end": a_end, "ms_per_step": ms,
            "value": w["n_parts"] ** 3 / (ms * 1e-3),
            "step_roofline_frac": step_bytes(w["n_cells"], w["n_parts"] ** 3, 32) / (ms * 1e-3) / 1e9
            / HBM_PEAK_GBS}


def traffic_from_profiles(kernel):
    """HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, if committed."""
    f = ROOT / "profiles" / "hbm_traffic_512.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text()).get(kernel)
        return None if d is None else float(d["bytes"])
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=WORKLOAD["steps"])
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--timing-steps", type=int, default=50)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=10)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--largest-steps", type=int, default=-1,
                    help="steps of the 512^3 / 1024^3 run to time (-1: the whole run, 0 skips it)")
    ap.add_argument("--serve-window", type=float, default=60.0,
                    help="seconds of the 600 s serving stream to replay (600 = all of it)")
    ap.add_argument("--ensemble-steps", type=int, default=None)
    ap.add_argument("--tiles", type=lambda t: tuple(int(x) for x in t.split(",")), default=None,
                    help="fixed tile shape TX,TY,TZ of the 256^3/512^3 runs instead of the autotuner's")
    ap.add_argument("--no-late", action="store_true",
                    help="skip the late-time window (a whole run to a = 1)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device (a known-size kernel for "
                         "calibrating FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); a multi-GPU PM step "
              "is not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy_mib:
        from pmgpu.device import DeviceArray
        nbytes = args.calibrate_copy_mib << 20
        src = DeviceArray.zeros((nbytes // 4,), np.float32)
        dst = DeviceArray.empty((nbytes // 4,), np.float32)
        from pmgpu import _lib
        for _ in range(3):
            _lib.call("pm_memcpy_dtod", dst.ptr, src.ptr, nbytes)
        log(f"calibration copy: 3 x {args.calibrate_copy_mib} MiB device-to-device")
        del src, dst
    w = WORKLOAD
    n, n_p = w["n_cells"], w["n_parts"] ** 3
    log("initial conditions (CPU, seeded)")
    pos, vel = make_ics(w)

    log(f"FP32 GPU: {args.warmup} warm-up + {args.steps} timed steps")
    sim32, dt32, a_end = run_gpu(32, pos, vel, w, args.steps, args.warmup, args.tiles)
    log(f"FP32: {dt32:.3f} s, a_end {a_end:.4f}")
    fp64_steps = args.fp64_steps or args.steps
    log(f"FP64 GPU: {args.warmup} warm-up + {fp64_steps} timed steps")
    sim64, dt64, _ = run_gpu(64, pos, vel, w, fp64_steps, args.warmup, args.tiles)
    log(f"FP64: {dt64:.3f} s")
    pk_err = None
    if fp64_steps == args.steps:
        pk_err = pk_error(sim32, sim64, w["box"])
    tiles32, tiles64 = sim32.tuned[0], sim64.tuned[0]
    sim32.close()
    sim64.close()

    log(f"per-stage HIP-event timing over {args.timing_steps} FP32 steps")
    stage_ms, stage_calls = stage_timing(pos, vel, w, args.timing_steps, tiles32)
    sb = stage_bytes(n, n_p, 32)
    dom = max(stage_ms, key=lambda k: stage_ms[k])
    ach = sb[dom] / (stage_ms[dom] * 1e-3) / 1e9
    v32 = n_p * args.steps / dt32
    v64 = n_p * fp64_steps / dt64
    ms32 = dt32 / args.steps * 1e3
    out = {
        "metric": METRIC,
        "value": v32,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "pm_step_256p_512g: 256^3 particles, 512^3 grid, box 250 Mpc/h, "
                               "seed 7, EH no-wiggle sigma8 0.8, a 0.02 -> 1",
                   "n_parts": w["n_parts"], "n_cells": n, "box": w["box"],
                   "a_init": w["a_init"], "da": (w["a_end"] - w["a_init"]) / w["steps"],
                   "parallelism": "single GPU", "tiles_fp32": list(tiles32),
                   "tiles_fp64": list(tiles64)},
        "fp64": {"value": v64, "ms_per_step": dt64 / fp64_steps * 1e3, "steps": fp64_steps},
        "fp32_over_fp64": v32 / v64,
        "pk_max_rel_err_fp32_vs_fp64_k_lt_half_nyquist": pk_err,
        "step_roofline": {"bytes_per_step": step_bytes(n, n_p, 32),
                          "achieved_GBs": step_bytes(n, n_p, 32) / (ms32 * 1e-3) / 1e9,
                          "frac": step_bytes(n, n_p, 32) / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "stages_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "stages_GBs": {k: round(sb[k] / (stage_ms[k] * 1e-3) / 1e9, 1) for k in stage_ms},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                     "traffic": traffic_from_profiles(dom),
                     "algorithmic_bytes_per_launch": sb[dom],
                     "avg_launch_ms": stage_ms[dom], "launches": stage_calls[dom]},
        "vs_reference_readme_cpu": v32 / 4.66e6,
    }
    log("copy ceiling")
    out["copy_ceiling"] = copy_ceiling()
    out["roofline"]["frac_of_copy_ceiling"] = ach / out["copy_ceiling"]["GBs"]
    out["step_roofline"]["frac_of_copy_ceiling"] = (out["step_roofline"]["achieved_GBs"]
                                                    / out["copy_ceiling"]["GBs"])
    total = loop_steps(w["a_init"], w["a_end"], w["steps"])
    if args.steps < total and not args.no_late:
        log(f"late-time window: the last {args.steps} steps of the run")
        out["late_time"] = late_time(pos, vel, w, args.steps, tiles32)
    log("plain FP64 GPU path (reference-layout entry points)")
    out["plain_fp64"] = plain_fp64_bench(pos, vel, w, args.plain_steps)
    out["fp32_over_plain_fp64"] = v32 / out["plain_fp64"]["value"]
    out["ensemble"] = None if args.no_ensemble else ensemble_bench(ENSEMBLE, args.ensemble_steps)
    out["serve"] = None if args.no_serve else serve_bench(SERVE, args.serve_window)
    if args.largest_steps != 0:
        log("largest run: " + ("all" if args.largest_steps < 0 else str(args.largest_steps))
            + " steps of 512^3 particles on 1024^3")
        out["largest"] = largest_bench(LARGEST, args.largest_steps)
    else:
        out["largest"] = None
    if not args.no_cpu:
        log("CPU baseline (port) on host cores")
        out["cpu_baseline"] = cpu_baseline(pos, vel, w, args.cpu_steps, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
