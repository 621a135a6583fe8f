"""Command-line driver: build (or reload) the particles, then advance them with the PM step.

``python pmesh.py`` reads configure_me.py and runs one simulation.  What the driver guarantees
(the behaviour of the reference's driver, pmesh.py:18-93):

* a step is ``density(positions, mass)`` followed by ``advance_time(...)``, with the code-unit
  particle mass (N_CELLS/N_PARTS)^3 that makes the mean density 1 (pmesh.py:28);
* the scale factor grows by da = (A_END - A_INIT)/STEPS, accumulated in float64, while
  a < A_END - da, so rounding decides whether STEPS - 1 or STEPS steps run (pmesh.py:56,63);
* output j is due after the first step that brings a to A_INIT + j*(A_END - A_INIT)/N or past
  it (N = N_SAVE_FILES for snapshots, N_PLOTS for plots); snapshot 0 holds the initial
  conditions; the density written or drawn with output j is the one that step solved for;
* a restart from snapshot n continues with the snapshot counter at n and the plot counter at
  n*N_PLOTS/N_SAVE_FILES.

With configure_me.RESIDENT (the default) the particles stay on the GPU for the whole run and
reach the host only when a snapshot is written.
"""
from __future__ import annotations

import time

import configure_me as cfg
from density import density
from fourier_utils import fourier_grid
from gaussian_random_field import gaussian_random_field
from integrate import advance_time
from plot_helper import plot_grf, plot_projection, plot_step
from save_data import from_file, save_file
from zeldovich import new_simulation, zeldovich

# h^2 M_sun / Mpc^3: critical density today, for the informational particle-mass line
RHO_CRIT_H2 = 2.775e11


class Schedule:
    """Outputs spaced evenly in scale factor: ``due(a)`` is true once for every mark
    a_first + j*spacing that a has reached, counting j from ``start``."""

    def __init__(self, a_first: float, a_last: float, count: int, start: float = 0):
        self.a_first = a_first
        self.spacing = (a_last - a_first) / count
        self.index = start

    def due(self, a: float) -> bool:
        if a >= self.a_first + self.index * self.spacing:
            self.index += 1
            return True
        return False


def _resident() -> bool:
    return bool(getattr(cfg, "RESIDENT", True))


def starting_point():
    """(positions, velocities, a, first snapshot number, first plot number)."""
    if cfg.RESTART:
        n = cfg.RESTART_FROM_N
        print(f"Resuming from snapshot {n}")
        pos, vel, a = from_file(n)
        if _resident():
            pos, vel = new_simulation(pos, vel)
        return pos, vel, float(a), n, n / cfg.N_SAVE_FILES * cfg.N_PLOTS
    field = gaussian_random_field()
    pos, vel = zeldovich(field)
    first = 0
    if cfg.SAVE_DATA:
        save_file(field, pos, vel, 0, cfg.A_INIT)
        first = 1
    if cfg.PLOT_GRF:
        plot_grf(field)
    return pos, vel, cfg.A_INIT, first, 0


def describe_run() -> None:
    n_p = cfg.N_PARTS ** 3
    m_sun = RHO_CRIT_H2 * cfg.OMEGA_M0 * cfg.H0 ** 2 * cfg.BOX_SIZE ** 3 / n_p
    print(f"PM run: {cfg.N_PARTS}^3 particles, {cfg.N_CELLS}^3 cells, box {cfg.BOX_SIZE}, "
          f"a {cfg.A_INIT} -> {cfg.A_END} in {cfg.STEPS} steps")
    print(f"  particle mass {m_sun:.3E} M_sun, {(cfg.N_CELLS / cfg.N_PARTS) ** 3:.3e} in code units")


def simulator():
    """Runs the configured simulation; returns (positions, velocities, final a, steps taken)."""
    describe_run()
    mass = (cfg.N_CELLS / cfg.N_PARTS) ** 3
    da = (cfg.A_END - cfg.A_INIT) / cfg.STEPS
    positions, velocities, a, first_file, first_plot = starting_point()
    snapshots = Schedule(cfg.A_INIT, cfg.A_END, cfg.N_SAVE_FILES, first_file)
    plots = Schedule(cfg.A_INIT, cfg.A_END, cfg.N_PLOTS, first_plot)
    green = fourier_grid()
    span = cfg.A_END - cfg.A_INIT
    steps = 0
    print("Stepping...")
    while a < cfg.A_END - da:
        t0 = time.perf_counter()
        rho = density(positions, mass)
        positions, velocities = advance_time(rho, positions, velocities, green, a, da)
        a += da
        steps += 1
        if snapshots.due(a) and cfg.SAVE_DATA:
            save_file(rho, positions, velocities, snapshots.index - 1, a)
        if plots.due(a):
            if cfg.PLOT_STEPS:
                plot_step(rho, plots.index - 1)
            if cfg.PLOT_PROJECTIONS:
                plot_projection(rho, plots.index - 1, 15)
        if cfg.PRINT_STATUS:
            print(f"step {steps:5d}  a = {a:.5f}  ({100 * (a - cfg.A_INIT) / span:6.2f} %)  "
                  f"{1e3 * (time.perf_counter() - t0):9.3f} ms")
    return positions, velocities, a, steps


if __name__ == "__main__":
    t_start = time.perf_counter()
    simulator()
    print(f"Done in {time.perf_counter() - t_start:.2f} s")
