"""Simulation driver — the reference's pmesh.py:18-93 step loop, unchanged in structure.

``python pmesh.py`` runs the simulation configured in configure_me.py.  With the default
configure_me.RESIDENT the particles and grids stay on the GPU for the whole loop; the host
only sees them when a snapshot is written.
"""
from time import time

import configure_me as _c
from density import density
from integrate import advance_time
from zeldovich import zeldovich, new_simulation
from fourier_utils import fourier_grid
from gaussian_random_field import gaussian_random_field
from cosmology import *  # noqa: F401,F403  (the reference imports the same names)
from save_data import save_file, from_file
from plot_helper import plot_step, plot_grf, plot_projection


def simulator():
    N_CELLS, N_PARTS, BOX_SIZE = _c.N_CELLS, _c.N_PARTS, _c.BOX_SIZE
    OMEGA_M0, H0, A_INIT, A_END = _c.OMEGA_M0, _c.H0, _c.A_INIT, _c.A_END
    print('Starting the simulation for {}^3 particles'.format(N_PARTS), 'with {}^3 grid cells'.format(N_CELLS))

    #Physical particle mass in solar masses
    particle_mass = 1.32*10**5*(OMEGA_M0*H0**2)*(BOX_SIZE/(N_PARTS/128))**3
    print('Particle mass in solar mass for this configuration: {:.3E}'.format(particle_mass))

    dens_contrast = (N_CELLS/N_PARTS)**3
    print('Particle mass in code units: {:.3e}'.format(dens_contrast))
    da = (A_END - A_INIT)/_c.STEPS
    da_save = (A_END - A_INIT)/_c.N_SAVE_FILES
    da_plot = (A_END - A_INIT)/_c.N_PLOTS
    a_current = A_INIT
    n_file = 0
    n_plot = 0

    if _c.RESTART:
        print("Restarting from file data.{}".format(_c.RESTART_FROM_N))
        n_file = _c.RESTART_FROM_N
        n_plot = (_c.RESTART_FROM_N)/_c.N_SAVE_FILES*_c.N_PLOTS
        positions, velocities, a_current = from_file(n_file)
        a_current = float(a_current)
        if getattr(_c, "RESIDENT", True):
            positions, velocities = new_simulation(positions, velocities)
    else:
        rho = gaussian_random_field()
        positions, velocities = zeldovich(rho)
        if _c.SAVE_DATA:
            save_file(rho, positions, velocities, 0, a_current)
            n_file += 1
        if _c.PLOT_GRF:
            plot_grf(rho)

    ksq_inverse = fourier_grid()
    print('Starting the integrations...')
    steps = 0
    while a_current<A_END-da:

        start_time = time()

        rho = density(positions, dens_contrast)
        positions, velocities = advance_time(rho, positions, velocities, ksq_inverse, a_current, da)

        a_current += da
        steps += 1

        if a_current >= A_INIT + n_file*da_save:
            if _c.SAVE_DATA:
                save_file(rho, positions, velocities, n_file, a_current)
            n_file += 1
        if a_current >= A_INIT + n_plot*da_plot:
            if _c.PLOT_STEPS:
                plot_step(rho, n_plot)
            if _c.PLOT_PROJECTIONS:
                plot_projection(rho, n_plot, 15)
            n_plot += 1

        if _c.PRINT_STATUS:
            print_status(a_current, start_time)

    return positions, velocities, a_current, steps


def print_status(a_current, start_time):

    percentile = 100*(a_current-_c.A_INIT)/(_c.A_END-_c.A_INIT)
    print("%.5f" % percentile,"%", " Save step time: --- %.5f seconds ---" % (time() - start_time))


if __name__ == "__main__":

    start_time = time()

    simulator()
    print("Finished in")
    print("--- %.2f seconds ---" % (time() - start_time))
