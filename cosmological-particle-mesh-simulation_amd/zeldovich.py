"""Reference name for the Zel'dovich initial conditions (zeldovich.py:10).

``zeldovich(density)`` returns (positions, velocities) in the reference layout (3, N_PARTS^3).
With configure_me.RESIDENT (the default) they are views of a new device-resident simulation,
so the PM loop in pmesh.py never copies particles to the host; otherwise NumPy float32 arrays.
"""
import configure_me
from initial_conditions import zeldovich_numpy
from pmgpu.sim import PMSim


def new_simulation(positions, velocities):
    c = configure_me
    sim = PMSim(c.N_CELLS, c.N_PARTS, int(getattr(c, "PRECISION", 32)),
                cic_deconvolve=bool(getattr(c, "CIC_DECONVOLVE", False)),
                fix_f_cosmology=bool(getattr(c, "FIX_F_COSMOLOGY", False)),
                omega_m0=c.OMEGA_M0, omega_lambda0=c.OMEGA_LAMBDA0, omega_k0=c.OMEGA_K0, h0=c.H0)
    sim.set_state(positions, velocities)
    return sim.views()


def zeldovich(density):
    pos, vel = zeldovich_numpy(density)
    if getattr(configure_me, "RESIDENT", True):
        return new_simulation(pos, vel)
    return pos, vel
