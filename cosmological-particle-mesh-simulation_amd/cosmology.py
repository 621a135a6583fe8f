"""Background cosmology scalars (reference: cosmology.py:1-36).

Host-side float64 scalars.  ``f`` sets the per-step kick/drift coefficient of the PM step
(integrate.py:12); ``H`` and ``Dt`` are used by the initial conditions only.
"""
import math

import configure_me


def H(a, H0, cosmology):
    """Hubble rate, cosmology.py:11-18: sqrt(H0^2 (Om/a^3 + Ok/a^2 + OL))."""
    omega_m0, omega_l0, omega_k0 = cosmology
    return math.sqrt(H0 ** 2 * (omega_m0 / a ** 3 + omega_k0 / a ** 2 + omega_l0))


def f(a, cosmology):
    """cosmology.py:20-27: 1/sqrt((c0 + c2*a + c1*a^3)/a) for cosmology = [c0, c1, c2]."""
    omegaM, omegaL, omegaK = cosmology
    return 1 / math.sqrt((omegaM + omegaK * a + omegaL * a ** 3) / a)


def Dt(a, cosmology):
    """Linear growth approximation, cosmology.py:29-36."""
    omegaM, omegaL, omegaK = cosmology
    return 5 / 2 / omegaM / (omegaM ** (4 / 7) - omegaL + (1 + omegaM / 2) * (1 + omegaL / 70)) * a


class cosmo:
    """cosmology.py:5-9."""

    def __init__(self):
        c = configure_me
        cos = [c.OMEGA_M0, c.OMEGA_LAMBDA0, c.OMEGA_K0]
        self.Dt = Dt(c.A_INIT, cos)
        self.H01 = H(c.A_INIT, c.H0, cos)
        self.f0 = f(c.A_INIT, cos)


def step_coefficient_cosmology():
    """The list advance_time passes to f(): [H0, OL, OK] as in integrate.py:12, or the
    physical [Om, OL, OK] when configure_me.FIX_F_COSMOLOGY is set."""
    c = configure_me
    first = c.OMEGA_M0 if getattr(c, "FIX_F_COSMOLOGY", False) else c.H0
    return [first, c.OMEGA_LAMBDA0, c.OMEGA_K0]
