"""Reference name for the initial density field (gaussian_random_field.py:9).

Returns the seeded linear density contrast of ``initial_conditions`` (Eisenstein-Hu no-wiggle
P(k), sigma_8 normalisation) on the N_PARTS^3 lattice at A_INIT.
"""
from initial_conditions import gaussian_random_field

__all__ = ["gaussian_random_field"]
