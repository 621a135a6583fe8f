"""The C/OpenMP CPU port (bench baseline) agrees with the float64 oracle."""
import numpy as np
import pytest

import pm_oracle as O

pm_port = pytest.importorskip("pm_port")


@pytest.fixture(scope="module", autouse=True)
def lib():
    try:
        pm_port.load()
    except RuntimeError as e:
        pytest.skip(str(e))


def test_port_step_matches_oracle():
    from helpers import make_ics, periodic_diff, rel
    n, npart = 32, 16
    pos, vel = make_ics(n, npart, seed=2)
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
    a, da = 0.02, 0.0098
    rho_ref = O.density(pos, p.mass, n)
    xr, vr = O.advance_time(rho_ref, pos, vel, O.fourier_grid(n), a, da, p)
    x, v = pos.copy(), vel.copy()
    rho, phi = pm_port.step(x, v, O.fourier_grid(n, ref_dtypes=True), p.mass, a, da, 0.3, 0.7, 0.0,
                            0.7, workers=2)
    assert rel(rho, rho_ref) < 1e-6
    assert rel(v - vel, vr - vel) < 1e-5
    assert np.abs(periodic_diff(x, xr, n)).max() < 1e-4
