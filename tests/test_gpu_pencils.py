"""Pencil shapes of the y / z FFT passes: the LDS (Stockham) passes have an alternative shape with
half the columns per block, chosen per pass by the autotuner or the caller.  The shape changes
only how many columns a block transforms together, not the arithmetic of a column, so a step in
either shape gives the same potential and velocities (checked here to 1e-6 at FP32, 1e-13 at
FP64, with the order-independent deposit) and the float64 oracle is met either way.
"""
import numpy as np
import pytest

import pm_oracle as O
from helpers import make_ics, rel

pytestmark = pytest.mark.gpu

KW = dict(omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)


def _step(n, npart, prec, pos, vel, pencils=None):
    from pmgpu.sim import PMSim
    s = PMSim(n, npart, prec, deterministic=True, **KW)
    if pencils:
        s.set_pencils(*pencils)
    shape = s.pencils()
    s.set_state(pos, vel)
    s.step(0.05, 0.01)
    out = (s.potential(np.float64), *s.get_state(np.float64))
    s.close()
    return shape, out


@pytest.mark.parametrize("n,npart,prec", [(384, 96, 32), (192, 48, 32), (96, 48, 32), (128, 32, 64)])
def test_alternative_pencil_shapes_give_the_same_step(n, npart, prec):
    from pmgpu.sim import PMSim
    pos, vel = make_ics(n, npart, seed=71, box=float(2 * n))
    base, ref = _step(n, npart, prec, pos, vel)
    alt = tuple(w // 2 for w in base)
    shape, got = _step(n, npart, prec, pos, vel, alt)
    assert shape == alt
    tol = 1e-6 if prec == 32 else 1e-13
    errs = [rel(a, b) for a, b in zip(got, ref)]
    print(f"N={n} FP{prec} pencils {base} -> {alt}: phi {errs[0]:.1e} x {errs[1]:.1e} v {errs[2]:.1e}"
          f"{' (bit-identical)' if all(np.array_equal(a, b) for a, b in zip(got, ref)) else ''}")
    assert max(errs) <= tol
    # and the oracle's potential from the same density
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
    rho = O.density(O.wrap_positions(pos.astype(np.float64), n), p.mass, n)
    phi = O.potential(rho, O.fourier_grid(n), 0.05, 0.3)
    assert rel(got[0], phi) < (1e-5 if prec == 32 else 1e-10)


def test_pencil_shapes_are_validated_and_autotuned():
    from pmgpu.sim import PMSim
    s = PMSim(512, 64, 32, **KW)             # register passes at 512^3: one shape each
    assert s.pencils() == (16, 16, 16)
    with pytest.raises(RuntimeError, match="columns per block"):
        s.set_pencils(8, 8, 8)
    s.close()
    n, npart = 384, 96
    pos, vel = make_ics(n, npart, seed=72, box=float(2 * n))
    s = PMSim(n, npart, 32, **KW)
    with pytest.raises(RuntimeError, match="columns per block"):
        s.set_pencils(4, 0, 0)
    s.set_state(pos, vel)
    s.autotune(2)
    chosen = s.pencils()
    assert all(w in (8, 16) for w in chosen)
    s.run(0.05, 0.01, 10)
    x, v = s.get_state(np.float64)
    s.close()
    ref = PMSim(n, npart, 32, **KW)
    ref.set_state(pos, vel)
    ref.run(0.05, 0.01, 10)
    xr, vr = ref.get_state(np.float64)
    ref.close()
    print(f"N=384 autotuned pencils {chosen}: 10 steps vs default shapes, v {rel(v, vr):.1e}")
    assert rel(v, vr) < 1e-5
