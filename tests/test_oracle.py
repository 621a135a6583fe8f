"""Pins for the float64 CPU oracle (oracle/pm_oracle.py).

The reference ships no tests, fixtures or golden vectors, and its own path cannot run here
(numba and pyFFTW are absent).  Each oracle function is therefore pinned by checks that do not
restate its formula: closed forms, conservation laws, symmetries and brute force on tiny
inputs, chosen so that a dropped term, a wrong sign, index or axis order fails one of them.

    function              pinned by
    density               test_deposit_* (closed form, mass, integer/edge positions, roll, brute force)
    fourier_grid          test_green_* (per-mode closed form, symmetry)
    density_k             test_dft_direct, test_dft_parseval
    potential_k/_real     test_poisson_plane_wave, test_poisson_laplacian
    weights/sweep/integ.  test_gather_plane_wave, test_zero_self_force, test_momentum,
                          test_pair_attraction, test_free_drift, test_axis_order
    f / advance_time      test_f_*, test_advance_time_uses_h0_slot
    step_count            test_step_count
"""
import math

import numpy as np
import pytest

import pm_oracle as O


def rng(seed=0):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------------------------- deposit
def test_deposit_single_particle_closed_form():
    n = 8
    x, y, z = 2.3, 5.75, 7.5
    rho = O.density(np.array([[x], [y], [z]]), 2.0, n)
    expect = np.zeros((n, n, n))
    for cz, wz in ((7, 0.5), (0, 0.5)):
        for cy, wy in ((5, 0.25), (6, 0.75)):
            for cx, wx in ((2, 0.7), (3, 0.3)):
                expect[cz, cy, cx] += 2.0 * wx * wy * wz
    np.testing.assert_allclose(rho, expect, atol=1e-14)


def test_deposit_mass_conservation():
    n = 16
    pos = rng(1).uniform(0, n, (3, 5000))
    rho = O.density(pos, 0.37, n)
    assert abs(rho.sum() - 0.37 * 5000) < 1e-9


def test_deposit_integer_position_hits_one_cell():
    n = 8
    rho = O.density(np.array([[3.0], [1.0], [6.0]]), 1.0, n)
    assert rho[6, 1, 3] == 1.0 and rho.sum() == 1.0


def test_deposit_periodic_edge_and_q4():
    n = 8
    rho = O.density(np.array([[7.25], [0.0], [0.0]]), 1.0, n)
    assert rho[0, 0, 7] == 0.75 and rho[0, 0, 0] == 0.25
    # a float32 position equal to N (Q4) is the same particle as 0
    rho_n = O.density(np.array([[8.0], [0.0], [0.0]], dtype=np.float32), 1.0, n)
    assert rho_n[0, 0, 0] == 1.0


def test_deposit_translation_is_roll():
    n = 12
    pos = rng(2).uniform(0, n, (3, 300))
    shift = np.array([[3.0], [-2.0], [5.0]])
    rho = O.density(pos, 1.0, n)
    rho2 = O.density(np.mod(pos + shift, n), 1.0, n)
    np.testing.assert_allclose(rho2, np.roll(rho, (5, -2, 3), axis=(0, 1, 2)), atol=1e-12)


def test_deposit_brute_force():
    n = 6
    pos = rng(3).uniform(0, n, (3, 20))
    expect = np.zeros((n, n, n))
    for p in range(pos.shape[1]):
        for cz in range(n):
            for cy in range(n):
                for cx in range(n):
                    w = 1.0
                    for c, xx in ((cx, pos[0, p]), (cy, pos[1, p]), (cz, pos[2, p])):
                        dist = abs(((xx - c + n / 2) % n) - n / 2)   # periodic distance
                        w *= max(0.0, 1.0 - dist)
                    expect[cz, cy, cx] += w
    np.testing.assert_allclose(O.density(pos, 1.0, n), expect, atol=1e-12)


# --------------------------------------------------------------------------------- Green
def test_green_closed_form_and_symmetry():
    n = 10
    g = O.fourier_grid(n)
    assert g[0, 0, 0] == 0.0
    for (i, j, l) in [(1, 0, 0), (3, 7, 2), (9, 5, 5)]:
        k = lambda m: 2 * math.pi * (m if m < n / 2 else m - n) / n
        val = 1 / (math.sin(k(i) / 2) ** 2 + math.sin(k(j) / 2) ** 2 + math.sin(k(l) / 2) ** 2)
        assert abs(g[i, j, l] - val) < 1e-12
    np.testing.assert_allclose(g, g.transpose(2, 1, 0))
    np.testing.assert_allclose(g, np.roll(g[::-1, ::-1, ::-1], 1, axis=(0, 1, 2)))


def test_green_ref_dtype_gap():
    g64 = O.fourier_grid(16)
    g32 = O.fourier_grid(16, ref_dtypes=True)
    assert g32.dtype == np.float32
    assert np.max(np.abs(g32 - g64) / np.maximum(g64, 1e-30)) < 1e-6


# ------------------------------------------------------------------------------------ FFT
def test_dft_direct():
    n = 4
    x = rng(4).normal(size=(n, n, n))
    out = O.density_k(x)
    idx = np.arange(n)
    for (a, b, c) in [(0, 0, 0), (1, 2, 3), (3, 1, 0)]:
        ph = np.exp(-2j * np.pi * (a * idx[:, None, None] + b * idx[None, :, None]
                                    + c * idx[None, None, :]) / n)
        assert abs(out[a, b, c] - (x * ph).sum()) < 1e-10


def test_dft_parseval_and_inverse():
    n = 8
    x = rng(5).normal(size=(n, n, n))
    X = O.density_k(x)
    assert abs((np.abs(X) ** 2).sum() / n ** 3 - (x ** 2).sum()) < 1e-9
    np.testing.assert_allclose(O.potential_real(X), x, atol=1e-12)


# -------------------------------------------------------------------------------- Poisson
def test_poisson_plane_wave():
    n, a, om = 16, 0.5, 0.3
    m = (2, 3, 1)   # mode numbers along axes (0, 1, 2)
    zz, yy, xx = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    arg = 2 * np.pi * (m[0] * zz + m[1] * yy + m[2] * xx) / n
    rho = 1.0 + 0.1 * np.cos(arg)
    phi = O.potential(rho, O.fourier_grid(n), a, om)
    S = sum(math.sin(math.pi * mi / n) ** 2 for mi in m)
    np.testing.assert_allclose(phi, -3 * om / (8 * a) * 0.1 * np.cos(arg) / S, atol=1e-12)


def test_poisson_laplacian():
    # -4 sum sin^2(k/2) is the eigenvalue of the 7-point Laplacian, so lap7(phi) = 3 Om/(2a) (rho - <rho>)
    n, a, om = 12, 0.2, 0.31
    rho = rng(6).uniform(0, 2, (n, n, n))
    phi = O.potential(rho, O.fourier_grid(n), a, om)
    lap = sum(np.roll(phi, s, axis=ax) for ax in range(3) for s in (1, -1)) - 6 * phi
    np.testing.assert_allclose(lap, 1.5 * om / a * (rho - rho.mean()), atol=1e-10)


# --------------------------------------------------------------------------------- gather
def _gather(pos, phi, n):
    """accelerations from the oracle with da*fa1 = 1 and no drift contribution"""
    x = np.asarray(pos, dtype=np.float64)
    v0 = np.zeros_like(x)
    _, v = O.integrate(x, v0, a_val=1.0, f_a1=1.0, da=1.0, phi=phi, n_cells=n)
    return v   # v = da*fa1*g = g


def test_gather_plane_wave():
    n = 16
    m = (1, 2, 3)  # along grid axes (z, y, x)
    zz, yy, xx = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    arg = 2 * np.pi * (m[0] * zz + m[1] * yy + m[2] * xx) / n
    phi = np.cos(arg)
    pos = rng(7).uniform(0, n, (3, 50))
    g = _gather(pos, phi, n)
    # central difference of cos: g_d(c) = sin(2 pi m_d/N) sin(arg(c)); CIC-interpolate by hand
    for p in range(pos.shape[1]):
        x, y, z = pos[:, p]
        cx, cy, cz = int(x), int(y), int(z)
        ref = np.zeros(3)
        for oz in (0, 1):
            for oy in (0, 1):
                for ox in (0, 1):
                    w = ((x - cx) if ox else 1 - (x - cx)) * ((y - cy) if oy else 1 - (y - cy)) \
                        * ((z - cz) if oz else 1 - (z - cz))
                    c_arg = 2 * np.pi * (m[0] * (cz + oz) + m[1] * (cy + oy) + m[2] * (cx + ox)) / n
                    # positions[0] is grid axis 2 (mode m[2]), positions[2] is axis 0 (m[0])
                    ref += w * np.sin(c_arg) * np.array([math.sin(2 * math.pi * m[2] / n),
                                                         math.sin(2 * math.pi * m[1] / n),
                                                         math.sin(2 * math.pi * m[0] / n)])
        np.testing.assert_allclose(g[:, p], ref, atol=1e-12)


def _accel(pos, n, om=0.3, a=1.0):
    rho = O.density(pos, 1.0, n)
    phi = O.potential(rho, O.fourier_grid(n), a, om)
    return _gather(pos, phi, n)


def test_zero_self_force():
    n = 16
    for x in ([3.3, 7.9, 0.2], [0.0, 15.999, 8.5], [4.5, 4.5, 4.5]):
        g = _accel(np.array(x)[:, None], n)
        assert np.max(np.abs(g)) < 1e-12


def test_momentum_conservation():
    n = 16
    pos = rng(8).uniform(0, n, (3, 400))
    g = _accel(pos, n)
    assert np.max(np.abs(g.sum(axis=1))) < 1e-10 * np.abs(g).sum()


def test_pair_attraction():
    n = 32
    pos = np.array([[10.2, 14.7], [16.1, 16.3], [5.5, 5.4]])
    g = _accel(pos, n)
    sep = pos[:, 1] - pos[:, 0]
    assert np.dot(g[:, 0], sep) > 0 and np.dot(g[:, 1], -sep) > 0
    np.testing.assert_allclose(g[:, 0], -g[:, 1], atol=1e-12)


def test_axis_order():
    # a potential varying only along grid axis 2 (x) accelerates positions[0] only
    n = 16
    xx = np.arange(n)[None, None, :] * np.ones((n, n, 1))
    phi = np.cos(2 * np.pi * xx / n)
    g = _gather(rng(9).uniform(0, n, (3, 20)), phi, n)
    assert np.abs(g[0]).max() > 1e-3 and np.abs(g[1:]).max() < 1e-14


def test_free_drift():
    n, a, da, fa1 = 16, 0.3, 0.01, 1.7
    x = rng(10).uniform(0, n, (3, 50))
    v = rng(11).normal(size=(3, 50))
    x2, v2 = O.integrate(x, v, a, fa1, da, np.zeros((n, n, n)), n)
    np.testing.assert_allclose(v2, v)
    np.testing.assert_allclose(x2, np.mod(x + da * v / (a + da) ** 2 * fa1, n), atol=1e-12)


# ----------------------------------------------------------------------------- cosmology
def test_f_is_inverse_aH():
    cos = [0.3, 0.7, 0.0]
    for a in (0.02, 0.5, 1.0):
        assert abs(O.f(a, cos) - 1 / (a * O.H(a, 1.0, cos))) < 1e-13


def test_advance_time_uses_h0_slot():
    # Q1: integrate.py:12 passes [H0, OL, OK] to f
    p = O.OracleParams(n_cells=16, n_parts=2, omega_m0=0.3, omega_lambda0=0.7, h0=0.55)
    pos = np.array([[3.2, 9.1], [4.4, 8.0], [1.0, 12.5]])
    v0 = np.zeros((3, 2))
    a, da = 0.4, 0.02
    rho = O.density(pos, 1.0, 16)
    _, v = O.advance_time(rho, pos, v0, O.fourier_grid(16), a, da, p)
    phi = O.potential(rho, O.fourier_grid(16), a, 0.3)
    g = _gather(pos, phi, 16)
    fa1 = 1 / math.sqrt((0.55 + 0.0 * (a + da) + 0.7 * (a + da) ** 3) / (a + da))
    np.testing.assert_allclose(v, da * fa1 * g, rtol=1e-12)


def test_step_count():
    # pmesh.py:56,63 replayed in float64 (Q7)
    assert O.step_count(0.02, 1.0, 100) == 99
    assert O.step_count(0.02, 1.0, 1000) == 1000
    assert O.step_count(0.02, 1.0, 200) == 199
    assert O.step_count(0.02, 1.0, 500) == 500


def test_reference_dtype_gap_small():
    n = 16
    pos = rng(12).uniform(0, n, (3, 2000)).astype(np.float32)
    r64 = O.density(pos, 1.0, n)
    r32 = O.density(pos, 1.0, n, ref_dtypes=True)
    assert np.linalg.norm(r32 - r64) / np.linalg.norm(r64) < 1e-6


def test_power_spectrum_of_plane_wave():
    n, box = 16, 100.0
    zz, yy, xx = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    rho = 1.0 + 0.2 * np.cos(2 * np.pi * 3 * xx / n)
    k, pk, cnt = O.power_spectrum(rho, box)
    # all the power sits in the |m| = 3 shell: 2 modes of amplitude 0.1*N^3
    total = (pk * cnt).sum() * n ** 6 / box ** 3
    assert abs(total - 2 * (0.1 * n ** 3) ** 2) < 1e-6 * total
    assert np.argmax(pk) == 2


# Philox4x32-10 known-answer vectors published with the Random123 library (Salmon et al.,
# SC'11; kat_vectors, philox4x32 with 10 rounds): (counter, key) -> output
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expect):
    out = O.philox4x32_10([[c] for c in ctr], key)[:, 0]
    assert [int(x) for x in out] == list(expect)


def test_white_noise_is_standard_normal_and_seeded():
    n = 1_000_003                      # not a multiple of 4: the last group is partial
    z = O.white_noise(n, 42)
    assert z.shape == (n,)
    assert abs(z.mean()) < 5 / math.sqrt(n)
    assert abs(z.var() - 1) < 5 * math.sqrt(2 / n)
    # tail fraction beyond 3 sigma: 0.27 %
    assert abs(np.mean(np.abs(z) > 3) - 0.0026998) < 3e-4
    np.testing.assert_array_equal(z[:1000], O.white_noise(1000, 42))
    assert np.abs(z[:1000] - O.white_noise(1000, 43)).max() > 1
    # element e is word e % 4 of counter e // 4: the first two come from one Box-Muller pair
    u = O.philox4x32_10([[0], [0], [0], [0]], (42, 0))[:, 0].astype(np.float64)
    r = math.sqrt(-2 * math.log((u[0] + 0.5) * 2.0 ** -32))
    t = 2 * math.pi * (u[1] + 0.5) * 2.0 ** -32
    assert z[0] == pytest.approx(r * math.cos(t), rel=1e-15)
    assert z[1] == pytest.approx(r * math.sin(t), rel=1e-15)
