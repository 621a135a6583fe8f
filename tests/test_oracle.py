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
    H                     test_f_is_inverse_aH (H(a) = 1/(a f) with h0 = 1), test_h_closed_form
    Dt                    test_dt_einstein_de_sitter_and_linear, test_dt_against_cpt92
    step_count            test_step_count
    run                   test_run_free_particle_closed_form, test_run_counts_and_limit
    lattice_q             test_lattice_q_explicit
    wrap_positions        test_deposit_periodic_edge_and_q4, test_wrap_positions_edges
    shaped_field          test_shaped_field_constant_amplitude, test_shaped_field_one_mode
    zeldovich_displacement  test_zeldovich_plane_wave_each_axis, test_zeldovich_divergence
    zeldovich_positions   test_zeldovich_positions_wrap_lattice
    power_spectrum        test_power_spectrum_of_plane_wave
    philox4x32_10 / white_noise  test_philox_known_answers, test_white_noise_*
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


# ----------------------------------------------------------------- growth factor and lattice
def test_dt_einstein_de_sitter_and_linear():
    # cosmology.py:29-36 in an Omega_m = 1 universe is the textbook D(a) = a, and linear in a
    # for any cosmology
    for a in (0.01, 0.3, 1.0):
        assert abs(O.Dt(a, [1.0, 0.0, 0.0]) - a) < 1e-15
    c = [0.3, 0.7, 0.0]
    assert abs(O.Dt(0.6, c) - 2 * O.Dt(0.3, c)) < 1e-14


def test_dt_against_cpt92():
    # Carroll, Press & Turner (1992, ARA&A 30, 499, eq. 29): g = 5/2 Om / (Om^(4/7) - OL +
    # (1 + Om/2)(1 + OL/70)).  The reference's expression divides by Omega_m where CPT92
    # multiplies (cosmology.py:36), i.e. Dt = g a / Om^2
    for om, ol in ((0.3, 0.7), (0.5, 0.5), (1.0, 0.0), (0.25, 0.0)):
        g = 2.5 * om / (om ** (4.0 / 7.0) - ol + (1 + om / 2) * (1 + ol / 70))
        assert abs(O.Dt(0.7, [om, ol, 0.0]) - g * 0.7 / om ** 2) < 1e-13


def test_h_closed_form():
    # cosmology.py:11-18 at a = 1 is h0 for a flat universe, and scales as a^-3/2 for matter
    assert abs(O.H(1.0, 0.7, [0.3, 0.7, 0.0]) - 0.7) < 1e-15
    assert abs(O.H(0.25, 2.0, [1.0, 0.0, 0.0]) - 2.0 * 8.0) < 1e-12


def test_lattice_q_explicit():
    # zeldovich.py:79-83 for N_PARTS = 2, N_CELLS = 8: axis = [0, 4] + 0.5, particle
    # p = (i*2 + j)*2 + k at (q_i, q_j, q_k)
    q = O.lattice_q(2, 8, np.float64)
    lo, hi = 0.5, 4.5
    np.testing.assert_array_equal(q[0], [lo, lo, lo, lo, hi, hi, hi, hi])
    np.testing.assert_array_equal(q[1], [lo, lo, hi, hi, lo, lo, hi, hi])
    np.testing.assert_array_equal(q[2], [lo, hi, lo, hi, lo, hi, lo, hi])
    q3 = O.lattice_q(3, 12, np.float64)    # spacing N_CELLS/N_PARTS = 4, not a power of two
    np.testing.assert_array_equal(q3[2, :3], [0.5, 4.5, 8.5])
    np.testing.assert_array_equal(q3[0, ::9], [0.5, 4.5, 8.5])
    assert O.lattice_q(4, 8).dtype == np.float32


def test_wrap_positions_edges():
    n = 8
    x = np.array([-1e-9, 0.0, 7.999, 8.0, 11.5, -3.25], np.float64)
    np.testing.assert_array_equal(O.wrap_positions(x, n), [8 - 1e-9, 0.0, 7.999, 0.0, 3.5, 4.75])
    tiny = np.array([-1e-9], np.float32)          # rounds to n after the add: mapped to 0
    assert O.wrap_positions(tiny, n)[0] == 0.0


# --------------------------------------------------------------------------------- the loop
def test_run_free_particle_closed_form():
    # one particle feels no force (test_zero_self_force), so run() drifts it by
    # v * sum_k da f(a_k + da) / (a_k + da)^2 with f of [H0, OL, OK] (integrate.py:12,95) over
    # the pmesh.py:56-63 steps
    n = 8
    p = O.OracleParams(n_cells=n, n_parts=1, omega_m0=0.3, omega_lambda0=0.7, h0=0.7,
                       a_init=0.1, a_end=1.0, steps=12)
    x0 = np.array([[2.3], [5.1], [7.6]])
    v0 = np.array([[0.4], [-0.25], [0.1]])
    x, v, a, k = O.run(x0, v0, p)
    da = 0.9 / 12
    kk = O.step_count(0.1, 1.0, 12)
    ak = 0.1 + da * np.arange(kk) + da                         # a + da of each step
    fa = 1 / np.sqrt((0.7 + 0.7 * ak ** 3) / ak)
    drift = np.sum(da * fa / ak ** 2)
    assert k == kk and abs(a - (0.1 + kk * da)) < 1e-12
    np.testing.assert_allclose(v, v0, atol=1e-12)
    np.testing.assert_allclose(x, np.mod(x0 + v0 * drift, n), atol=1e-10)


def test_run_counts_and_limit():
    n = 8
    p = O.OracleParams(n_cells=n, n_parts=2, omega_m0=0.3, omega_lambda0=0.7, h0=0.7,
                       a_init=0.02, a_end=1.0, steps=100)
    x0 = O.lattice_q(2, n, np.float64) + 0.3
    v0 = np.zeros_like(x0)
    _, _, a, k = O.run(x0, v0, p, n_steps=5)
    assert k == 5 and abs(a - (0.02 + 5 * 0.0098)) < 1e-12
    *_, k_all = O.run(x0, v0, p)
    assert k_all == 99                     # Q7: STEPS = 100 from 0.02 runs 99 steps


# ----------------------------------------------------------------- initial conditions
def test_shaped_field_constant_amplitude():
    # amp = c for every |m|^2 >= 1 and amp[0] = 0: the field is c times the noise minus its mean
    n = 8
    w = rng(20).normal(size=(n, n, n))
    amp = np.full(3 * (n // 2) ** 2 + 1, 1.7)
    amp[0] = 0.0
    np.testing.assert_allclose(O.shaped_field(w, amp), 1.7 * (w - w.mean()), atol=1e-12)


def test_shaped_field_one_mode():
    # a plane wave of mode m = (1, 2, 0) has |m|^2 = 5: only amp[5] scales it
    n = 8
    i = np.arange(n)
    w = np.cos(2 * np.pi * (1 * i[:, None, None] + 2 * i[None, :, None]) / n) * np.ones((1, 1, n))
    amp = np.zeros(3 * (n // 2) ** 2 + 1)
    amp[5] = 3.0
    np.testing.assert_allclose(O.shaped_field(w.reshape(-1), amp), 3.0 * w, atol=1e-12)
    amp[5], amp[4] = 0.0, 1.0
    assert np.abs(O.shaped_field(w, amp)).max() < 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_zeldovich_plane_wave_each_axis(axis):
    # delta = A cos(k q) along lattice axis `axis` (k = 2 pi m / BOX): the Zel'dovich
    # displacement is Psi = -(A/k) sin(k q) along that axis only (div Psi = -delta), in cells
    # N_CELLS/BOX per Mpc/h, and moves positions[axis] (Q8)
    n, n_cells, box, m, A = 8, 16, 40.0, 2, 0.3
    i = np.arange(n)
    q = i * box / n
    k = 2 * np.pi * m / box
    shape = [1, 1, 1]
    shape[axis] = n
    delta = np.broadcast_to((A * np.cos(k * q)).reshape(shape), (n, n, n))
    psi = O.zeldovich_displacement(delta, n_cells, box)
    expect = np.broadcast_to((-(A / k) * np.sin(k * q)).reshape(shape), (n, n, n)).reshape(-1)
    for a in range(3):
        np.testing.assert_allclose(psi[a], expect * (n_cells / box) if a == axis else 0.0,
                                   atol=1e-12)


def test_zeldovich_divergence():
    # a sum of two waves, one oblique: delta = A cos(k.q + p) gives Psi = -A k/|k|^2 sin(k.q + p)
    # and B sin(k2 x) gives Psi_x = (B/k2) cos(k2 x); both satisfy div Psi = -delta
    n, n_cells, box = 8, 8, 8.0        # one cell per Mpc/h, so Psi is in the field's units
    i = np.arange(n)
    z, y, x = np.meshgrid(i, i, i, indexing="ij")
    delta = 0.2 * np.cos(2 * np.pi * (z + 2 * y - x) / n + 0.3) + 0.1 * np.sin(2 * np.pi * 3 * x / n)
    psi = O.zeldovich_displacement(delta, n_cells, box).reshape(3, n, n, n)
    kz, ky, kx = (2 * np.pi / n) * np.array([1, 2, -1])
    kk = kz ** 2 + ky ** 2 + kx ** 2
    ph = 2 * np.pi * (z + 2 * y - x) / n + 0.3
    expect_z = -0.2 * kz / kk * np.sin(ph)
    k2 = 2 * np.pi * 3 / n
    expect_x = -0.2 * kx / kk * np.sin(ph) + 0.1 / k2 * np.cos(k2 * x)
    np.testing.assert_allclose(psi[0], expect_z, atol=1e-12)
    np.testing.assert_allclose(psi[2], expect_x, atol=1e-12)


def test_zeldovich_positions_wrap_lattice():
    n_parts, n_cells = 2, 8
    psi = np.zeros((3, 8))
    psi[0, 0] = -1.0          # 0.5 - 1 wraps to 7.5
    psi[2, 7] = 3.5           # 4.5 + 3.5 = 8 wraps to 0
    x = O.zeldovich_positions(psi, n_cells)
    q = O.lattice_q(n_parts, n_cells, np.float64)
    assert x[0, 0] == 7.5 and x[2, 7] == 0.0
    np.testing.assert_array_equal(x[1], q[1])
