"""Batched ensemble: every member advanced in the batch must equal the same sim run alone.

The batched kernels run the single-sim device bodies on each member's own buffers, so a member
differs from its solo run only through the order of the floating-point sums in the deposit
(LDS accumulation order is not fixed), which rounds identically almost everywhere.
Tolerance: positions within 1e-4 cells, velocities and density within 1e-5 (norm-wise).
"""
import numpy as np
import pytest

from helpers import make_ics, periodic_diff, rel

pytestmark = pytest.mark.gpu

# (N_CELLS, N_PARTS, a0, da, flags): both FFT size classes, an odd lattice, a flag
MEMBERS = [
    (32, 16, 0.02, 0.0098, {}),
    (48, 24, 0.03, 0.0100, {"cic_deconvolve": True}),
    (64, 32, 0.05, 0.0050, {}),
    (96, 36, 0.02, 0.0120, {}),
    (384, 16, 0.04, 0.0080, {}),
]


def _sim(n, npart, **kw):
    from pmgpu.sim import PMSim
    return PMSim(n, npart, 32, omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7, **kw)


def _pair(i, seed0=40):
    n, npart, a0, da, flags = MEMBERS[i]
    pos, vel = make_ics(n, npart, seed=seed0 + i, box=float(n))
    alone, member = _sim(n, npart, **flags), _sim(n, npart, **flags)
    alone.set_state(pos, vel)
    member.set_state(pos, vel)
    return alone, member


def _assert_same(alone, member, what=""):
    n = alone.n_cells
    xa, va = alone.get_state()
    xm, vm = member.get_state()
    dx = np.abs(periodic_diff(xm, xa, n)).max()
    ev = rel(vm, va)
    er = rel(member.density(np.float64), alone.density(np.float64))
    print(f"{what} N={n}: max|dx| {dx:.2e} vel {ev:.2e} rho {er:.2e}")
    assert dx < 1e-4 and ev < 1e-5 and er < 1e-5, what


def test_ensemble_matches_members_run_alone():
    from pmgpu.ensemble import PMEnsemble
    pairs = [_pair(i) for i in range(len(MEMBERS))]
    ens = PMEnsemble([m for _, m in pairs])
    assert len(ens) == len(MEMBERS)
    a0 = [m[2] for m in MEMBERS]
    da = [m[3] for m in MEMBERS]
    # 11 steps: enqueued launch by launch (too short to capture a graph); then 70 steps: a
    # captured graph of 8 steps replayed 8 times plus 6 table-driven steps
    for steps in (11, 70):
        a_out = ens.run(a0, da, steps)
        for i, (alone, member) in enumerate(pairs):
            a_alone = alone.run(a0[i], da[i], steps)
            assert a_out[i] == a_alone
            _assert_same(alone, member, f"{steps} steps, member {i}")
        a0 = a_out
    # the potential of the last step is kept per member as well
    for alone, member in pairs:
        assert rel(member.potential(np.float64), alone.potential(np.float64)) < 1e-5
    ens.close()
    for p in pairs:
        for s in p:
            s.close()


def test_ensemble_membership_changes_and_solo_steps():
    from pmgpu.ensemble import PMEnsemble
    pairs = {i: _pair(i, seed0=60) for i in (0, 2, 3)}
    a = {0: 0.02, 2: 0.03, 3: 0.04}
    da = {0: 0.01, 2: 0.01, 3: 0.005}
    ens = PMEnsemble()
    order = []
    for i in (0, 2, 3):
        ens.add(pairs[i][1])
        order.append(i)

    def run_all(steps, what):
        a_new = ens.run([a[i] for i in order], [da[i] for i in order], steps)
        for i, an in zip(order, a_new):
            assert pairs[i][0].run(a[i], da[i], steps) == an
            a[i] = an
        for i in order:                      # eager work between replays: read back, compare
            _assert_same(*pairs[i], f"{what} member {i}")

    run_all(9, "all")
    ens.remove(pairs[2][1])                  # leaves and steps alone meanwhile
    order.remove(2)
    assert len(ens) == 2
    a_solo = pairs[2][1].run(a[2], da[2], 5)
    assert pairs[2][0].run(a[2], da[2], 5) == a_solo
    a[2] = a_solo
    run_all(9, "two")
    ens.add(pairs[2][1])                     # rejoins at its own scale factor
    order.append(2)
    run_all(9, "rejoined")
    ens.close()
    for p in pairs.values():
        for s in p:
            s.close()


def test_ensemble_graph_replays_equal_eager_steps():
    # the captured graph is replayed in three calls, with eager work and synchronisation
    # between them; every call must agree with the same steps enqueued launch by launch
    from pmgpu.ensemble import PMEnsemble
    p1 = [_pair(i, seed0=80) for i in (0, 2)]
    p2 = [_pair(i, seed0=80) for i in (0, 2)]
    e1 = PMEnsemble([m for _, m in p1])
    e2 = PMEnsemble([m for _, m in p2])
    e2.set_timing(True)                                 # timing: launch by launch, no graph
    a, da = [0.02, 0.05], [0.001, 0.0005]
    for rep, steps in enumerate((70, 16, 9)):
        a1 = e1.run(a, da, steps)
        a2 = e2.run(a, da, steps)
        assert a1 == a2
        for (_, m1), (_, m2) in zip(p1, p2):
            _assert_same(m1, m2, f"graph vs eager, call {rep}")
        a = a1
    t = e2.timing()
    assert t["z_green"][1] == 95 and t["gather_kick_drift"][0] > 0
    e1.close()
    e2.close()


def test_deterministic_ensemble_members_equal_solo_runs_bitwise():
    # With the order-independent deposit nothing in a step depends on the order of work: a
    # member advanced in the batch (16^3 tiles; the tile scatter for members with 4 or more
    # tiles per axis, the particle-parallel scatter for the others) then gives the bits of its
    # solo run in its own tile shape, over a captured graph's replays and launch-by-launch steps.
    from pmgpu.ensemble import PMEnsemble
    pairs = []
    for i in (0, 1, 2, 3):                  # 32^3 / 48^3: particle-parallel, 64^3 / 96^3: tile scatter
        n, npart, a0, da, flags = MEMBERS[i]
        pos, vel = make_ics(n, npart, seed=90 + i, box=float(n))
        pr = (_sim(n, npart, deterministic=True, **flags), _sim(n, npart, deterministic=True, **flags))
        for s in pr:
            s.set_state(pos, vel)
        pairs.append(pr)
    ens = PMEnsemble([m for _, m in pairs])
    a0 = [MEMBERS[i][2] for i in (0, 1, 2, 3)]
    da = [MEMBERS[i][3] for i in (0, 1, 2, 3)]
    for steps in (11, 20):
        a_out = ens.run(a0, da, steps)
        for k, (alone, member) in enumerate(pairs):
            assert alone.run(a0[k], da[k], steps) == a_out[k]
            for name, x, y in zip(("x", "v"), alone.get_state(np.float64), member.get_state(np.float64)):
                assert np.array_equal(x, y), f"{steps} steps, member {k}: {name}"
            assert np.array_equal(alone.density(np.float64), member.density(np.float64)), f"member {k}: rho"
        a0 = a_out
    ens.close()
    for p in pairs:
        for s in p:
            s.close()


def test_ensemble_rejects_unsupported_members():
    from pmgpu.ensemble import PMEnsemble
    from pmgpu.sim import PMSim
    ens = PMEnsemble()
    s64 = PMSim(32, 16, 64)
    with pytest.raises(RuntimeError, match="FP32"):
        ens.add(s64)
    s16 = PMSim(16, 8, 32)
    with pytest.raises(RuntimeError, match="batched grid size"):
        ens.add(s16)
    s = _sim(32, 16)
    ens.add(s)
    with pytest.raises(RuntimeError, match="already"):
        ens.add(s)
    other = PMEnsemble()
    with pytest.raises(RuntimeError, match="another ensemble"):
        other.add(s)
    with pytest.raises(RuntimeError, match="no state"):
        ens.run([0.1], [0.01], 2)
    s.close()                                           # leaves the ensemble on destruction
    assert len(ens) == 0
    for x in (s64, s16):
        x.close()
    ens.close()
    other.close()
