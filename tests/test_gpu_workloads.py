"""The batched-ensemble and serving workloads at their own configurations.

* Ensemble: the 48 realisations of the bench's ensemble (seeds 100-147, N_CELLS drawn from
  128 / 192 / 256 with (N_CELLS/2)^3 particles, box U(50, 300), sigma8 U(0.7, 0.9), device
  initial conditions) advanced together for 20 steps (a captured graph of 8 steps replayed
  twice plus 4 table-driven steps).  Every member equals the same realisation run alone on the
  ensemble's kernels (positions within 1e-4 cells, velocities and density within 1e-5: only the
  order of the deposit's LDS sums differs); one member of each grid size then takes a step
  that is checked against the float64 oracle (potential and accelerations within 1e-5).
* Serving: the first two requests of each grid size (64, 128, 192, 256, 384) of the bench's
  serving stream (Poisson 2 req/s, seed 4), served together with continuous batching and
  compared with each request run alone on the same kernels: same step count and final scale
  factor, positions within 1e-4 cells, velocities and P(k) within 1e-5.
"""
import numpy as np
import pytest

import pm_oracle as O
from helpers import periodic_diff, rel

pytestmark = pytest.mark.gpu


def test_ensemble_of_48_matches_members_alone_and_oracle():
    import bench
    import initial_conditions as ic
    from pmgpu.ensemble import PMEnsemble
    from pmgpu.sim import PMSim
    e = bench.ENSEMBLE
    members = bench.ensemble_members(e)
    assert len(members) == 48
    assert [m["seed"] for m in members] == list(range(100, 148))
    assert {m["n_cells"] for m in members} == {128, 192, 256}
    da = (e["a_end"] - e["a_init"]) / e["steps"]
    kw = lambda m: dict(omega_m0=m["om"], omega_lambda0=m["ol"], omega_k0=m["ok"], h0=m["h"])
    sims, alone = [], []
    for m in members:
        s = PMSim(m["n_cells"], m["n_parts"], 32, **kw(m))
        ic.zeldovich_device(s, bench._ic_params(m))
        pos, vel = s.get_state()
        s.set_state(pos, vel)                 # both runs start from the same float32 state
        a = PMSim(m["n_cells"], m["n_parts"], 32, **kw(m))
        a.set_tiles(16, 16, 16)               # the ensemble's tile kernels
        a.set_state(pos, vel)
        sims.append(s)
        alone.append(a)
    ens = PMEnsemble(sims)
    steps = 20
    a_out = ens.run([e["a_init"]] * 48, [da] * 48, steps)
    worst = [0.0, 0.0, 0.0]
    for i, (s, a) in enumerate(zip(sims, alone)):
        a_alone = a.run(e["a_init"], da, steps)
        assert a_out[i] == a_alone
        (xm, vm), (xa, va) = s.get_state(), a.get_state()
        dx = float(np.abs(periodic_diff(xm, xa, s.n_cells)).max())
        ev = rel(vm, va)
        er = rel(s.density(np.float64), a.density(np.float64))
        worst = [max(worst[0], dx), max(worst[1], ev), max(worst[2], er)]
        assert dx < 1e-4 and ev < 1e-5 and er < 1e-5, (i, dx, ev, er)
    print(f"48 members vs alone after {steps} steps: max|dx| {worst[0]:.1e} vel {worst[1]:.1e} "
          f"rho {worst[2]:.1e}")
    ens.close()
    for a in alone:
        a.close()
    # one member per grid size: a further step from its ensemble state against the oracle
    a_now = e["a_init"] + steps * da
    for n in (128, 192, 256):
        i = next(k for k, m in enumerate(members) if m["n_cells"] == n)
        s, m = sims[i], members[i]
        x0, v0 = s.get_state(np.float64)
        p = O.OracleParams(n_cells=n, n_parts=m["n_parts"], omega_m0=m["om"],
                           omega_lambda0=m["ol"], h0=m["h"])
        rho = O.density(x0, p.mass, n)
        _, v1r, phi_r = O.advance_time(rho, x0, v0, O.fourier_grid(n), a_now, da, p,
                                       return_phi=True)
        s.step(a_now, da)
        e_phi = rel(s.potential(np.float64), phi_r)
        e_acc = rel(s.get_state(np.float64)[1] - v0, v1r - v0)
        print(f"member {i} (N={n}, box {m['box']:.1f}) step vs oracle: phi {e_phi:.2e} accel {e_acc:.2e}")
        assert e_phi < 1e-5 and e_acc < 1e-5
    for s in sims:
        s.close()


def test_serving_stream_subset_equals_runs_alone():
    import bench
    from pmgpu.serve import PMServer, SimRequest, poisson_stream, run_alone
    sv = bench.SERVE
    stream = poisson_stream(sv["rate"], sv["duration"], seed=sv["seed"], sizes=sv["sizes"],
                            weights=sv["weights"], box=sv["box"], steps=sv["steps"])
    picked = []
    for n in sv["sizes"]:
        picked += [r for r in stream if r.n_cells == n][:2]
    assert sorted({r.n_cells for r in picked}) == [64, 128, 192, 256, 384] and len(picked) == 10
    server = PMServer(chunk_steps=16, keep_state=True)
    served = server.serve(picked, realtime=False)
    assert [r.id for r in served] == sorted(r.id for r in picked)
    by_id = {r.id: r for r in picked}
    for r in served:
        q = by_id[r.id]
        alone = run_alone(q, keep_state=True, tiles=(16, 16, 16))
        assert r.steps == alone.steps == q.loop_steps
        assert r.a_final == alone.a_final
        dx = float(np.abs(periodic_diff(r.positions, alone.positions, q.n_cells)).max())
        ev = rel(r.velocities, alone.velocities)
        ep = float(np.max(np.abs(r.pk / alone.pk - 1)))
        print(f"request {q.id} N={q.n_cells} steps={r.steps}: dx {dx:.1e} vel {ev:.1e} P(k) {ep:.1e}")
        assert dx < 1e-4 and ev < 1e-5 and ep < 1e-5
    print(f"served {len(served)} requests, peak {server.stats['peak_members']} in flight")
