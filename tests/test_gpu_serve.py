"""Serving loop: continuous batching must not change any result.

Every request served in the mix (admitted while others run, advanced in shared batched
launches, retired at a chunk boundary) must equal the same request run alone: final
positions within 1e-4 cells, velocities and P(k) within 1e-5 (the two runs differ only in
the order of floating-point sums inside the deposit), same step count and scale factor.
"""
import numpy as np
import pytest

from helpers import periodic_diff, rel

pytestmark = pytest.mark.gpu


def _requests():
    from pmgpu.serve import SimRequest
    spec = [(64, 20, 60.0), (96, 35, 150.0), (128, 9, 80.0), (64, 50, 300.0), (96, 17, 90.0),
            (128, 28, 200.0), (64, 3, 50.0)]
    return [SimRequest(id=i, seed=500 + i, box=box, n_cells=n, n_parts=n // 2, steps=steps,
                       arrival=0.01 * i) for i, (n, steps, box) in enumerate(spec)]


def test_served_results_equal_runs_alone():
    from pmgpu.serve import PMServer, run_alone
    reqs = _requests()
    server = PMServer(max_members=3, chunk_steps=8, keep_state=True)
    served = server.serve(reqs, realtime=False)
    assert [r.id for r in served] == [q.id for q in reqs]
    assert server.stats["peak_members"] == 3
    for q, r in zip(reqs, served):
        alone = run_alone(q, keep_state=True)
        assert r.steps == alone.steps == q.loop_steps
        assert r.a_final == alone.a_final
        dx = np.abs(periodic_diff(r.positions, alone.positions, q.n_cells)).max()
        ev = rel(r.velocities, alone.velocities)
        ep = np.max(np.abs(r.pk / alone.pk - 1))
        print(f"request {q.id} N={q.n_cells} steps={r.steps}: dx {dx:.1e} vel {ev:.1e} P(k) {ep:.1e}")
        assert dx < 1e-4 and ev < 1e-5 and ep < 1e-5


def test_admission_respects_hbm_budget():
    from pmgpu.serve import PMServer, SimRequest
    reqs = [SimRequest(id=i, seed=700 + i, box=100.0, n_cells=128, n_parts=64, steps=12)
            for i in range(5)]
    budget = 2 * reqs[0].footprint() + reqs[0].footprint() // 2     # room for two at a time
    server = PMServer(hbm_budget=budget, chunk_steps=4)
    res = server.serve(reqs, realtime=False)
    assert len(res) == 5 and server.stats["peak_members"] == 2
    assert all(r.steps == reqs[0].loop_steps for r in res)
    too_big = SimRequest(id=9, seed=1, box=100.0, n_cells=256, n_parts=128, steps=10)
    with pytest.raises(ValueError, match="HBM budget"):
        server.serve([too_big])


def test_realtime_arrivals_are_admitted_after_they_arrive():
    from pmgpu.serve import PMServer, SimRequest, latency_summary
    reqs = [SimRequest(id=i, seed=800 + i, box=80.0, n_cells=64, n_parts=32, steps=40,
                       arrival=0.05 * i) for i in range(6)]
    res = PMServer(chunk_steps=8).serve(reqs, realtime=True)
    for q, r in zip(reqs, res):
        assert r.arrival == q.arrival
        assert r.admitted >= q.arrival - 1e-6
        assert r.finished >= r.admitted
    lat = latency_summary(res)
    assert 0 < lat["p50_s"] <= lat["p99_s"] <= lat["max_s"]


def test_cancelling_an_active_request_frees_it_and_the_rest_are_unchanged():
    # request 1 (300 loop steps) is cancelled when request 0 (10 steps) completes: it ends at
    # the next chunk boundary, partly stepped, its memory released; request 2 still equals its
    # run alone
    from pmgpu.serve import PMServer, SimRequest, run_alone
    reqs = [SimRequest(id=0, seed=900, box=60.0, n_cells=64, n_parts=32, steps=11),
            SimRequest(id=1, seed=901, box=90.0, n_cells=96, n_parts=48, steps=301),
            SimRequest(id=2, seed=902, box=80.0, n_cells=64, n_parts=32, steps=41)]
    server = PMServer(chunk_steps=8, keep_state=True)
    res = server.serve(reqs, realtime=False,
                       on_result=lambda r: server.cancel(1) if r.id == 0 else None)
    assert [r.error for r in res] == [None, "cancelled", None]
    assert 0 < res[1].steps < reqs[1].loop_steps
    assert server.stats["cancelled"] == 1
    alone = run_alone(reqs[2], keep_state=True)
    assert res[2].steps == alone.steps and res[2].a_final == alone.a_final
    assert np.abs(periodic_diff(res[2].positions, alone.positions, 64)).max() < 1e-4
    assert rel(res[2].velocities, alone.velocities) < 1e-5
