"""Host-side logic that runs without a GPU: the serving scheduler (admission by footprint,
continuous batching, chunking), the request stream, the step-count replay and snapshots.

The scheduler is driven with stand-ins for the device objects (an ensemble that records its
membership at every batched call, simulations that only count steps): what is checked is the
order and the accounting of the host loop, not any numerics.
"""
import numpy as np
import pytest

from pmgpu import serve
from pmgpu.serve import PMServer, SimRequest, latency_summary, poisson_stream


class FakeSim:
    def __init__(self, req):
        self.req = req
        self.steps = 0
        self.closed = False

    def power_spectrum(self, box):
        return np.arange(3.0), np.full(3, float(self.req.id)), None

    def get_state(self):
        return np.zeros((3, 1), np.float32), np.zeros((3, 1), np.float32)

    def close(self):
        self.closed = True


class FakeEnsemble:
    """Records the members and their footprints at every batched call."""
    log = []

    def __init__(self):
        self.members = []

    def add(self, sim):
        assert sim not in self.members
        self.members.append(sim)

    def remove(self, sim):
        self.members.remove(sim)

    def run(self, a, da, n_steps):
        assert len(a) == len(da) == len(self.members)
        FakeEnsemble.log.append(([m.req.id for m in self.members], n_steps))
        out = []
        for m, a0, d in zip(self.members, a, da):
            m.steps += n_steps
            for _ in range(n_steps):
                a0 += d
            out.append(a0)
        return out

    def close(self):
        pass


FOOT = {64: 1, 128: 4, 192: 9, 256: 16, 384: 40}


@pytest.fixture
def fake_device(monkeypatch):
    FakeEnsemble.log = []
    started = []

    def start(req, prepared):
        sim = FakeSim(req)
        started.append(sim)
        return sim

    monkeypatch.setattr(serve, "PMEnsemble", FakeEnsemble)
    monkeypatch.setattr(serve, "_start", start)
    monkeypatch.setattr(serve, "_prepare", lambda req: None)
    monkeypatch.setattr(SimRequest, "footprint", lambda self: FOOT[self.n_cells])
    return started


def _req(i, n, steps, arrival=0.0):
    return SimRequest(id=i, seed=i, box=100.0, n_cells=n, n_parts=n // 2, steps=steps, arrival=arrival)


def _replay(req):
    a = req.a_init
    for _ in range(req.loop_steps):
        a += req.da
    return a


def test_server_runs_every_request_to_its_loop_count(fake_device):
    reqs = [_req(i, n, s) for i, (n, s) in enumerate([(64, 60), (128, 100), (256, 37), (64, 500), (192, 51)])]
    out = PMServer(hbm_budget=100, max_members=8, chunk_steps=16).serve(reqs, realtime=False)
    assert [r.id for r in out] == [r.id for r in reqs]
    for r, q in zip(out, reqs):
        assert r.steps == q.loop_steps
        assert r.a_final == _replay(q)  # the same float64 sequence as the reference's loop
        assert r.pk[0] == q.id
    assert all(s.closed for s in fake_device)
    # every simulation stepped exactly its loop count, in chunks of at most 16 steps
    assert sorted(s.steps for s in fake_device) == sorted(q.loop_steps for q in reqs)
    assert all(0 < n <= 16 for _, n in FakeEnsemble.log)


def test_admission_respects_budget_and_slots(fake_device):
    reqs = [_req(i, n, 120) for i, n in enumerate([256, 256, 192, 64, 128, 64, 384, 64])]
    budget, slots = 45, 3
    PMServer(hbm_budget=budget, max_members=slots, chunk_steps=8).serve(reqs, realtime=False)
    by_id = {q.id: q for q in reqs}
    for ids, _ in FakeEnsemble.log:
        assert len(ids) <= slots
        assert sum(FOOT[by_id[i].n_cells] for i in ids) <= budget


def test_a_request_that_does_not_fit_waits_and_smaller_ones_pass(fake_device):
    # 0 (256: 16) runs; 1 (384: 40) does not fit beside it in 50; 2 (64: 1) passes it
    reqs = [_req(0, 256, 300), _req(1, 384, 100), _req(2, 64, 100)]
    PMServer(hbm_budget=50, max_members=8, chunk_steps=32).serve(reqs, realtime=False)
    first = FakeEnsemble.log[0][0]
    assert first == [0, 2]
    # 1 is admitted only once 0 has retired
    with_1 = [ids for ids, _ in FakeEnsemble.log if 1 in ids]
    assert with_1 and all(0 not in ids for ids in with_1)


def test_chunks_end_at_the_first_completion(fake_device):
    reqs = [_req(0, 64, 100), _req(1, 64, 30)]  # 99 and 30 loop steps
    PMServer(hbm_budget=10, max_members=8, chunk_steps=64).serve(reqs, realtime=False)
    assert FakeEnsemble.log[0] == ([0, 1], 30)
    assert FakeEnsemble.log[1] == ([0], 64)
    assert FakeEnsemble.log[2] == ([0], 5)


def test_a_failed_start_is_reported_and_the_others_are_served(fake_device):
    # a start that raises (an allocation failure, say) ends that request with an error; its
    # budget is released and every other request is served
    def start(req, prepared):
        if req.id == 1:
            raise RuntimeError("libpm pm_sim_create failed (code 3): device allocation failed")
        sim = FakeSim(req)
        fake_device.append(sim)
        return sim

    reqs = [_req(i, 64, 40) for i in range(4)]
    server = PMServer(hbm_budget=2, max_members=8, chunk_steps=16, start=start)
    out = server.serve(reqs, realtime=False)
    assert [r.id for r in out] == [0, 1, 2, 3]
    assert out[1].error and "allocation failed" in out[1].error and out[1].steps == 0
    assert all(r.error is None and r.steps == reqs[0].loop_steps for r in (out[0], out[2], out[3]))
    assert server.stats["failed"] == 1
    assert all(s.closed for s in fake_device)


class SickEnsemble(FakeEnsemble):
    """Member 1 of the first batched call goes non-finite after 3 of its steps."""
    fired = False

    def run(self, a, da, n_steps):
        if not SickEnsemble.fired and len(self.members) > 1:
            SickEnsemble.fired = True
            out = super().run(a, da, 3)
            exc = RuntimeError("libpm pm_ensemble_run failed (code 6): ensemble member 1 "
                               "(after 3 steps): particle state is corrupt")
            exc.a_after = out
            raise exc
        return super().run(a, da, n_steps)


def test_a_member_that_turns_non_finite_is_retired_with_an_error(fake_device):
    SickEnsemble.fired = False
    reqs = [_req(i, 64, 40) for i in range(3)]
    server = PMServer(hbm_budget=10, max_members=8, chunk_steps=16, ensemble=SickEnsemble)
    out = server.serve(reqs, realtime=False)
    assert out[1].error and "corrupt" in out[1].error
    for r in (out[0], out[2]):
        assert r.error is None and r.steps == reqs[0].loop_steps
        assert r.a_final == _replay(reqs[0])       # the 3 steps before the failure count
    assert server.stats["failed"] == 1
    assert all(s.closed for s in fake_device)


class CancellingEnsemble(FakeEnsemble):
    """Asks the server to cancel request 1 during the first batched call."""
    server = None

    def run(self, a, da, n_steps):
        if len(FakeEnsemble.log) == 0:
            CancellingEnsemble.server.cancel(1)
        return super().run(a, da, n_steps)


def test_cancelled_requests_release_their_slot_and_report_their_steps(fake_device):
    # 0 and 1 run together (budget 32 = two 256^3 requests); 2 waits for a slot; 3 is
    # cancelled before it is ever admitted.  Request 1 is cancelled during the first chunk:
    # it ends at that chunk boundary with the steps it took, its sim is closed, and 2 is
    # admitted into its budget without waiting for 0 to finish.
    reqs = [_req(0, 256, 200), _req(1, 256, 200), _req(2, 256, 60), _req(3, 64, 40)]
    server = PMServer(hbm_budget=32, max_members=8, chunk_steps=16, ensemble=CancellingEnsemble)
    CancellingEnsemble.server = server
    server.cancel(3)
    out = server.serve(reqs, realtime=False)
    assert [r.id for r in out] == [0, 1, 2, 3]
    assert out[1].error == "cancelled" and out[1].steps == 16
    assert out[3].error == "cancelled" and out[3].steps == 0
    assert out[0].error is None and out[0].steps == reqs[0].loop_steps
    assert out[2].error is None and out[2].steps == reqs[2].loop_steps
    assert FakeEnsemble.log[0][0] == [0, 1]
    # 2 was admitted into the released budget while 0 was still running (its start then
    # runs in the starter thread and it joins at a later chunk boundary)
    assert out[2].admitted < out[0].finished
    assert all(3 not in ids for ids, _ in FakeEnsemble.log)
    assert server.stats["cancelled"] == 2 and server.stats["failed"] == 0
    assert [s.req.id for s in fake_device] == [0, 1, 2]   # 3 was never started
    assert all(s.closed for s in fake_device)


def test_cancelling_a_finished_or_unknown_request_changes_nothing(fake_device):
    reqs = [_req(0, 64, 40), _req(1, 64, 40)]
    server = PMServer(hbm_budget=10, max_members=8, chunk_steps=16)
    server.cancel(7)
    out = server.serve(reqs, realtime=False, on_result=lambda r: server.cancel(r.id))
    assert all(r.error is None and r.steps == reqs[0].loop_steps for r in out)
    assert server.stats["cancelled"] == 0


def test_request_larger_than_the_budget_is_rejected(fake_device):
    with pytest.raises(ValueError, match="HBM budget"):
        PMServer(hbm_budget=10).serve([_req(0, 384, 50)], realtime=False)


def test_poisson_stream_matches_its_specification():
    reqs = poisson_stream(2.0, 600.0, seed=4)
    again = poisson_stream(2.0, 600.0, seed=4)
    assert [(r.arrival, r.n_cells, r.box, r.steps) for r in reqs] == \
           [(r.arrival, r.n_cells, r.box, r.steps) for r in again]
    assert abs(len(reqs) - 1200) < 5 * np.sqrt(1200)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0) and t[-1] <= 600.0
    assert [r.id for r in reqs] == list(range(len(reqs)))
    assert [r.seed for r in reqs] == [1000 + i for i in range(len(reqs))]
    n = np.array([r.n_cells for r in reqs])
    assert set(n) <= {64, 128, 192, 256, 384}
    assert all(r.n_parts == r.n_cells // 2 for r in reqs)
    assert all(50.0 <= r.box <= 400.0 and 50 <= r.steps <= 500 for r in reqs)
    frac = np.array([np.mean(n == s) for s in (64, 128, 192, 256, 384)])
    assert np.all(np.abs(frac - [0.3, 0.3, 0.2, 0.15, 0.05]) < 0.05)
    # compressed arrivals: the same requests, arriving faster
    fast = poisson_stream(2.0, 600.0, seed=4, time_scale=0.1)
    assert [r.n_cells for r in fast] == list(n)
    assert np.allclose([r.arrival for r in fast], t * 0.1)


def test_latency_summary():
    res = [serve.SimResult(i, arrival=0.0, admitted=0.0, finished=float(i + 1), steps=1, a_final=1.0,
                           k=np.zeros(1), pk=np.zeros(1)) for i in range(100)]
    s = latency_summary(res)
    assert s["p50_s"] == pytest.approx(50.5) and s["max_s"] == 100.0
    assert s["p99_s"] == pytest.approx(np.percentile(np.arange(1, 101), 99))
    assert s["mean_s"] == pytest.approx(50.5)


@pytest.mark.parametrize("steps, executed", [(100, 99), (1000, 1000), (200, 199), (500, 500)])
def test_request_loop_count_replays_the_reference_loop(steps, executed):
    # pmesh.py:56-63 runs `while a < A_END - da` in float64: 99 / 1000 / 199 / 500 steps for
    # STEPS = 100 / 1000 / 200 / 500 from a = 0.02 to 1
    assert _req(0, 64, steps).loop_steps == executed


def test_snapshot_roundtrip_npz(tmp_path, monkeypatch):
    import configure_me
    import save_data
    monkeypatch.setattr(save_data, "h5py", None)
    monkeypatch.setattr(configure_me, "DATA_DIR", str(tmp_path), raising=False)
    monkeypatch.setattr(configure_me, "SAVE_DENSITY", True)
    monkeypatch.setattr(configure_me, "N_CELLS", 32)
    monkeypatch.setattr(configure_me, "N_PARTS", 16)
    monkeypatch.setattr(configure_me, "BOX_SIZE", 100.0)
    monkeypatch.setattr(configure_me, "H0", 0.7)
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 32, (3, 16 ** 3)).astype(np.float32)
    vel = rng.normal(0, 1, (3, 16 ** 3)).astype(np.float32)
    rho = rng.uniform(0, 2, (32, 32, 32)).astype(np.float32)
    a = 0.5
    save_data.save_file(rho, pos, vel, 7, a)
    with np.load(tmp_path / "data.7.npz", allow_pickle=False) as z:
        assert sorted(z.files) == sorted(["x1", "x2", "x3", "vx1", "vx2", "vx3", "a", "density"])
        # unit factors of save_data.py:10-11
        up = 7.8 * (100.0 / (32 / 128)) / 10 ** 3
        uv = 0.781 * 100.0 * 0.7 / (a * 32 / 128)
        assert np.allclose(z["x2"], pos[1] * up, rtol=1e-6)
        assert np.allclose(z["vx3"], vel[2] * uv, rtol=1e-6)
        assert float(z["a"]) == a
        assert np.array_equal(z["density"], rho)
    p2, v2, a2 = save_data.from_file(7)
    assert a2 == np.float32(a) and p2.dtype == np.float32 and v2.dtype == np.float32
    assert np.allclose(p2, pos, rtol=1e-6, atol=1e-5)
    assert np.allclose(v2, vel, rtol=1e-6, atol=1e-6)
