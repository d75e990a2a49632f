"""Serving: request stream, admission control and the continuous-batching loop.

CPU tests cover the stream generators and the HBM accounting; GPU tests check
that every shot served in the mix (queued under a tight budget, joining a
running batch) returns exactly the spectra of the same shot run alone, by the
fused path and by the reference-named per-step operators, and that gradient
requests served beside the forward batch (in a worker thread, under a budget
that forces recursive checkpointing) return exactly the gradient computed alone.
"""
import numpy as np
import pytest

from mittet_main_amd import serving


def small_stream(n=6, seed=3):
    return serving.request_stream(seed, n, rate=200.0, sizes=(16, 20, 24), ncpml=6, max_shots=4,
                                  nf=3, nrec_side=3, steps_per_cell=5)


def test_stream_is_seeded_and_poisson_ordered():
    a = serving.request_stream(7, 50)
    b = serving.request_stream(7, 50)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(np.array_equal(x.sources, y.sources) for x, y in zip(a, b))
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    assert 1.0 < arr[-1] < 6.0                      # 50 arrivals at 20/s
    for r in a:
        assert r.n_int in (96, 128, 192, 256)
        assert 1 <= r.nshots <= 16
        assert r.it == 30 * r.n_int
        n = r.n_int + 2 * r.ncpml
        assert r.sources.min() >= r.ncpml and r.sources.max() < n - r.ncpml


def test_request_bytes_counts_model_and_every_shot():
    req = small_stream(1)[0]
    grid, sig = serving.build_model(req)
    one = serving.shot_bytes(grid, req.it, nrec=len(req.receivers))
    assert serving.request_bytes(req, grid) == grid.ncells * 24 + req.nshots * one
    assert sig.shape == grid.shape and sig.min() > 0


def test_request_work_orders_queue():
    reqs = small_stream(8)
    w = [serving.request_work(r) for r in reqs]
    assert all(x > 0 for x in w)
    r = reqs[0]
    assert w[0] == (r.n_int + 2 * r.ncpml) ** 3 * (r.it - 1) * r.nshots


@pytest.mark.gpu
@pytest.mark.parametrize("policy,target_cells,launch", [("swf", 24e6, "cache"),
                                                         ("swf", 24e6, "merged"),
                                                         ("swf", 1.0, "interleave"),
                                                         ("fifo", 60e3, "interleave")])
def test_served_mix_equals_alone(policy, target_cells, launch):
    reqs = small_stream()
    # budget for roughly two requests at a time: forces queueing and joins mid-run;
    # target_cells 1 admits one shot at a time, 60e3 about two small grids; the
    # shots in flight advance in groups that fit the cache (all of them, at these
    # sizes), in one merged batch, or each in turn
    budget = 2.2 * max(serving.request_bytes(r, serving.build_model(r)[0]) for r in reqs)
    srv = serving.Server(budget_bytes=budget, chunk=8, target_cells=target_cells, policy=policy,
                         launch=launch)
    served = srv.run(reqs, realtime=False)
    srv.close()
    alone = serving.run_sequential(reqs, mode="fused", realtime=False)
    ops = serving.run_sequential(reqs[:2], mode="ops", realtime=False)
    assert served["shots"] == sum(r.nshots for r in reqs)
    assert served["peak_hbm_gb"] * 1e9 <= budget
    for rid, res in alone["results"].items():
        got = served["results"][rid]
        assert len(got.spectra) == len(res.spectra)
        for a, b in zip(got.spectra, res.spectra):
            np.testing.assert_array_equal(a, b)
    for rid, res in ops["results"].items():
        for a, b in zip(res.spectra, alone["results"][rid].spectra):
            np.testing.assert_array_equal(a, b)
    assert served["p99_s"] >= served["p50_s"] > 0


def test_fwi_stream_mixes_kinds_on_one_grid():
    a = serving.fwi_stream(5, 200)
    assert [r.arrival for r in a] == [r.arrival for r in serving.fwi_stream(5, 200)]
    kinds = [r.kind for r in a]
    assert 0.2 < kinds.count("gradient") / len(a) < 0.4
    for r in a:
        assert r.n_int == 384 and r.ncpml == 16 and r.it == 1000 and len(r.freqs_hz) == 8
        assert 1 <= r.nshots <= 2 and r.d_obs is None
    g = [r for r in a if r.kind == "gradient"][0]
    assert serving.request_work(g) == 6 * (416 ** 3) * 999 * g.nshots


def test_gradient_admission_picks_levels_for_the_budget():
    req = serving.fwi_stream(5, 1, n_int=32, ncpml=6, it=400, nrec_side=3)[0]
    grid, _ = serving.build_model(req)
    fixed = grid.ncells * serving.MODEL_BYTES_PER_CELL + \
        serving.shot_bytes(grid, req.it, nrec=len(req.receivers))
    b = [serving.plan_bytes(grid, req.it - 1, len(req.receivers), lv) for lv in (1, 2, 3)]
    assert b[0] > b[1] > b[2]
    assert serving.gradient_bytes(req, fixed + b[0]) == (1, fixed + b[0])
    assert serving.gradient_bytes(req, fixed + b[0] - 1) == (2, fixed + b[1])
    assert serving.gradient_bytes(req, fixed + b[2])[0] == 3
    assert serving.gradient_bytes(req, fixed) is None


def grad_stream():
    reqs = serving.fwi_stream(9, 6, n_int=14, ncpml=6, rate=500.0, grad_frac=0.5, max_shots=2,
                              nf=3, it=150, nrec_side=3)
    assert {r.kind for r in reqs} == {"forward", "gradient"}
    for r in reqs:
        if r.kind == "gradient":
            r.d_obs = serving.observed_data(r)
    return reqs


@pytest.mark.gpu
def test_plan_bytes_mirror_the_library():
    import mittet_main_amd._lib as lib
    req = grad_stream()[0]
    grid, sig = serving.build_model(req)
    m = serving.Model(grid)
    sh = m.shot(req.it, np.zeros(req.it), tuple(int(v) for v in req.sources[0]), 0, req.receivers)
    for lv in (1, 2, 3, 5):
        assert serving.plan_bytes(grid, req.it - 1, len(req.receivers), lv) == \
            lib.load().fwi3d_grad_plan_bytes(sh._h, lv)
    sh.close()
    m.close()


@pytest.mark.gpu
def test_served_gradients_equal_alone():
    reqs = grad_stream()
    grads = [r for r in reqs if r.kind == "gradient"]
    grid, _ = serving.build_model(grads[0])
    fixed = grid.ncells * serving.MODEL_BYTES_PER_CELL + \
        serving.shot_bytes(grid, grads[0].it, nrec=len(grads[0].receivers))
    lv1 = fixed + serving.plan_bytes(grid, grads[0].it - 1, len(grads[0].receivers), 1)
    fwd = max(serving.request_bytes(r, serving.build_model(r)[0]) for r in reqs)
    # too small for the sqrt(Nt) plan; a recursive plan leaves room for forwards
    budget = lv1 - 1
    lv2 = fixed + serving.plan_bytes(grid, grads[0].it - 1, len(grads[0].receivers), 2)
    assert budget - lv2 > fwd
    print(f"served gradients: {len(grads)} of {len(reqs)} requests, budget {budget / 1e6:.1f} MB",
          flush=True)
    srv = serving.Server(budget_bytes=budget, chunk=8)
    served = srv.run(reqs, realtime=False)
    srv.close()
    alone = serving.run_sequential(reqs, mode="fused", realtime=False, budget_bytes=1e12)
    assert served["peak_hbm_gb"] * 1e9 <= budget
    assert served["gradient"]["requests"] == len(grads)
    for rid, res in alone["results"].items():
        got = served["results"][rid]
        if res.grad is not None:
            assert res.levels == 1 and got.levels >= 2       # the budget forced recomputation
            assert got.misfit == res.misfit > 0
            np.testing.assert_array_equal(got.grad, res.grad)
        else:
            for a, b in zip(got.spectra, res.spectra):
                np.testing.assert_array_equal(a, b)
