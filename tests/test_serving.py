"""Serving: request stream, admission control and the continuous-batching loop.

CPU tests cover the stream generator and the HBM accounting; GPU tests check
that every shot served in the mix (queued under a tight budget, joining a
running batch) returns exactly the spectra of the same shot run alone, by the
fused path and by the reference-named per-step operators.
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
@pytest.mark.parametrize("policy,target_cells", [("swf", 24e6), ("swf", 1.0), ("fifo", 60e3)])
def test_served_mix_equals_alone(policy, target_cells):
    reqs = small_stream()
    # budget for roughly two requests at a time: forces queueing and joins mid-run;
    # target_cells 1 admits one shot at a time, 60e3 about two small grids
    budget = 2.2 * max(serving.request_bytes(r, serving.build_model(r)[0]) for r in reqs)
    srv = serving.Server(budget_bytes=budget, chunk=8, target_cells=target_cells, policy=policy)
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
