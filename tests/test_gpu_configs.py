"""The benchmark workloads at their full grid sizes, with bounded time loops.
The oracle cannot step these grids in seconds, so each check is a property
that holds at any size and that a wrong index, a lost flush or a shared
buffer breaks; the same kernels are pinned bit for bit to the oracle on
small grids (test_gpu_parity.py, test_gpu_traces.py).

* 256^3 + 16 (288^3), Nt 8000 cut to 256 steps, 8 volume frequencies,
  256 receivers: the volume accumulators and traces are the same bits for
  E rings of 128 and 16 slots (one flush vs sixteen), and exactly linear in
  the source amplitude.
* served stream on {96, 128, 192}^3 + 12 (Nt cut to 160-300 steps):
  every request's receiver spectra in the continuous batch equal the same
  shots run alone, bit for bit.
* 128^3 + 12 (152^3) gradient, 2 of the 8 shots, Nt 1600 of 3840:
  dJ/dlog(sigma) against central finite differences (h = 1e-6) at four
  seeded cells where the gradient is large and at the source cell, <= 1e-6.
* 384^3 + 16 (416^3), 1 shot, Nt 200: gradients from 2- and 3-level
  checkpoint recursion are the same bits, and the forward is linear.
Each test prints its case before it starts."""
import numpy as np
import pytest

from mittet_main_amd import Model, configs, serving
from mittet_main_amd.adjoint import GradientProblem, ShotSpec

pytestmark = pytest.mark.gpu


def _exactly_doubled(a, b, what):
    """b == 2 a, bit for bit, wherever |a| > 1e-200, and to 1e-30 of the largest
    value elsewhere.  Ahead of the wave front the fields pass through
    subnormal numbers, whose rounding is absolute (2^-1074) and so does not
    scale with the source; the fictitious field grows by many orders of
    magnitude as the front arrives, and values of 1e-260 at 416^3 were seen
    to carry that rounding at 1e-12 relative.  Values far above it cannot."""
    fa, fb = a.view(np.float64).reshape(-1), b.view(np.float64).reshape(-1)
    big = np.abs(fa) > 1e-200
    assert np.count_nonzero(big) > 0, what
    np.testing.assert_array_equal(fb[big], 2.0 * fa[big], err_msg=what)
    top = np.abs(fb).max()
    assert np.abs(fb[~big] - 2.0 * fa[~big]).max(initial=0.0) <= 1e-30 * top, what


def test_roofline_256_volume_dft_ring_independent_and_linear():
    wl = configs.roofline_256(it=257)
    g = wl.grid
    print(f"288^3 forward, {wl.it - 1} steps, {len(wl.omegas)} volume frequencies, "
          f"{len(wl.receivers)} receivers: ring 128 vs 16, source x2", flush=True)
    m = Model(g)
    m.media_coeff_3d(wl.sigma)

    def run(scale, ring):
        sh = m.shot(wl.it, scale * wl.signal, wl.src, 0, wl.receivers, wl.omegas, ring=ring)
        sh.run()
        out = sh.volume_dft(), sh.traces()
        sh.close()
        return out

    a, ta = run(1.0, 128)
    assert np.abs(a).max() > 0 and np.abs(ta).max() > 0
    b, tb = run(1.0, 16)
    for d in range(3):
        np.testing.assert_array_equal(b[d], a[d], err_msg=f"component {d}: ring 16 vs 128")
    np.testing.assert_array_equal(tb, ta)
    del b
    c, tc = run(2.0, 128)
    for d in range(3):
        _exactly_doubled(a[d], c[d], f"component {d}")
    _exactly_doubled(ta, tc, "traces")
    m.close()


def _served_requests():
    """Six requests on three grid sizes with unequal time loops, arriving
    together, so batch membership changes while shots of every size run."""
    reqs = serving.request_stream(seed=21, n_requests=6, sizes=(96, 128), max_shots=3)
    steps = (300, 220, 260, 180, 240, 200)
    for r, it in zip(reqs, steps):
        r.it = it
        r.arrival = 0.0
    big = serving.request_stream(seed=22, n_requests=1, sizes=(192,), max_shots=1)[0]
    big.rid, big.it, big.arrival = 6, 160, 0.0
    return reqs + [big]


def test_served_stream_equals_each_shot_alone():
    reqs = _served_requests()
    print("served stream: " + ", ".join(f"{r.n_int}^3 x{r.nshots} Nt {r.it}" for r in reqs),
          flush=True)
    srv = serving.Server(budget_bytes=200e9, chunk=32)
    try:
        served = srv.run(reqs, realtime=False)["results"]
    finally:
        srv.close()
    alone = serving.run_sequential(reqs, mode="fused", realtime=False)["results"]
    assert len(served) == len(reqs)
    for r in reqs:
        got, want = served[r.rid].spectra, alone[r.rid].spectra
        assert len(got) == r.nshots
        for q in range(r.nshots):
            assert np.abs(want[q]).max() > 0
            np.testing.assert_array_equal(got[q], want[q], err_msg=f"request {r.rid} shot {q}")


def test_gradient_128_matches_central_fd():
    wl = configs.gradient_128(it=1600, nshots=2)
    print(f"152^3 gradient: 2 shots, {len(wl.receivers)} receivers, Nt {wl.it}, "
          "FD h = 1e-6 at 4 cells + the source", flush=True)
    specs = [ShotSpec(tuple(int(v) for v in s), wl.signal) for s in wl.sources]
    P = GradientProblem(wl.grid, specs, wl.receivers, wl.it, wl.omegas)
    try:
        P.d_obs = P.forward_data(wl.sigma_true)
        logs = np.log(wl.sigma_start)
        J, grad = P.value_and_grad(logs)
        assert J > 0 and np.all(np.isfinite(grad))
        gmax = np.abs(grad).max()
        rng = np.random.default_rng(128)
        big = np.argwhere(np.abs(grad) > 1e-2 * gmax)
        cells = [tuple(int(v) for v in big[q]) for q in rng.choice(len(big), 4, replace=False)]
        src = specs[0].src
        cells.append((src[2], src[1], src[0]))
        h = 1e-6
        for (k, j, i) in cells:
            lp = logs.copy()
            lp[k, j, i] += h
            jp = P.misfit(np.exp(lp))
            lp[k, j, i] -= 2 * h
            jm = P.misfit(np.exp(lp))
            fd = (jp - jm) / (2 * h)
            g = grad[k, j, i]
            err = abs(fd - g) / max(abs(g), 1e-3 * gmax)
            print(f"  cell {(k, j, i)}: adjoint {g:.9e} FD {fd:.9e} rel {err:.2e}", flush=True)
            assert err <= 1e-6, ((k, j, i), g, fd, err)
    finally:
        P.close()


def test_fwi_384_levels_bitwise_and_forward_linear():
    wl = configs.fwi_384(it=201, nshots=1)
    g = wl.grid
    print(f"416^3 gradient: 1 shot, Nt {wl.it}, checkpoint levels 2 vs 3; forward x2",
          flush=True)
    spec = ShotSpec(tuple(int(v) for v in wl.sources[0]), wl.signal)
    grads = []
    for levels in (2, 3):
        P = GradientProblem(g, [spec], wl.receivers, wl.it, wl.omegas, levels=levels)
        try:
            if levels == 2:
                d_obs = P.forward_data(wl.sigma_true)
            P.d_obs = d_obs
            J, gr = P.value_and_grad(np.log(wl.sigma_start))
            grads.append((J, gr))
        finally:
            P.close()
    assert grads[0][0] == grads[1][0] and grads[0][0] > 0
    np.testing.assert_array_equal(grads[1][1], grads[0][1])
    assert np.abs(grads[0][1]).max() > 0
    m = Model(g)
    m.media_coeff_3d(wl.sigma_true)
    tr = []
    for scale in (1.0, 2.0):
        sh = m.shot(wl.it, scale * wl.signal, spec.src, 0, wl.receivers)
        sh.run()
        tr.append(sh.traces())
        sh.close()
    m.close()
    _exactly_doubled(tr[0], tr[1], "416^3 traces")
