"""All six receiver traces (E and H) at every step, against the oracle stepped
one step at a time and sampled at the receivers after each step (reference
output(), function.c:765-808).  The fused loop gathers step j's receivers in
the E launch of step j + 1 and with a launch of its own after a chunk's last
step, so the runs below end chunks early (step_n pieces) and in the middle of
the ring, for the LDS family and the per-cell family."""
import numpy as np
import pytest

from .test_gpu_parity import setup_pair

pytestmark = pytest.mark.gpu

NAMES = ("EX", "EY", "EZ", "HX", "HY", "HZ")


def oracle_traces(S, signal, src, rec, it):
    c = rec[:, 2] * S.p.ix * S.p.iy + rec[:, 1] * S.p.ix + rec[:, 0]
    out = np.zeros((6, len(rec), it))
    src1 = tuple(q + 1 for q in src)
    for n in range(it - 1):
        S.forward(0, signal, src1, 1, it=it, step0=n)
        for q, nm in enumerate(NAMES):
            out[q, :, n] = S.arr(nm).reshape(-1)[c]
    return out


@pytest.mark.parametrize("pieces", [None, (5, 13, 1, 30)])
@pytest.mark.parametrize("family", ["lds", "auto"])
def test_every_trace_sample_matches_oracle_steps(family, pieces, monkeypatch):
    if family == "auto":
        monkeypatch.delenv("MITTET_FDTD_LDS", raising=False)
    it = 45
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(seed=41, it=it, ring=8)
    print(f"traces: {family} family, pieces {pieces}, it {it}", flush=True)
    if pieces is None:
        shot.run()
    else:
        for p in pieces:
            shot.step_n(p)
    assert shot.step == it - 1
    want = oracle_traces(S, signal, src, rec, it)
    got = shot.traces()
    assert got.shape == want.shape
    np.testing.assert_array_equal(got, want)
    assert np.abs(got[3:]).max() > 0
