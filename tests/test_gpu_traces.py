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


@pytest.mark.parametrize("family", ["lds", "auto"])
def test_ten_shot_batch_traces_equal_alone(family, monkeypatch):
    """More shots than one group of 8 gather blocks: ten shots of four grid
    sizes and three lengths, four of them joining after 11 steps, give every
    trace sample, field and volume DFT of each run alone."""
    from mittet_main_amd import Batch
    if family == "auto":
        monkeypatch.delenv("MITTET_FDTD_LDS", raising=False)
    dims = [(26, 24, 20), (30, 22, 24), (24, 28, 22), (28, 26, 26)]
    specs = [dict(ix=dims[q % 4][0], iy=dims[q % 4][1], iz=dims[q % 4][2], seed=50 + q,
                  it=(30, 41, 52)[q % 3]) for q in range(10)]
    print(f"ten-shot batch: {family} family", flush=True)
    alone = []
    for sp in specs:
        m, shot, *_ = setup_pair(ring=8, **sp)
        shot.run()
        alone.append((shot.traces(), shot.volume_dft(), [shot.field(c) for c in range(6)]))
    keep = [setup_pair(ring=8, **sp)[:2] for sp in specs]
    b = Batch(12, chunk=8)
    for m, shot in keep[:6]:
        b.add(shot)
    b.step(11)
    for m, shot in keep[6:]:
        b.add(shot)
    for _ in range(8):
        b.step(8)
    b.sync()
    for m, shot in keep:
        b.remove(shot)
    for q, (m, shot) in enumerate(keep):
        tr, vol, f = alone[q]
        assert shot.step == specs[q]["it"] - 1
        np.testing.assert_array_equal(shot.traces(), tr, err_msg=f"shot {q}")
        np.testing.assert_array_equal(shot.volume_dft(), vol, err_msg=f"shot {q}")
        for c in range(6):
            np.testing.assert_array_equal(shot.field(c), f[c], err_msg=f"shot {q} field {c}")
    b.close()


@pytest.mark.parametrize("family", ["lds", "auto"])
def test_ring_one_traces_many_receivers(family, monkeypatch):
    """Shots without a volume DFT keep E in one slot updated in place; their
    receiver gathers must come after each step's E launch, never inside the
    next one.  256 receivers (a gather loop of 6 x 256 samples per block)
    against the oracle stepped one step at a time, alone, cut into pieces, and
    as three ring-1 shots in a batch beside a ring-8 shot."""
    from mittet_main_amd import Batch
    if family == "auto":
        monkeypatch.delenv("MITTET_FDTD_LDS", raising=False)
    it = 40
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(seed=61, it=it, nrec=256, omegas=())
    print(f"ring-1 traces: {family} family, 256 receivers, it {it}", flush=True)
    assert shot.nf == 0   # no volume DFT: one E slot
    want = oracle_traces(S, signal, src, rec, it)
    shot.run()
    np.testing.assert_array_equal(shot.traces(), want)
    shot.set_zero_eh()
    for p in (7, 1, 19, 12):
        shot.step_n(p)
    np.testing.assert_array_equal(shot.traces(), want)

    specs = [dict(seed=61, nrec=256, omegas=()), dict(seed=62, nrec=200, omegas=(), ix=28),
             dict(seed=63, nrec=256, omegas=(), iz=24), dict(seed=64, nrec=64, ring=8)]
    alone, keep = [], []
    for sp in specs:
        m2, sh2, S2, _, sg2, src2, rec2, *_ = setup_pair(it=it, **sp)
        alone.append(oracle_traces(S2, sg2, src2, rec2, it))
        keep.append((m2, sh2))
    b = Batch(4, chunk=8)
    for _, sh2 in keep:
        b.add(sh2)
    for _ in range(5):
        b.step(8)
    b.sync()
    for q, (_, sh2) in enumerate(keep):
        b.remove(sh2)
        assert sh2.step == it - 1
        np.testing.assert_array_equal(sh2.traces(), alone[q], err_msg=f"batch shot {q}")
    b.close()
