"""The fused E+H interior step against the oracle and the two-launch schedule.

With MITTET_FDTD_FUSE=1 (set for every test here), shots of the LDS family
whose E lives in a ring of more than one slot run one launch over the
CPML-free interior that computes E of plane p and then H of
plane p - 2; its blocks recompute the E halo their H update needs and read H
from one buffer while writing the other (the current buffer alternates with
every step).  Everything below must come out bit for bit: fields, all six
traces at every step, and the volume DFT, over partial tiles, z-chunks of 1 to
8 planes, the minimum CPML width (3), odd ring sizes and odd run pieces (the H
buffer is the other one after every odd piece), a source on the halo of a
neighbouring tile, per-operator calls and set_field between fused steps, the
event-timed step loop, and a batch."""
import numpy as np
import pytest

import oracle
from mittet_main_amd import Batch

from .test_gpu_parity import OM0, setup_pair
from .test_gpu_traces import oracle_traces

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fused(monkeypatch):
    monkeypatch.setenv("MITTET_FDTD_FUSE", "1")

# (dims, ncpml, z-chunk, ring, source x offset from the region's x start)
CASES = [((90, 40, 33), 6, "8", 5, 64),      # source on tile 2's first column (tile 1's halo)
         ((100, 30, 26), 3, "3", 7, 63),     # ncpml 3; source on tile 1's last column
         ((41, 31, 41), 4, "1", 8, 5),       # one-plane chunks
         ((84, 44, 30), 5, "5", 6, 20)]


def is_fused(shot) -> bool:
    """Runs one event-timed step and reports whether it used the fused kernel."""
    return shot.profile_steps(1)["fused"] == 1.0


@pytest.mark.parametrize("dims,nc,lkb,ring,sx", CASES)
def test_fused_every_trace_and_field_match_oracle(dims, nc, lkb, ring, sx, monkeypatch):
    monkeypatch.setenv("MITTET_FDTD_LKB", lkb)
    ix, iy, iz = dims
    it = 40
    xlo = (nc + 15) // 16 * 16
    src = (xlo + sx, iy // 2 - 1, iz // 2 + 1)
    print(f"fused vs oracle: {dims} nc {nc} lkb {lkb} ring {ring} src {src}", flush=True)
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(ix=ix, iy=iy, iz=iz, nc=nc, seed=61,
                                                            it=it, ring=ring, src=src)
    assert is_fused(shot)
    for piece in (4, 13, 1, 6):
        shot.step_n(piece)
    shot.run()
    assert shot.step == it - 1
    want = oracle_traces(S, signal, src, rec, it)
    np.testing.assert_array_equal(shot.traces(), want)
    for q, nm in enumerate(oracle.FIELD_NAMES):
        np.testing.assert_array_equal(shot.field(q), S.arr(nm), err_msg=nm)
    S2 = oracle.OracleSim(ix, iy, iz, nc, 50.0, 40.0, 30.0, dt, OM0, m.grid.cmax)
    S2.set_sigma(sig)
    S2.init_cpml()
    S2.media_coeff_3d()
    w = oracle.dft_weights(it, dt, OM0, om)
    _, acc = S2.forward(0, signal, tuple(c + 1 for c in src), it - 1, rec1=rec + 1, it=it,
                        weights=w)
    np.testing.assert_array_equal(shot.volume_dft().reshape(3, len(om), -1), acc)
    assert np.abs(shot.traces()[3:]).max() > 0
    shot.close()
    m.close()


def test_fused_equals_two_launch_schedule_with_operators_between(monkeypatch):
    """Per-operator calls and set_field on H between fused pieces: the current
    H buffer is the one they touch, and both buffers agree where no step
    writes."""
    ix, iy, iz, it, ring = 88, 36, 30, 36, 6
    runs = []
    for fuse in ("1", "0"):
        monkeypatch.setenv("MITTET_FDTD_FUSE", fuse)
        m, shot, S, sig, signal, src, rec, om, dt = setup_pair(ix=ix, iy=iy, iz=iz, seed=62,
                                                                it=it, ring=ring)
        assert is_fused(shot) == (fuse == "1")
        shot.step_n(6)
        rng = np.random.default_rng(5)
        hz = shot.field(5) + 1e-3 * rng.standard_normal(m.grid.shape)
        shot.set_field(5, hz)
        shot.e_field_cpml42()
        shot.epml_abcs4()
        shot.h_field_cpml42()
        shot.hpml_abcs4()
        shot.step_n(9)
        mid = [shot.field(q) for q in range(6)]
        shot.run()
        runs.append((mid, [shot.field(q) for q in range(6)], shot.traces(), shot.volume_dft()))
        shot.close()
        m.close()
    (ma, fa, ta, va), (mb, fb, tb, vb) = runs
    for q in range(6):
        np.testing.assert_array_equal(ma[q], mb[q], err_msg=f"mid {q}")
        np.testing.assert_array_equal(fa[q], fb[q], err_msg=f"end {q}")
    np.testing.assert_array_equal(ta, tb)
    np.testing.assert_array_equal(va, vb)


def test_fused_batch_equals_each_shot_alone():
    """Three fused shots (two grids, a source on a tile boundary), one joining
    the batch after an odd number of steps, equal their runs alone."""
    specs = [dict(ix=90, iy=40, iz=28, seed=70, src=(80, 20, 14)),
             dict(ix=90, iy=40, iz=28, seed=71, src=(79, 21, 12)),
             dict(ix=72, iy=34, iz=30, seed=72, src=(40, 17, 15))]
    it, ring = 33, 8
    alone = []
    for sp in specs:
        m, shot, *_ = setup_pair(it=it, ring=ring, **sp)
        assert is_fused(shot)
        shot.set_zero_eh()
        shot.run()
        alone.append((shot.traces(), [shot.field(q) for q in range(6)], shot.volume_dft()))
        shot.close()
        m.close()
    held = [setup_pair(it=it, ring=ring, **sp) for sp in specs]
    b = Batch(4, chunk=ring)
    b.add(held[0][1])
    b.add(held[1][1])
    b.step(5)
    b.add(held[2][1])
    b.step(it)
    b.sync()
    for q, h in enumerate(held):
        shot = h[1]
        b.remove(shot)
        if q == 2:
            shot.run()
        tr, fl, vd = alone[q]
        np.testing.assert_array_equal(shot.traces(), tr, err_msg=f"shot {q}")
        for c in range(6):
            np.testing.assert_array_equal(shot.field(c), fl[c], err_msg=f"shot {q} field {c}")
        np.testing.assert_array_equal(shot.volume_dft(), vd, err_msg=f"shot {q}")
    b.close()
    for h in held:
        h[1].close()
        h[0].close()
