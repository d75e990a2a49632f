"""CPU side of the measurement: the CPU-oracle baseline stays inside the
workload's time steps however fast the host is, and the oracle refuses step
ranges its arrays cannot hold (a range past the last sample once overran the
signal, trace and weight arrays of the 84^3 workload on a 16-thread host).
The forward mode's default step counts fit each workload."""
import os
import sys

import numpy as np
import pytest

import oracle
from mittet_main_amd import configs

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def tiny(it=6):
    return configs.make("tiny", 8, 4, it, [1.0], nrec_line=2, seed=1)


def test_cpu_baseline_stops_at_the_last_step():
    wl = tiny(it=14)
    out = bench.cpu_baseline(wl, 5.0)      # far longer than 12 steps of a 16^3 grid take
    assert out["last_step"] <= wl.it - 2   # timed steps 1 .. it-2 after the warm-up step 0
    assert len(out["windows_with_dft"]) == 3 and len(out["windows_plain"]) == 3
    assert out["kind"] == "port" and out["value"] > 0 and out["plain_step_value"] > 0
    assert out["one_thread"]["with_dft"] > 0 and out["stream_triad_GBps"] > 0
    assert out["sockets"] >= 1 and out["cores"] >= 1
    with pytest.raises(ValueError):
        bench.cpu_baseline(tiny(it=8), 1.0)


def test_oracle_refuses_steps_past_the_arrays():
    wl = tiny(it=6)
    g = wl.grid
    S = oracle.OracleSim(g.ix, g.iy, g.iz, g.ncpml, g.dx, g.dy, g.dz, g.dt, g.omega0, g.cmax)
    S.set_sigma(wl.sigma)
    S.init_cpml()
    S.media_coeff_3d()
    src1 = tuple(c + 1 for c in wl.src)
    w = oracle.dft_weights(wl.it, g.dt, g.omega0, wl.omegas)
    # the whole run fits: steps 0 .. it-2
    tr, _ = S.forward(0, wl.signal, src1, wl.it - 1, rec1=wl.receivers + 1, it=wl.it, weights=w)
    assert np.all(np.isfinite(tr))
    with pytest.raises(ValueError):       # signal[n + 1] past the waveform
        S.forward(0, wl.signal, src1, 1, rec1=wl.receivers + 1, it=wl.it, step0=wl.it - 1)
    with pytest.raises(ValueError):       # trace sample past it
        S.forward(0, np.zeros(wl.it + 5), src1, 1, rec1=wl.receivers + 1, it=wl.it,
                  step0=wl.it)
    with pytest.raises(ValueError):       # weight row past the table
        S.forward(0, None, src1, 2, it=wl.it + 4, step0=wl.it - 1, weights=w)
    with pytest.raises(ValueError):
        S.forward(0, wl.signal, src1, -1, it=wl.it)


def test_forward_defaults_fit_each_workload():
    import argparse
    for name, make, steps, ring in (("roofline_256", configs.roofline_256, 2048, 128),
                                    ("plain_64", configs.plain_64, 1024, 64)):
        wl_steps = make().steps
        a = argparse.Namespace(workload=name, steps=0, ring=0, warmup=128, profile_steps=0)
        assert bench.forward_step_counts(a, wl_steps) is None
        assert (a.steps, a.ring, a.profile_steps) == (steps, ring, ring)
        assert a.warmup + a.steps + a.profile_steps <= wl_steps
    a = argparse.Namespace(workload="plain_64", steps=2048, ring=0, warmup=128, profile_steps=0)
    assert "exceeds" in bench.forward_step_counts(a, configs.plain_64().steps)
