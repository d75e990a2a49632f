"""HIP path vs the CPU oracle on the same seeded inputs (through the C ABI).

The library is built with -ffp-contract=off and recomputes the reference's
coefficients with the reference's operation order, so fields, psi-driven
corrections, traces and the volume running DFT must agree BIT FOR BIT.  The
spectra whose weights the GPU evaluates itself (exp/sincos on the device)
agree to 1e-12 relative.  Sizes are chosen so the oracle finishes in seconds.
"""
import numpy as np
import pytest

import oracle
from mittet_main_amd import Batch, Grid, Model, configs
from mittet_main_amd.fwi3d import firstderiv_gauss

pytestmark = pytest.mark.gpu

OM0 = 2 * np.pi


def setup_pair(ix=30, iy=26, iz=22, nc=5, dx=50.0, dy=40.0, dz=30.0, seed=0, it=64,
               nrec=5, omegas=(0.5, 1.3, 3.0), ring=0, src=None, comp=0):
    rng = np.random.default_rng(seed)
    sig = np.exp(rng.uniform(np.log(0.02), np.log(3.0), (iz, iy, ix)))
    dt, cmax = oracle.lattice_time_3d(float(sig.min()), OM0, dx, dy, dz, factor=0.99)
    g = Grid(ix, iy, iz, nc, dx, dy, dz, dt, OM0, cmax)
    m = Model(g)
    m.media_coeff_3d(sig)
    S = oracle.OracleSim(ix, iy, iz, nc, dx, dy, dz, dt, OM0, cmax)
    S.set_sigma(sig)
    S.init_cpml()
    S.media_coeff_3d()
    signal = firstderiv_gauss(it, dt, 6.0) * 1e3
    src = src or (ix // 2, iy // 2 - 1, iz // 2 + 1)
    rec = np.stack([rng.integers(0, ix, nrec), rng.integers(0, iy, nrec),
                    rng.integers(0, iz, nrec)], axis=1).astype(np.int32)
    rec[0] = src
    om = 2 * np.pi * np.asarray(omegas)
    shot = m.shot(it, signal, src, comp, rec, om if len(om) else None, ring)
    return m, shot, S, sig, signal, src, rec, om, dt


def test_profiles_equal_oracle():
    m, shot, S, *_ = setup_pair()
    for a in range(3):
        for nm in oracle.PROFILE_IDS:
            np.testing.assert_array_equal(m.profile(a, nm), S.prof(a, nm), err_msg=f"{a} {nm}")


def test_each_operator_bitwise():
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(seed=1)
    rng = np.random.default_rng(2)
    for q, nm in enumerate(oracle.FIELD_NAMES):
        v = rng.standard_normal(S.shape) * (1e-3 if q < 3 else 1.0)
        S.arr(nm)[...] = v
        shot.set_field(q, v)
    src1 = tuple(c + 1 for c in src)
    rec1 = rec + 1
    traces = np.zeros((3, len(rec), shot.it))
    for step in range(6):
        for op in ("e", "epml", "src", "h", "hpml", "out"):
            if op == "e":
                shot.e_field_cpml42(1); S.e_field_cpml42(1)
            elif op == "epml":
                shot.epml_abcs4(); S.epml_abcs4()
            elif op == "src":
                shot.read_source_3d(step); S.read_source_3d(0, signal, src1, step)
            elif op == "h":
                shot.h_field_cpml42(1); S.h_field_cpml42(1)
            elif op == "hpml":
                shot.hpml_abcs4(); S.hpml_abcs4()
            else:
                shot.output(step)
                oracle.lib().or_output(S._s, len(rec), oracle._ip(np.ascontiguousarray(rec1, np.int32)),
                                       shot.it, step, *[oracle._dp(traces[c]) for c in range(3)])
            for q, nm in enumerate(oracle.FIELD_NAMES):
                np.testing.assert_array_equal(shot.field(q), S.arr(nm), err_msg=f"step {step} {op} {nm}")
    np.testing.assert_array_equal(shot.traces()[:3], traces)


def test_backprop_sign_variants_bitwise():
    m, shot, S, *_ = setup_pair(seed=3)
    rng = np.random.default_rng(4)
    for q, nm in enumerate(oracle.FIELD_NAMES):
        v = rng.standard_normal(S.shape)
        S.arr(nm)[...] = v
        shot.set_field(q, v)
    for _ in range(3):
        shot.e_field_cpml42(-1); S.e_field_cpml42(-1)
        shot.h_field_cpml42(-1); S.h_field_cpml42(-1)
    for q, nm in enumerate(oracle.FIELD_NAMES):
        np.testing.assert_array_equal(shot.field(q), S.arr(nm), err_msg=nm)


@pytest.mark.parametrize("family", ["lds", "auto", "lds-nt", "auto-nt"])
@pytest.mark.parametrize("ring,pieces", [(0, None), (8, (5, 3, 19, 8, 100)), (1, (7, 50, 100))])
def test_fused_run_bitwise(ring, pieces, family, monkeypatch):
    """Forced LDS family, and the size-based choice (per-cell kernels alone on
    this small grid); "-nt": the kernels with non-temporal field loads and
    stores that grids beyond the Infinity Cache take (MITTET_FDTD_NT=1)."""
    if family.startswith("auto"):
        monkeypatch.delenv("MITTET_FDTD_LDS", raising=False)
    monkeypatch.setenv("MITTET_FDTD_NT", "1" if family.endswith("-nt") else "0")
    it = 90
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(seed=5, it=it, ring=ring)
    if pieces is None:
        shot.run()
    else:
        for p in pieces:
            shot.step_n(p)
    assert shot.step == it - 1
    w = oracle.dft_weights(it, dt, OM0, om)
    tr, acc = S.forward(0, signal, tuple(c + 1 for c in src), it - 1, rec1=rec + 1, it=it, weights=w)
    got = shot.traces()
    np.testing.assert_array_equal(got[:3], tr)
    for q, nm in enumerate(oracle.FIELD_NAMES):
        np.testing.assert_array_equal(shot.field(q), S.arr(nm), err_msg=nm)
        if q >= 3:   # H traces = H at the receivers after each step (final sample check)
            c = rec[:, 2] * S.p.ix * S.p.iy + rec[:, 1] * S.p.ix + rec[:, 0]
            np.testing.assert_array_equal(got[q][:, it - 2], S.arr(nm).reshape(-1)[c])
    vol = shot.volume_dft()
    np.testing.assert_array_equal(vol.reshape(3, len(om), -1), acc)


def test_receiver_spectra_vs_oracle():
    it = 120
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(seed=6, it=it, omegas=())
    shot.run()
    tr = shot.traces()
    freqs = 2 * np.pi * np.array([0.0, 0.25, 1.0, 2.5, 7.0])
    spec = shot.receiver_dft(freqs, h_factor=False)
    w = oracle.dft_weights(it, dt, OM0, freqs)
    want = np.einsum("crn,nf->crf", tr, w)
    np.testing.assert_allclose(spec, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    # f_to_d_h factor on the magnetic rows
    spec_h = shot.receiver_dft(freqs[1:], h_factor=True)
    fac = np.sqrt(-2 * OM0 / (1j * freqs[1:]))
    np.testing.assert_allclose(spec_h[3:], want[3:, :, 1:] * fac, rtol=1e-12,
                               atol=1e-12 * np.abs(want[3:]).max())


def test_laplace_to_freq_vs_oracle():
    it = 100
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(seed=7, it=it, omegas=())
    shot.run()
    tr = shot.traces()
    strue = oracle.ricker(it, dt)
    nb = 20
    got = shot.laplace_to_freq(nb, strue)
    for c in range(3):
        for r in range(len(rec)):
            want = oracle.laplace_to_freq(tr[c, r], signal, strue, dt, OM0, nb)
            for key in ("E", "G", "OBS"):
                ref = want[key]
                np.testing.assert_allclose(got[key][c, r], ref, rtol=1e-10,
                                           atol=1e-10 * np.abs(ref).max() + 1e-300,
                                           err_msg=f"{key} {c} {r}")
            np.testing.assert_allclose(got["J"], want["J"], rtol=1e-10)


@pytest.mark.parametrize("families", ["lds", "mixed", "lds-nt"])
def test_batch_mixed_sizes_equal_alone(families, monkeypatch):
    """Shots of different grid sizes in one persistent launch give the same
    bits as each run alone; one joins mid-way (continuous batching).  "mixed":
    the second shot runs the per-cell family alone (as grids below 2.5 M cells
    do), the others the LDS family, in the same launches.  "lds-nt": the
    batch launches load and store non-temporally (MITTET_FDTD_NT=1), the
    lone runs do not."""
    monkeypatch.setenv("MITTET_FDTD_NT", "0")
    specs = [dict(ix=30, iy=26, iz=22, seed=10, it=70), dict(ix=24, iy=28, iz=20, seed=11, it=50),
             dict(ix=34, iy=22, iz=26, seed=12, it=90)]

    def make(q, sp):
        if families == "mixed" and q == 1:
            monkeypatch.setenv("MITTET_FDTD_NOLDS", "1")
        out = setup_pair(ring=8, **sp)
        monkeypatch.delenv("MITTET_FDTD_NOLDS", raising=False)
        return out

    alone, mixed = [], []
    for q, sp in enumerate(specs):
        m, shot, *_ = make(q, sp)
        shot.run()
        alone.append((shot.traces(), shot.volume_dft(), [shot.field(q) for q in range(6)]))
    keep = []
    if families == "lds-nt":
        monkeypatch.setenv("MITTET_FDTD_NT", "1")
    b = Batch(8, chunk=8)
    for q, sp in enumerate(specs):
        m, shot, *_ = make(q, sp)
        keep.append((m, shot))
    b.add(keep[0][1])
    b.add(keep[1][1])
    b.step(13)
    b.add(keep[2][1])         # joins while the others are mid-run
    for _ in range(20):
        b.step(8)
    b.sync()
    for (m, shot) in keep:
        b.remove(shot)
    for q, (m, shot) in enumerate(keep):
        tr, vol, f = alone[q]
        assert shot.step == specs[q]["it"] - 1
        np.testing.assert_array_equal(shot.traces(), tr)
        np.testing.assert_array_equal(shot.volume_dft(), vol)
        for c in range(6):
            np.testing.assert_array_equal(shot.field(c), f[c])
    b.close()


def test_plain_64cube_whole_run_vs_oracle():
    """The 64^3 + 10-cell CPML workload (seeded layers + boxes), its whole
    time loop (1999 steps, 4 volume frequencies, 32 receivers): traces and
    volume spectra bit-identical to the oracle (a few seconds of the oracle
    on the host's threads)."""
    wl = configs.plain_64()
    it = wl.it
    print(f"84^3 workload vs oracle: {it - 1} steps, {oracle.num_threads()} oracle threads", flush=True)
    g = wl.grid
    m = Model(g)
    m.media_coeff_3d(wl.sigma)
    shot = m.shot(it, wl.signal, wl.src, 0, wl.receivers, wl.omegas)
    shot.run()
    S = oracle.OracleSim(g.ix, g.iy, g.iz, g.ncpml, g.dx, g.dy, g.dz, g.dt, g.omega0, g.cmax)
    S.set_sigma(wl.sigma)
    S.init_cpml()
    S.media_coeff_3d()
    w = oracle.dft_weights(it, g.dt, g.omega0, wl.omegas)
    tr, acc = S.forward(0, wl.signal, tuple(c + 1 for c in wl.src), it - 1,
                        rec1=wl.receivers + 1, it=it, weights=w)
    np.testing.assert_array_equal(shot.traces()[:3], tr)
    np.testing.assert_array_equal(shot.volume_dft().reshape(acc.shape), acc)


def test_full_size_256_linearity_and_zero():
    """At the 256^3 + 16 roofline size: doubling the source doubles every trace
    sample exactly; a zero source leaves every field zero (80 steps)."""
    wl = configs.roofline_256(it=81)
    g = wl.grid
    m = Model(g)
    m.media_coeff_3d(wl.sigma)
    rec = wl.receivers
    a = m.shot(81, wl.signal, wl.src, 0, rec)
    a.run()
    ta = a.traces()
    a.close()
    b = m.shot(81, 2.0 * wl.signal, wl.src, 0, rec)
    b.run()
    tb = b.traces()
    # exact doubling wherever no intermediate underflows; subnormals (< 2.2e-308)
    # carry absolute rounding only
    normal = np.abs(ta) > 1e-290
    np.testing.assert_array_equal(tb[normal], 2.0 * ta[normal])
    assert np.abs(tb[~normal] - 2.0 * ta[~normal]).max(initial=0.0) < 1e-300
    b.close()
    z = m.shot(81, 0.0 * wl.signal, wl.src, 0, rec)
    z.run()
    assert not np.any(z.traces())
    assert not np.any(z.field(0))
    z.close()
    assert np.abs(ta).max() > 0


@pytest.mark.parametrize("dims,pair", [((29, 26, 22), 0), ((31, 23, 25), 0), ((30, 26, 22), 1),
                                       ((32, 23, 25), 1)])
def test_fused_run_bitwise_tilings(dims, pair, monkeypatch):
    """Ragged tiles of the one-cell-per-lane fused kernels, and the opt-in
    two-cells-per-lane (x-pair) kernels, give the oracle's bits."""
    monkeypatch.setenv("MITTET_FDTD_PAIR", str(pair))
    ix, iy, iz = dims
    it = 70
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(ix=ix, iy=iy, iz=iz, seed=21, it=it,
                                                            ring=8)
    shot.run()
    w = oracle.dft_weights(it, dt, OM0, om)
    tr, acc = S.forward(0, signal, tuple(c + 1 for c in src), it - 1, rec1=rec + 1, it=it, weights=w)
    np.testing.assert_array_equal(shot.traces()[:3], tr)
    for q, nm in enumerate(oracle.FIELD_NAMES):
        np.testing.assert_array_equal(shot.field(q), S.arr(nm), err_msg=nm)
    np.testing.assert_array_equal(shot.volume_dft().reshape(3, len(om), -1), acc)


@pytest.mark.parametrize("dims,lkb", [((41, 31, 41), "1"), ((41, 31, 41), "3"),
                                      ((41, 31, 41), "16"), ((40, 30, 42), "5")])
def test_lds_chunk_lengths_bitwise(dims, lkb, monkeypatch):
    """The LDS kernels' z-chunk length (chosen per grid; MITTET_FDTD_LKB forces
    it) only changes which block marches which planes: 1-, 3-, 5- and
    16-plane chunks, ragged last chunks, and the DFT flush with an odd cell
    count (one cell per lane) and an even one (cell pairs), give the oracle's
    bits."""
    monkeypatch.setenv("MITTET_FDTD_LKB", lkb)
    ix, iy, iz = dims
    it = 50
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(ix=ix, iy=iy, iz=iz, seed=23, it=it,
                                                            ring=8)
    shot.run()
    w = oracle.dft_weights(it, dt, OM0, om)
    tr, acc = S.forward(0, signal, tuple(c + 1 for c in src), it - 1, rec1=rec + 1, it=it, weights=w)
    np.testing.assert_array_equal(shot.traces()[:3], tr)
    for q, nm in enumerate(oracle.FIELD_NAMES):
        np.testing.assert_array_equal(shot.field(q), S.arr(nm), err_msg=nm)
    np.testing.assert_array_equal(shot.volume_dft().reshape(3, len(om), -1), acc)


def test_zmarch_kernels_bitwise():
    """The z-marching kernel family (zchunk > 0) gives the same bits."""
    it = 60
    m, shot, S, sig, signal, src, rec, om, dt = setup_pair(seed=22, it=it, ring=8)
    shot.close()
    shot = m.shot(it, signal, src, 0, rec, om, ring=8, zchunk=5)
    shot.run()
    w = oracle.dft_weights(it, dt, OM0, om)
    tr, acc = S.forward(0, signal, tuple(c + 1 for c in src), it - 1, rec1=rec + 1, it=it, weights=w)
    np.testing.assert_array_equal(shot.traces()[:3], tr)
    np.testing.assert_array_equal(shot.volume_dft().reshape(3, len(om), -1), acc)
