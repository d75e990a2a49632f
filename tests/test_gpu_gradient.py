"""Exact discrete adjoint: gradient of the receiver misfit vs central finite
differences of the forward model (h = 1e-6 in log sigma) at 16 seeded random
cells, through the C ABI; and the torch.autograd wrapper returns the same
gradient.  The forward model itself is pinned bit for bit to the oracle
(test_gpu_parity.py), so the FD check pins the adjoint."""
import numpy as np
import pytest

from mittet_main_amd import Grid
from mittet_main_amd.adjoint import GradientProblem, ShotSpec
from mittet_main_amd.configs import dt_for
from mittet_main_amd.fwi3d import firstderiv_gauss

pytestmark = pytest.mark.gpu

OM0 = 2 * np.pi


def problem(seed=0, ix=26, iy=24, iz=22, nc=5, it=200, segment=0, levels=1, budget=None):
    rng = np.random.default_rng(seed)
    sig_true = np.exp(rng.uniform(np.log(0.1), np.log(2.0), (iz, iy, ix)))
    sig0 = np.full((iz, iy, ix), float(np.exp(np.mean(np.log(sig_true)))))
    sig0[: iz // 2] *= 0.7
    dx, dy, dz = 50.0, 45.0, 40.0
    dt, cmax = dt_for(dx, dy, dz, sig_min=0.1)
    g = Grid(ix, iy, iz, nc, dx, dy, dz, dt, OM0, cmax)
    sigw = firstderiv_gauss(it, dt, 6.0)
    shots = [ShotSpec((ix // 2, iy // 2, iz // 2), sigw),
             ShotSpec((ix // 2 - 4, iy // 2 + 3, iz // 2 - 2), sigw, src_comp=1)]
    rec = np.stack([rng.integers(nc, ix - nc, 8), rng.integers(nc, iy - nc, 8),
                    rng.integers(nc, iz - nc, 8)], axis=1)
    omegas = 2 * np.pi * np.array([1.0, 2.5, 4.0])
    P = GradientProblem(g, shots, rec, it, omegas, segment=segment, levels=levels,
                        budget_bytes=budget)
    P.d_obs = P.forward_data(sig_true)
    return P, sig0, rng


def test_gradient_matches_central_fd():
    P, sig0, rng = problem()
    logs = np.log(sig0)
    J, grad = P.value_and_grad(logs)
    assert J > 0 and np.all(np.isfinite(grad))
    iz, iy, ix = sig0.shape
    cells = [tuple(int(v) for v in rng.integers(0, (iz, iy, ix))) for _ in range(16)]
    src = P.specs[0].src
    cells.append((src[2], src[1], src[0]))
    h = 1e-6
    gmax = np.abs(grad).max()
    for (k, j, i) in cells:
        lp = logs.copy()
        lp[k, j, i] += h
        jp = P.misfit(np.exp(lp))
        lp[k, j, i] -= 2 * h
        jm = P.misfit(np.exp(lp))
        fd = (jp - jm) / (2 * h)
        g = grad[k, j, i]
        # relative error against the gradient's own scale at that cell, floored
        # at 1e-3 of the largest component (roundoff of J(+h) - J(-h) swamps
        # cells whose gradient is smaller than that)
        err = abs(fd - g) / max(abs(g), 1e-3 * gmax)
        assert err <= 1e-6, ((k, j, i), g, fd, err)
    P.close()


def test_lds_adjoint_matches_the_per_cell_adjoint(monkeypatch):
    """The adjoint's LDS kernels over the CPML-free core (lE update, E-phase
    gradient, lH update) and the per-cell passes everywhere
    (MITTET_FDTD_NOLDS=1) give the same gradient up to rounding.  90 x 40 x 40
    with a 6-cell CPML puts 2 x 4 tiles of 64 x 8 (partial on both axes) and
    several z-chunks in the core; both sources lie inside it."""
    def run():
        P, sig0, _ = problem(seed=7, ix=90, iy=40, iz=40, nc=6, it=120)
        J, g = P.value_and_grad(np.log(sig0))
        P.close()
        return J, g

    print("LDS adjoint vs per-cell adjoint: 90 x 40 x 40, ncpml 6", flush=True)
    J1, g1 = run()
    monkeypatch.setenv("MITTET_FDTD_NOLDS", "1")
    J0, g0 = run()
    assert J1 == J0                      # the forward is bitwise either way
    assert np.all(np.isfinite(g1)) and np.abs(g0).max() > 0
    assert np.abs(g1 - g0).max() <= 1e-12 * np.abs(g0).max()


def test_segment_length_does_not_change_gradient():
    P1, sig0, _ = problem(seed=3, segment=7)
    P2, _, _ = problem(seed=3, segment=64)
    J1, g1 = P1.value_and_grad(np.log(sig0))
    J2, g2 = P2.value_and_grad(np.log(sig0))
    assert J1 == J2
    np.testing.assert_array_equal(g1, g2)
    P1.close()
    P2.close()


def test_odd_cell_count_gradient():
    """27 x 25 x 23 cells (an odd count): the H and psi parts of every stored
    state start off 16-byte alignment, so the state store copies them as
    doubles, and the 256-byte padding between states is not a no-op.  Segment
    length and recursion depth still give the same gradient bits, and central
    FD still agrees."""
    print("odd cell count: 27 x 25 x 23, segment 7 vs 2 levels, FD at 6 cells", flush=True)
    P1, sig0, rng = problem(seed=11, ix=27, iy=25, iz=23, segment=7)
    P2, _, _ = problem(seed=11, ix=27, iy=25, iz=23, levels=2)
    logs = np.log(sig0)
    J1, g1 = P1.value_and_grad(logs)
    J2, g2 = P2.value_and_grad(logs)
    assert J1 == J2
    np.testing.assert_array_equal(g1, g2)
    P2.close()
    iz, iy, ix = sig0.shape
    h = 1e-6
    gmax = np.abs(g1).max()
    for _ in range(6):
        k, j, i = (int(v) for v in rng.integers(0, (iz, iy, ix)))
        lp = logs.copy()
        lp[k, j, i] += h
        jp = P1.misfit(np.exp(lp))
        lp[k, j, i] -= 2 * h
        jm = P1.misfit(np.exp(lp))
        fd = (jp - jm) / (2 * h)
        g = g1[k, j, i]
        err = abs(fd - g) / max(abs(g), 1e-3 * gmax)
        assert err <= 1e-6, ((k, j, i), g, fd, err)
    P1.close()


def test_gradient_with_the_x_pair_forward_family(monkeypatch):
    """The replays of the LDS and per-cell forward families write E, H and
    psi straight into the stored states; other families update H and psi in
    place and copy them into each state.  With the x-pair family
    (MITTET_FDTD_PAIR=1) the gradient is bitwise that of the default family:
    the forward is bitwise the same and the adjoint kernels are shared."""
    print("gradient: x-pair forward family vs default, 26 x 24 x 22", flush=True)
    P1, sig0, _ = problem(seed=9, segment=9)
    J1, g1 = P1.value_and_grad(np.log(sig0))
    P1.close()
    monkeypatch.setenv("MITTET_FDTD_PAIR", "1")
    P2, _, _ = problem(seed=9, segment=9)
    J2, g2 = P2.value_and_grad(np.log(sig0))
    P2.close()
    assert J1 == J2
    np.testing.assert_array_equal(g1, g2)


def test_gradient_with_non_temporal_forward_kernels(monkeypatch):
    """Grids whose steps outgrow the Infinity Cache run the forward and its
    replays with non-temporal field loads and stores (MITTET_FDTD_NT=1 forces
    them here).  A cache hint only: misfit and gradient are bitwise those of
    the cached kernels, with the replays writing into the stored states."""
    print("gradient: non-temporal forward kernels vs cached, 26 x 24 x 22", flush=True)
    monkeypatch.setenv("MITTET_FDTD_NT", "0")
    P1, sig0, _ = problem(seed=13, segment=9)
    J1, g1 = P1.value_and_grad(np.log(sig0))
    P1.close()
    monkeypatch.setenv("MITTET_FDTD_NT", "1")
    P2, _, _ = problem(seed=13, segment=9)
    J2, g2 = P2.value_and_grad(np.log(sig0))
    P2.close()
    assert J1 == J2
    np.testing.assert_array_equal(g1, g2)


@pytest.mark.parametrize("levels", [2, 3, 4])
def test_recursive_checkpoints_give_the_same_gradient(levels):
    """Recursive recomputation replays the same forward states, so misfit and
    gradient are bitwise those of the sqrt(Nt) scheme, with less storage.
    it = 200 (199 steps) leaves partial segments at every depth."""
    P1, sig0, _ = problem(seed=4)
    P2, _, _ = problem(seed=4, levels=levels)
    print(f"recursive checkpoints: levels {levels}", flush=True)
    J1, g1 = P1.value_and_grad(np.log(sig0))
    J2, g2 = P2.value_and_grad(np.log(sig0))
    assert J1 == J2
    np.testing.assert_array_equal(g1, g2)
    assert P2.storage_bytes() < P1.storage_bytes()
    P1.close()
    P2.close()


def test_budget_selects_the_fewest_levels_that_fit():
    from mittet_main_amd.adjoint import levels_for_budget
    import mittet_main_amd._lib as lib

    P, sig0, _ = problem(seed=6)
    L = lib.load()
    sh = P.shots[0]
    b = [int(L.fwi3d_grad_plan_bytes(sh._h, lv)) for lv in (1, 2, 3)]
    assert b[0] > b[1] > b[2] > 0
    assert levels_for_budget(sh, b[0]) == 1
    assert levels_for_budget(sh, b[0] - 1) == 2
    assert levels_for_budget(sh, b[2]) == 3
    with pytest.raises(ValueError):
        levels_for_budget(sh, 1e3)
    P.budget_bytes = b[1]
    J2, g2 = P.value_and_grad(np.log(sig0))
    assert P.last_levels == 2 and P.storage_bytes() == b[1]
    P.budget_bytes = None
    J1, g1 = P.value_and_grad(np.log(sig0))
    assert J1 == J2
    np.testing.assert_array_equal(g1, g2)
    P.close()


def test_autograd_function_returns_adjoint_gradient():
    torch = pytest.importorskip("torch")
    from mittet_main_amd.adjoint import misfit

    P, sig0, _ = problem(seed=5)
    logs = torch.tensor(np.log(sig0), dtype=torch.float64, requires_grad=True)
    J = misfit(logs, P)
    J.backward()
    J_ref, g_ref = P.value_and_grad(np.log(sig0))
    assert float(J) == J_ref
    np.testing.assert_array_equal(logs.grad.numpy(), g_ref)
    P.close()
