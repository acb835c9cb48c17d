"""GPU parity on even grids that are not powers of two (the reference's Wigner
classes only assert evenness, wigner_moyal_cuda_1d.py:57-62; its density-matrix
classes only check that X_gridDIM is given, rho_vneumann_cuda_1d.py:53-56).
Such lines take odd-radix and prime-factor stages in the FFT engine and the
spectral step through LDS; lines whose thread count has an odd factor run
blocks with a partial last wave.

Bars as for the power-of-two grids: FP64 relative L2 <= 1e-12 against the
oracle, series to 1e-11 (relative to their scale); FP32 form <= 1e-4 against
the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import cexpr
from oracle import fw_oracle as fo
from tests.test_gpu_moyal import DK_Q, DV_Q, K_Q, V_Q, check_series, gaussian, make_pair, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,npp,passes,steps", [
    (1000, 1000, 0, 100),   # E = 10: radix-10 stages, 100 threads per line (a partial second wave)
    (96, 120, 0, 300),      # 2^5 3 by 2^3 3 5 (E = 12 and 30)
    (768, 250, 3, 100),     # three passes per step; 250 = 2 5^3
    (112, 104, 0, 300),     # 2^4 7 by 2^3 13 (E = 14, 26)
    (30, 22, 0, 300),       # one radix-30 and one radix-22 stage per line
])
def test_quartic_parity_even_grids(gpu, nx, npp, passes, steps):
    g, o = make_pair(nx, 8., npp, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, step_passes=passes)
    W0 = gaussian(o)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(steps).get()
    o.propagate(steps)
    err = rel_l2(Wg, o.W)
    print("quartic %dx%d (%d passes), %d steps: rel L2 %.2e" % (nx, npp, passes or 2, steps, err))
    assert err <= 1e-12
    check_series(g, o)


@pytest.mark.parametrize("tiles", [(6, 2, 24, 1), (12, 1, 12, 4), (24, 4, 6, 2)])
def test_tile_schedules_agree_on_a_mixed_radix_grid(gpu, tiles):
    """192 = 2^6 3: every valid E (6, 12, 24) gives the oracle's answer."""
    g, o = make_pair(192, 8., 192, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, tiles=tiles)
    W0 = gaussian(o)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(100).get()
    o.propagate(100)
    assert rel_l2(Wg, o.W) <= 1e-12
    check_series(g, o)


@pytest.mark.parametrize("xp", ["x", "xp"])
def test_masks_and_field_on_an_even_grid(gpu, xp):
    """The three absorbing-boundary strings and a strong-field V(x, t) =
    x^2/2 + x E(t) (a table of x^2/2 times a per-step field factor) on a
    600 x 360 grid; an x-p mask of x alone keeps the fused column pass on the
    theta spectrum, one that depends on p sends it through x-p space."""
    consts = {"E0": 0.05, "om": 0.3}
    mxp = "pow(fabs(cos(0.5*M_PI*X/X_amplitude)), 0.125)"
    if xp == "xp":
        mxp += " * exp(-1e-3*P*P*P*P/(P_amplitude*P_amplitude))"
    masks = dict(abs_boundary_x_theta="pow(fabs(cos(0.5*M_PI*X_minus/X_amplitude)*cos(0.5*M_PI*X_plus/X_amplitude)), 0.01)",
                 abs_boundary_lambda_p="exp(-1e-4*Lambda*Lambda)",
                 abs_boundary_x_p=mxp)
    g, o = make_pair(600, 10., 360, 6., 0.02, "0.5*X*X + X*E0*sin(om*t)", "0.5*P*P",
                     "X + E0*sin(om*t)", "P", consts=consts, masks=masks)
    W0 = np.exp(-(o.X + 0.5) ** 2 - (o.P - 0.3) ** 2)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(150).get()
    o.propagate(150)
    assert rel_l2(Wg, o.W) <= 1e-12
    dip = np.array([s["P_average_RHS"] for s in o.series])
    assert np.max(np.abs(np.array(g.P_average_RHS) - dip)) <= 1e-11


def test_fp32_form_on_an_even_grid(gpu):
    """FP32 W and FFTs with the compensated odd-radix constants, FP64 phases
    and sums, against the FP64 oracle (500 steps at 250 x 300)."""
    g, o = make_pair(250, 8., 300, 8., 0.01, V_Q, K_Q, DV_Q, DK_Q, precision="fp32")
    W0 = gaussian(o)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(500).get()
    o.propagate(500)
    err = rel_l2(Wg, o.W)
    print("quartic 250x300 fp32, 500 steps: rel L2 vs fp64 oracle %.2e" % err)
    assert err <= 1e-4


def test_bloch_and_random_collision_on_an_even_grid(gpu):
    """The double-well Gibbs state (CUDA convention) on 300 x 300 and a
    random-collision run (FFTW-convention Gibbs source, three-pass dissipative
    steps) on 200 x 120, against the oracle."""
    from fastwigner_amd.random_collision import WignerRandomCollisionHIP1D
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    V, dV = "-0.5*X*X + X*X*X*X/20.", "-X + X*X*X/5."
    b = WignerBlochHIP1D(X_gridDIM=300, X_amplitude=10., P_gridDIM=300, P_amplitude=10., kT=0.5, dbeta=0.01,
                         t=0., V=V, K="0.5*P*P", diff_V=dV, diff_K="P")
    W = b.get_gibbs_state().get()
    ob = fo.BlochOracle(300, 10., 300, 10., 0.01, kT=0.5, dbeta=0.01, convention="cuda",
                        V=cexpr.make(V), K=cexpr.make("0.5*P*P", args=("P", "t")))
    assert rel_l2(W, ob.get_gibbs_state()) <= 1e-12

    Vr = "0.5*X*X + 0.05*X*X*X*X"
    g = WignerRandomCollisionHIP1D(X_gridDIM=200, X_amplitude=8., P_gridDIM=120, P_amplitude=8., dt=0.01, t=0.,
                                   gamma=0.05, kT=0.5, dbeta=0.01, V=Vr, K="0.5*P*P", diff_V="X + 0.2*X*X*X",
                                   diff_K="P")
    o = fo.RandomCollisionOracle(200, 8., 120, 8., 0.01, gamma=0.05, kT=0.5, dbeta=0.01, V=cexpr.make(Vr),
                                 K=cexpr.make("0.5*P*P", args=("P", "t")), diff_V=cexpr.make("X + 0.2*X*X*X"),
                                 diff_K=cexpr.make("P", args=("P", "t")), phase_time_shift=0.0)
    assert rel_l2(g.gibbs_state, o.gibbs) <= 1e-12
    W0 = np.exp(-(o.X - 1.) ** 2 - (o.P - 0.5) ** 2)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(200).get()
    o.propagate(200)
    assert rel_l2(Wg, o.W) <= 1e-12


# ---------------------------------------------------------------------------
# Density matrix on even grids that are not powers of two
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,K,dK,steps", [
    (200, "0.5*P*P", "P", 100),                                        # 2^3 5^2: momentum sums by weights
    (200, "0.5*P*P + 0.05*P*sin(0.3*t)", "P + 0.05*sin(0.3*t)", 60),  # the one-block inverse FFT form
    (1000, "0.5*P*P", "P", 30),                                        # 100-thread lines: partial waves
])
def test_density_matrix_parity_even_grids(gpu, n, K, dK, steps):
    from tests.test_gpu_rho import gaussian_rho, make
    g, o = make(n=n, K=K, dK=dK)
    r0 = gaussian_rho(o.Xv)
    g.set_rho(r0)
    o.set_rho(r0)
    g.propagate(steps)
    xs, ps, hs = [], [], []
    for _ in range(steps):
        o.propagate(1)
        xs.append(o.get_average((lambda X, t: X,)))
        ps.append(o.get_average((None, lambda P, t: P)))
        hs.append(o.get_average((None, o.K)) + o.get_average((o.V,)))
    err = rel_l2(g.rho.get(), o.rho)
    print("von Neumann %d^2, %d steps: rel L2 %.2e" % (n, steps, err))
    assert err <= 1e-12
    for name, ref in (("X_average", xs), ("P_average", ps), ("hamiltonian_average", hs)):
        assert np.max(np.abs(np.array(getattr(g, name)) - np.array(ref))) <= 1e-11, name


def test_density_matrix_wigner_fp32_and_bloch_even_grids(gpu):
    """The Wigner transform (three shears, sign-flip FFT) at 240 = 2^4 3 5
    against the oracle and the Gaussian's closed form; the FP32 form at 240
    against the FP64 oracle; the harmonic Gibbs state at 120 against the
    oracle."""
    from fastwigner_amd.rho_bloch import RhoBlochHIP1D
    from tests.test_gpu_rho import gaussian_rho, make
    g, o = make(n=240)
    r0 = gaussian_rho(o.Xv, 1.0, 0.5, 1.0)
    g.set_rho(r0)
    o.set_rho(r0)
    W = g.get_wignerfunction().get()
    assert rel_l2(W, o.get_wignerfunction()) <= 1e-12
    Wa = np.exp(-(g.X_wigner - 1.0) ** 2 - (g.P_wigner - 0.5) ** 2) / np.pi
    Wa /= Wa.sum() * g.wigner_dxdp
    assert rel_l2(W.real, Wa) <= 1e-12

    g32, o32 = make(n=240, precision="fp32")
    g32.set_rho(r0)
    o32.set_rho(r0)
    g32.propagate(200)
    o32.propagate(200)
    err = rel_l2(g32.rho.get(), o32.rho)
    print("von Neumann 240^2 fp32, 200 steps: rel L2 vs fp64 oracle %.2e" % err)
    assert err <= 1e-4

    V, K = "0.5*X*X + 0.05*X*X*X*X", "0.5*P*P"
    b = RhoBlochHIP1D(X_gridDIM=120, X_amplitude=10., kT=0.5, dbeta=0.005, t=0., V=V, K=K)
    rho = b.get_gibbs_state().get()
    ob = fo.RhoBlochOracle(120, 10., 0.005, kT=0.5, dbeta=0.005, V=cexpr.make(V), K=cexpr.make(K, args=("P", "t")))
    assert rel_l2(rho, ob.get_gibbs_state()) <= 1e-12
