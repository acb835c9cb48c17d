"""GPU parity of the density-matrix propagators (von Neumann, Bloch), their
observables and the Wigner transform against the CPU oracle.
FP64 bar: relative L2 <= 1e-12; FP32 form: <= 1e-4."""
import numpy as np
import pytest

from oracle import cexpr
from oracle import fw_oracle as fo

pytestmark = pytest.mark.gpu

VQ = "0.5*X*X + 0.05*X*X*X*X"
KQ = "0.5*P*P"


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def gaussian_rho(x, x0=1.0, p0=0.5, s=1.0):
    psi = np.exp(-(x - x0) ** 2 / (2 * s * s) + 1j * p0 * x)
    return np.outer(psi.conj(), psi)


def make(n=256, amp=10., dt=0.01, V=VQ, K=KQ, dV="X + 0.2*X*X*X", dK="P", consts=None, precision="fp64", **kw):
    from fastwigner_amd.rho_vneumann import RhoVNeumannHIP1D
    cv = dict(consts or {})
    g = RhoVNeumannHIP1D(X_gridDIM=n, X_amplitude=amp, dt=dt, t=0., V=V, K=K, diff_V=dV, diff_K=dK,
                         precision=precision, **cv, **kw)
    cx = dict(cv, X_amplitude=amp, dt=dt)
    o = fo.RhoOracle(n, amp, dt, V=cexpr.make(V, cx), K=cexpr.make(K, cx, ("P", "t")),
                     diff_V=cexpr.make(dV, cx), diff_K=cexpr.make(dK, cx, ("P", "t")),
                     mask_x=cexpr.make(kw.get("abs_boundary_x"), cx, ("X", "X_prime", "t")),
                     mask_p=cexpr.make(kw.get("abs_boundary_p"), cx, ("P", "P_prime", "t")))
    return g, o


@pytest.mark.parametrize("K,dK", [(KQ, "P"), ("0.5*P*P + 0.05*cos(P)", "P - 0.05*sin(P)"),
                                  ("0.5*P*P + 0.05*P*sin(0.3*t)", "P + 0.05*sin(0.3*t)")])
def test_vneumann_parity_and_ehrenfest(gpu, K, dK):
    """State and every Ehrenfest series against the oracle. The momentum family
    goes through precomputed weights when K and K' do not depend on t (the
    first two cases; the second is not a polynomial) and through the diagonal
    group sums and one inverse FFT otherwise (the third)."""
    g, o = make(K=K, dK=dK)
    r0 = gaussian_rho(o.Xv)
    g.set_rho(r0)
    o.set_rho(r0)
    n = 150
    g.propagate(n)
    xs, ps, dks, dvs, hs = [], [], [], [], []
    for _ in range(n):
        o.propagate(1)
        xs.append(o.get_average((lambda X, t: X,)))
        ps.append(o.get_average((None, lambda P, t: P)))
        dks.append(o.get_average((None, o.diff_K)))
        dvs.append(-o.get_average((o.diff_V,)))
        hs.append(o.get_average((None, o.K)) + o.get_average((o.V,)))
    err = rel_l2(g.rho.get(), o.rho)
    print("vN 256^2 rel L2 vs oracle:", err)
    assert err <= 1e-12
    for name, ref in (("X_average", xs), ("P_average", ps), ("X_average_RHS", dks), ("P_average_RHS", dvs),
                      ("hamiltonian_average", hs)):
        assert np.max(np.abs(np.array(getattr(g, name)) - np.array(ref))) <= 1e-11, name
    assert abs(g.get_purity() - o.get_purity()) <= 1e-12


def test_vneumann_masks_and_time_dependent_field(gpu):
    """Strong-field pattern: V with X*E(t) from `functions`, absorbing boundary
    in x and x' (strong_field_physics_cuda_1d.py:49-82)."""
    functions = "__device__ double E(double t)\n{{\n  return -F * sin(omega * t) * pow(sin(omega * t / 16.), 2);\n}}\n"
    F, om = 0.04, 0.5
    mask = "pow(fabs(sin(0.5*M_PI*(X + X_amplitude)/X_amplitude)*sin(0.5*M_PI*(X_prime + X_amplitude)/X_amplitude)), dt*0.05)"
    from fastwigner_amd.rho_vneumann import RhoVNeumannHIP1D
    g = RhoVNeumannHIP1D(X_gridDIM=128, X_amplitude=15., dt=0.05, t=0., F=F, omega=om, functions=functions,
                         V="-1. / sqrt(X * X + 1.37) + X * E(t)", K="0.5 * P * P", abs_boundary_x=mask)
    E = lambda t: -F * np.sin(om * t) * np.sin(om * t / 16.) ** 2
    cx = dict(X_amplitude=15., dt=0.05)
    o = fo.RhoOracle(128, 15., 0.05, V=lambda X, t: -1. / np.sqrt(X * X + 1.37) + X * E(t),
                     K=lambda P, t: 0.5 * P * P, mask_x=cexpr.make(mask, cx, ("X", "X_prime", "t")))
    r0 = gaussian_rho(o.Xv, 0.5, 0.2, 1.5)
    g.set_rho(r0)
    o.set_rho(r0)
    g.propagate(100)
    o.propagate(100)
    assert rel_l2(g.rho.get(), o.rho) <= 1e-12


def test_get_average_tuples(gpu):
    g, o = make(n=128)
    r0 = gaussian_rho(o.Xv, -0.5, 1.0, 0.8)
    g.set_rho(r0)
    o.set_rho(r0)
    g.propagate(20)
    o.propagate(20)
    X = lambda X, t: X
    P = lambda P, t: P
    for obs_s, obs_f in ((("X",), (X,)), ((None, "P * P"), (None, lambda P, t: P * P)),
                         (("X", "P"), (X, P)), ((None, "P", "X"), (None, P, X)),
                         (("X * X", None), (lambda X, t: X * X, None))):
        assert abs(g.get_average(obs_s) - o.get_average(obs_f)) <= 1e-12, obs_s
    assert abs(g.get_sigma_x_sigma_p() - o.get_sigma_x_sigma_p()) <= 1e-11


def test_wigner_transform_and_time(gpu):
    g, o = make(n=256)
    r0 = gaussian_rho(o.Xv, 1.0, 0.5, 1.0)
    g.set_rho(r0)
    o.set_rho(r0)
    W = g.get_wignerfunction().get()
    Wo = o.get_wignerfunction()
    assert rel_l2(W, Wo) <= 1e-12
    Wa = np.exp(-(g.X_wigner - 1.0) ** 2 - (g.P_wigner - 0.5) ** 2) / np.pi
    Wa /= Wa.sum() * g.wigner_dxdp
    assert rel_l2(W.real, Wa) <= 1e-12
    # Wigner time after some propagation, with the reference's indicator
    from fastwigner_amd.rho_vneumann import RhoVNeumannHIP1D
    g2 = RhoVNeumannHIP1D(X_gridDIM=256, X_amplitude=10., dt=0.01, t=0., V=VQ, K=KQ, Ip=0.6)
    g2.set_rho(r0)
    w_init = g2.get_wignerfunction().get().real
    g2.propagate(30)
    o.propagate(30)
    wc = g2.get_wignerfunction()
    got = g2.get_wigner_time(wc, w_init, g2.t)
    ref = o.get_wigner_time(o.get_wignerfunction(), w_init, o.t, o.K, o.V, 0.6)
    assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))


def test_rho_bloch_gibbs(gpu):
    from fastwigner_amd.rho_bloch import RhoBlochHIP1D
    V, K = "0.5*X*X", "0.5*P*P"
    g = RhoBlochHIP1D(X_gridDIM=128, X_amplitude=10., kT=0.5, dbeta=0.005, t=0., V=V, K=K)
    rho = g.get_gibbs_state().get()
    o = fo.RhoBlochOracle(128, 10., 0.005, kT=0.5, dbeta=0.005, V=cexpr.make(V), K=cexpr.make(K, args=("P", "t")))
    ro = o.get_gibbs_state()
    assert g.gibbs_steps == 400
    assert rel_l2(rho, ro) <= 1e-12
    assert abs(g.get_V_min() - o.V_min) == 0.0 and g.get_K_min() == 0.0
    W = g.get_wignerfunction().get().real
    th = np.tanh(1.0)
    Wa = th / np.pi * np.exp(-2 * th * 0.5 * (g.P_wigner ** 2 + g.X_wigner ** 2))
    assert np.max(np.abs(W - Wa)) * g.wigner_dxdp <= 1e-8


def test_rho_ground_state(gpu):
    from fastwigner_amd.rho_bloch import RhoBlochHIP1D
    g = RhoBlochHIP1D(X_gridDIM=128, X_amplitude=8., dbeta=0.05, t=0., V="0.5*X*X", K="0.5*P*P")
    g.get_ground_state(abs_tol_purity=1e-9)
    assert abs(g.get_purity() - 1.0) <= 1e-8
    assert abs(g.get_average((None, "0.5*P*P")) + g.get_average(("0.5*X*X",)) - 0.5) <= 1e-3


def test_double_well_density_matrix_1024(gpu):
    """Double-well density matrix at its full size: the thermal state of the
    double well at 1024^2 (x in [-10, 10)) from rho = I, then 2000 von Neumann
    steps with dt = 0.005. Checked by size-independent properties: trace,
    hermiticity, purity constant, energy conserved (the thermal state is
    stationary), normalisation of its Wigner function."""
    from fastwigner_amd.rho_bloch import RhoBlochHIP1D
    from fastwigner_amd.rho_vneumann import RhoVNeumannHIP1D
    V, K = "-0.5*X*X + X*X*X*X/20.", "0.5*P*P"
    b = RhoBlochHIP1D(X_gridDIM=1024, X_amplitude=10., kT=0.2, dbeta=0.005, t=0., V=V, K=K)
    rho_th = b.get_gibbs_state().get()
    Wb = b.get_wignerfunction().get().real
    vn = RhoVNeumannHIP1D(X_gridDIM=1024, X_amplitude=10., dt=0.005, t=0., V=V, K=K, diff_V="-X + X*X*X/5.",
                          diff_K="P")
    vn.set_rho(rho_th)
    vn.propagate(2000)
    r = vn.rho.get()
    assert abs(np.trace(r) * vn.dX - 1.0) <= 1e-12
    assert np.max(np.abs(r - r.conj().T)) <= 1e-12 * np.max(np.abs(r))
    pur = np.array(vn.purity_series)
    assert np.max(np.abs(pur - pur[0])) <= 1e-10
    H = np.array(vn.hamiltonian_average)
    print("double well vN 1024^2: energy drift over 2000 steps", np.max(np.abs(H - H[0])) / abs(H[0]))
    assert np.max(np.abs(H - H[0])) <= 1e-6 * abs(H[0])    # a thermal state is stationary
    assert abs(np.sum(Wb) * b.wigner_dxdp - 1.0) <= 1e-12


def test_fp32_density_matrix(gpu):
    g, o = make(n=256, precision="fp32")
    r0 = gaussian_rho(o.Xv)
    g.set_rho(r0)
    o.set_rho(r0)
    g.propagate(500)
    o.propagate(500)
    err = rel_l2(g.rho.get(), o.rho)
    print("vN fp32 256^2 500 steps rel L2 vs fp64 oracle:", err)
    assert err <= 1e-4
