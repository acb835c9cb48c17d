"""GPU parity of the Bloch (thermal / ground state) and random-collision
propagators against the CPU oracle. FP64 bar: relative L2 <= 1e-12."""
import ctypes

import numpy as np
import pytest

from oracle import cexpr
from oracle import fw_oracle as fo

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


HO = dict(V="0.5*X*X", K="0.5*P*P", diff_V="X", diff_K="P")
HO_NP = dict(V=lambda X, t: 0.5 * X * X, K=lambda P, t: 0.5 * P * P)


def test_gibbs_cuda_convention(gpu):
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    b = WignerBlochHIP1D(X_gridDIM=256, X_amplitude=10., P_gridDIM=256, P_amplitude=10., kT=0.5, dbeta=0.005,
                         t=0., t_initial=0., **HO)
    W = b.get_gibbs_state().get()
    o = fo.BlochOracle(256, 10., 256, 10., 0.005, kT=0.5, dbeta=0.005, convention="cuda", **HO_NP)
    Wo = o.get_gibbs_state()
    assert b.gibbs_steps == 400          # n = 1/(kT dbeta)
    assert rel_l2(W, Wo) <= 1e-12
    assert rel_l2(W, fo.harmonic_gibbs(o, 1.0, 0.5)) <= 2e-6
    assert abs(b.get_V_min() - o.V_min) <= 1e-15 and abs(b.get_K_min() - o.K_min) <= 1e-15


def test_gibbs_fftw_convention_with_absorbing_mask(gpu):
    """FFTW exponents (2x the CUDA ones) with the alpha absorbing boundary
    (wigner_moyal_fftw_1d.py:135-156) in the V factor."""
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    b = WignerBlochHIP1D(X_gridDIM=128, X_amplitude=8., P_gridDIM=128, P_amplitude=8., kT=0.5, dbeta=0.01,
                         t=0., alpha=2.0, bloch_convention="fftw", **HO)
    W = b.get_gibbs_state().get()
    mask = cexpr.make(b.abs_boundary_x_theta, {}, ("X", "Theta", "t"), cexpr.x_theta_names)
    o = fo.BlochOracle(128, 8., 128, 8., 0.01, kT=0.5, dbeta=0.01, convention="fftw", mask_x_theta=mask, **HO_NP)
    ref_mask, _, _ = fo.absorbing_alpha_mask(o, 2.0, 0.01)
    assert np.max(np.abs(mask(o.X, o.Theta, 0.) - ref_mask)) <= 1e-14
    Wo = o.get_gibbs_state()
    assert rel_l2(W, Wo) <= 1e-12


def test_device_purity_guard_stops_at_the_right_step(gpu):
    """The guard runs on the device: with threshold 0.5 the run stops right after
    the first step whose purity reaches 0.5, like the reference's assert
    (wigner_bloch_cuda_1d.py:149-163) does at 1."""
    from fastwigner_amd import _lib
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    b = WignerBlochHIP1D(X_gridDIM=128, X_amplitude=10., P_gridDIM=128, P_amplitude=10., kT=0.05, dbeta=0.01,
                         t=0., **HO)
    b.set_wignerfunction(1. / (128 * 128))
    series = np.empty((2000, 9))
    done = ctypes.c_int()
    _lib.call("fw_wigner_bloch_propagate", b._handle, 0.01, 2000, 0.5, _lib.dptr(series), ctypes.byref(done))
    o = fo.BlochOracle(128, 10., 128, 10., 0.01, kT=0.05, dbeta=0.01, **HO_NP)
    o.set_wignerfunction(np.full((128, 128), 1.0))
    k = 0
    while True:
        o.bloch_single_step_propagation(0.01)
        k += 1
        if not o.get_purity() < 0.5:
            break
    assert done.value == k
    assert rel_l2(b.wignerfunction.get(), o.W) <= 1e-12
    pur = 2 * np.pi * series[:k, 8] / (series[:k, 0] ** 2 * o.dXdP)
    assert pur[k - 1] >= 0.5 and np.all(pur[:k - 1] < 0.5)


def test_ground_state_with_rollback(gpu):
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    b = WignerBlochHIP1D(X_gridDIM=128, X_amplitude=8., P_gridDIM=128, P_amplitude=8., dbeta=0.1, t=0., **HO)
    W = b.get_ground_state(abs_tol_purity=1e-10).get()
    exact = np.exp(-(b.X ** 2) - b.P ** 2)
    exact /= exact.sum() * b.dXdP
    assert rel_l2(W, exact) <= 2e-3
    assert abs(b.get_average("0.5*P*P + 0.5*X*X") - 0.5) <= 1e-3
    assert abs(b.get_purity() - 1.0) <= 1e-9


@pytest.fixture(scope="module")
def double_well_gibbs_oracle():
    o = fo.BlochOracle(512, 10., 512, 10., 0.005, kT=0.2, dbeta=0.005, V=cexpr.make("-0.5*X*X + X*X*X*X/20."),
                       K=cexpr.make("0.5*P*P", args=("P", "t")))
    return o.get_gibbs_state()


@pytest.mark.parametrize("precision,bar", [("fp64", 1e-12), ("fp32", 1e-4)])
def test_double_well_gibbs_512(gpu, double_well_gibbs_oracle, precision, bar):
    """Double-well Bloch state: 512^2, x,p in [-10,10), V = -x^2/2 + x^4/20,
    kT = 0.2, dbeta = 0.005 -> 1000 steps (CUDA convention, state at kT), in
    FP64 and in the FP32 form (W and FFTs in FP32, exponents and sums in FP64),
    against the FP64 oracle."""
    from fastwigner_amd.wigner_bloch import WignerBlochHIP1D
    V, K = "-0.5*X*X + X*X*X*X/20.", "0.5*P*P"
    b = WignerBlochHIP1D(X_gridDIM=512, X_amplitude=10., P_gridDIM=512, P_amplitude=10., kT=0.2, dbeta=0.005,
                         t=0., V=V, K=K, precision=precision)
    W = b.get_gibbs_state().get()
    err = rel_l2(W, double_well_gibbs_oracle)
    print("double well Bloch 512^2 %s rel L2 vs FP64 oracle: %.2e" % (precision, err))
    assert err <= bar


def make_rc(gamma, n=128, **kw):
    from fastwigner_amd.random_collision import WignerRandomCollisionHIP1D
    V, K = "0.5*X*X + 0.05*X*X*X*X", "0.5*P*P"
    g = WignerRandomCollisionHIP1D(X_gridDIM=n, X_amplitude=8., P_gridDIM=n, P_amplitude=8., dt=0.01, t=0.,
                                   gamma=gamma, kT=0.5, dbeta=0.01, V=V, K=K, diff_V="X + 0.2*X*X*X",
                                   diff_K="P", **kw)
    o = fo.RandomCollisionOracle(n, 8., n, 8., 0.01, gamma=gamma, kT=0.5, dbeta=0.01, V=cexpr.make(V),
                                 K=cexpr.make(K, args=("P", "t")), diff_V=cexpr.make("X + 0.2*X*X*X"),
                                 diff_K=cexpr.make("P", args=("P", "t")), phase_time_shift=0.0)
    return g, o


def test_random_collision_parity(gpu):
    g, o = make_rc(0.05)
    assert rel_l2(g.gibbs_state, o.gibbs) <= 1e-12
    W0 = np.exp(-(o.X - 1.) ** 2 - (o.P - 0.5) ** 2)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(300).get()
    o.propagate(300)
    assert rel_l2(Wg, o.W) <= 1e-12
    for name in ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average"):
        ref = np.array([s[name] for s in o.series])
        assert np.max(np.abs(np.array(getattr(g, name)) - ref)) <= 1e-11, name


def test_random_collision_single_step_on_an_unnormalised_state(gpu):
    """single_step_propagation with the dissipator
    (wigner_random_collision_model_fftw_1d.py:56-75): W <- w W + W0s around the
    unitary step, W left unnormalised. The dissipator is affine, so a state of
    norm 2 (set without normalisation) must go through it as it is, in the
    single step and in the propagate that follows."""
    from fastwigner_amd import _lib
    g, o = make_rc(0.05)
    W0 = np.exp(-(o.X - 1.) ** 2 - (o.P - 0.5) ** 2)
    W0 = 2.0 * W0 / (W0.sum() * o.dXdP)
    _lib.call("fw_wigner_set_state", g._handle, _lib.dptr(np.ascontiguousarray(W0)))
    o.W = W0.copy()
    Wg = g.single_step_propagation().get()
    o.single_step_propagation()
    assert rel_l2(Wg, o.W) <= 1e-12
    norm = Wg.sum() * o.dXdP
    assert abs(norm - o.W.sum() * o.dXdP) <= 1e-12 and norm > 1.5
    Wg = g.propagate(50).get()
    o.propagate(50)
    assert rel_l2(Wg, o.W) <= 1e-12


def test_random_collision_gamma0_is_moyal(gpu):
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    g, o = make_rc(0.0)
    m = WignerMoyalHIP1D(X_gridDIM=128, X_amplitude=8., P_gridDIM=128, P_amplitude=8., dt=0.01, t=0.,
                         V="0.5*X*X + 0.05*X*X*X*X", K="0.5*P*P", phase_time_shift=0.0)
    W0 = np.exp(-(o.X - 1.) ** 2 - (o.P - 0.5) ** 2)
    g.set_wignerfunction(W0)
    m.set_wignerfunction(W0)
    # separately compiled kernels (the dissipator set) may round one ulp differently
    a, b = g.propagate(100).get(), m.propagate(100).get()
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))


def test_random_collision_driven_split_potential(gpu):
    """A driven potential V0(X) + X E0 sin(om t), which splits into a table of V0
    and a per-step field factor, under the dissipator (three-pass steps reading
    the field values filled per chunk), 300 steps against the oracle."""
    from fastwigner_amd.random_collision import WignerRandomCollisionHIP1D
    consts = {"E0": 0.05, "om": 0.3}
    V, K = "0.5*X*X + 0.05*X*X*X*X + X*E0*sin(om*t)", "0.5*P*P"
    dV = "X + 0.2*X*X*X + E0*sin(om*t)"
    n = 128
    g = WignerRandomCollisionHIP1D(X_gridDIM=n, X_amplitude=8., P_gridDIM=n, P_amplitude=8., dt=0.01, t=0.,
                                   gamma=0.05, kT=0.5, dbeta=0.01, V=V, K=K, diff_V=dV, diff_K="P", **consts)
    o = fo.RandomCollisionOracle(n, 8., n, 8., 0.01, gamma=0.05, kT=0.5, dbeta=0.01, V=cexpr.make(V, consts),
                                 K=cexpr.make(K, args=("P", "t")), diff_V=cexpr.make(dV, consts),
                                 diff_K=cexpr.make("P", args=("P", "t")), phase_time_shift=0.0)
    assert rel_l2(g.gibbs_state, o.gibbs) <= 1e-12
    W0 = np.exp(-(o.X - 1.) ** 2 - (o.P - 0.5) ** 2)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    Wg = g.propagate(300).get()
    o.propagate(300)
    assert rel_l2(Wg, o.W) <= 1e-12
    for name in ("X_average", "P_average", "P_average_RHS", "hamiltonian_average"):
        ref = np.array([s[name] for s in o.series])
        assert np.max(np.abs(np.array(getattr(g, name)) - ref)) <= 1e-11, name


def test_dissipator_cleared_between_graph_runs(gpu):
    """Three-pass steps with the dissipator for a full graph chunk (64 steps),
    then fw_wigner_set_dissipator(NULL) and 64 more steps: the second run must
    be plain Moyal (the graph captured with the dissipative kernels is not
    replayed), against the oracle doing the same."""
    from fastwigner_amd import _lib
    from fastwigner_amd.wigner_moyal import WignerMoyalHIP1D
    n, gamma = 128, 0.05
    kw = dict(V="0.5*X*X + 0.05*X*X*X*X", K="0.5*P*P")
    g = WignerMoyalHIP1D(X_gridDIM=n, X_amplitude=8., P_gridDIM=n, P_amplitude=8., dt=0.01, t=0.,
                         phase_time_shift=0.0, step_passes=3, **kw)
    o = fo.RandomCollisionOracle(n, 8., n, 8., 0.01, gamma=gamma, kT=0.5, dbeta=0.01, phase_time_shift=0.0,
                                 **{k: cexpr.make(v, args=(("X", "t") if k == "V" else ("P", "t")))
                                    for k, v in kw.items()})
    w0s = np.ascontiguousarray(o.scaled_gibbs_state)
    _lib.call("fw_wigner_set_dissipator", g._handle, _lib.dptr(w0s), float(o.w_coeff))
    W0 = np.exp(-(o.X - 1.) ** 2 - (o.P - 0.5) ** 2)
    g.set_wignerfunction(W0)
    o.set_wignerfunction(W0)
    g.propagate(64)
    o.propagate(64)
    assert rel_l2(g.wignerfunction.get(), o.W) <= 1e-12
    _lib.call("fw_wigner_set_dissipator", g._handle, None, 1.0)
    m = fo.MoyalOracle(n, 8., n, 8., 0.01, phase_time_shift=0.0, t=o.t,
                       **{k: cexpr.make(v, args=(("X", "t") if k == "V" else ("P", "t"))) for k, v in kw.items()})
    m.set_wignerfunction(o.W)
    Wg = g.propagate(64).get()
    m.propagate(64)
    assert rel_l2(Wg, m.W) <= 1e-12
