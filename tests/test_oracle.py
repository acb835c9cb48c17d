"""Pins the CPU oracle (oracle/fw_oracle.py) before it is trusted as the checker.

Each check is chosen so a plausible slip (dropped factor, wrong sign, swapped
axis, transposed operand) fails it; ``test_planted_mistakes_are_caught`` plants
such slips in the oracle and shows the checks catch them.
"""
import numpy as np
import pytest
import scipy.fft as sfft

from oracle import dft
from oracle import fw_oracle as fo


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# --------------------------------------------------------------------------- transforms
@pytest.mark.parametrize("n", [8, 16, 32, 64])
def test_rfft_is_the_direct_dft(n):
    x = np.random.default_rng(n).standard_normal((n, 5))
    assert np.max(np.abs(sfft.rfft(x, axis=0) - dft.rdft(x, axis=0))) < 1e-12


@pytest.mark.parametrize("n", [8, 16, 64])
def test_irfft_drops_imaginary_dc_and_nyquist(n):
    rng = np.random.default_rng(n + 1)
    y = rng.standard_normal((n // 2 + 1, 3)) + 1j * rng.standard_normal((n // 2 + 1, 3))
    assert np.max(np.abs(sfft.irfft(y, n=n, axis=0) - dft.irdft(y, n, axis=0))) < 1e-14
    # the sign of the Nyquist frequency (FFTW: -N/2, CUDA: +N/2) does not matter
    z = y.copy()
    z[n // 2] = np.conj(z[n // 2])
    assert np.max(np.abs(sfft.irfft(y, n=n, axis=0) - sfft.irfft(z, n=n, axis=0))) == 0.0


# --------------------------------------------------------------------------- Moyal
def dense_moyal_step(o, W):
    """Independent formulation: full complex DFT matrices and the full-spectrum
    multiplier continued to negative frequencies by Hermitian symmetry; the
    Nyquist bin keeps the real part of the factor."""
    tau = o.t + o.shift * o.dt

    def v_full(k_idx, n):
        k = np.where(k_idx <= n // 2, k_idx, k_idx - n).astype(float)
        return k

    npp, nx = o.Np, o.Nx
    th = v_full(np.arange(npp), npp)[:, None] * o.dTheta
    ph_v = -0.5 * o.dt * (o.V(o.X - 0.5 * th, tau) - o.V(o.X + 0.5 * th, tau))
    fv = np.exp(1j * ph_v)
    fv[npp // 2] = fv[npp // 2].real
    lam = v_full(np.arange(nx), nx)[None, :] * o.dLambda
    ph_k = -o.dt * (o.K(o.P + 0.5 * lam, tau) - o.K(o.P - 0.5 * lam, tau))
    fk = np.exp(1j * ph_k)
    fk[:, nx // 2] = fk[:, nx // 2].real

    def vpass(W):
        return dft.dft(dft.dft(W, axis=0) * fv, axis=0, inverse=True).real

    W = vpass(W)
    W = dft.dft(dft.dft(W, axis=1) * fk, axis=1, inverse=True).real
    return vpass(W)


@pytest.mark.parametrize("nx,npp", [(8, 8), (16, 8), (16, 16)])
def test_moyal_step_against_dense_dft(nx, npp):
    V = lambda X, t: 0.3 * X * X + 0.05 * X ** 4 + 0.1 * X * np.sin(t)
    K = lambda P, t: 0.5 * P * P + 0.02 * P ** 4
    o = fo.MoyalOracle(nx, 4., npp, 3., 0.05, V=V, K=K, t=0.3)
    W = np.random.default_rng(7).random((npp, nx))
    o.W = W.copy()
    ref = dense_moyal_step(o, W)
    o.single_step_propagation()
    assert np.max(np.abs(o.W - ref)) < 1e-13


def leapfrog_gaussian(o, n, om2, x0=1.0, p0=0.5, var=0.5):
    a = 0.5 * o.dt * om2
    kick = np.array([[1, 0], [-a, 1]])
    M = kick @ np.array([[1, o.dt], [0, 1]]) @ kick
    Mn = np.linalg.matrix_power(M, n)
    mu = Mn @ np.array([x0, p0])
    S = Mn @ (var * np.eye(2)) @ Mn.T
    Si = np.linalg.inv(S)
    dx, dp = o.X - mu[0], o.P - mu[1]
    Wa = np.exp(-0.5 * (Si[0, 0] * dx * dx + 2 * Si[0, 1] * dx * dp + Si[1, 1] * dp * dp))
    return Wa / (Wa.sum() * o.dXdP)


def test_harmonic_leapfrog_closed_form():
    om2 = 1.3
    o = fo.MoyalOracle(256, 8., 256, 8., 0.01, V=lambda X, t: 0.5 * om2 * X * X, K=lambda P, t: 0.5 * P * P)
    o.set_wignerfunction(np.exp(-(o.X - 1) ** 2 - (o.P - 0.5) ** 2))
    o.propagate(400)
    assert rel_l2(o.W, leapfrog_gaussian(o, 400, om2)) < 1e-12


def test_unitary_invariants_and_ehrenfest():
    o = fo.MoyalOracle(256, 8., 256, 8., 0.01, V=lambda X, t: 0.5 * X * X + 0.1 * X ** 4,
                       K=lambda P, t: 0.5 * P * P, diff_V=lambda X, t: X + 0.4 * X ** 3, diff_K=lambda P, t: P)
    o.set_wignerfunction(np.exp(-(o.X - 1) ** 2 - (o.P - 0.5) ** 2) / np.pi)
    o.propagate(300)
    pur = np.array([s["purity"] for s in o.series])
    H = np.array([s["hamiltonian_average"] for s in o.series])
    assert np.max(np.abs(pur - 1.0)) < 1e-9          # pure Gaussian stays pure
    assert (H.max() - H.min()) / H.max() < 1e-4      # energy conserved to O(dt^2)
    x = np.array([s["X_average"] for s in o.series])
    p_rhs = np.array([s["X_average_RHS"] for s in o.series])
    dxdt = np.gradient(x, o.dt)
    assert np.max(np.abs(dxdt[2:-2] - p_rhs[2:-2])) < 1e-3  # d<x>/dt = <p>


# --------------------------------------------------------------------------- Bloch
def test_bloch_harmonic_gibbs_and_conventions():
    om, kT = 1.0, 0.5
    args = (256, 10., 256, 10., 0.01)
    kw = dict(V=lambda X, t: 0.5 * om * om * X * X, K=lambda P, t: 0.5 * P * P)
    b = fo.BlochOracle(*args, kT=kT, dbeta=0.005, convention="cuda", **kw)
    W = b.get_gibbs_state()
    assert rel_l2(W, fo.harmonic_gibbs(b, om, kT)) < 2e-6
    bf = fo.BlochOracle(*args, kT=kT, dbeta=0.005, convention="fftw", **kw)
    Wf = bf.get_gibbs_state()
    assert rel_l2(Wf, fo.harmonic_gibbs(b, om, kT / 2)) < 2e-5
    # FFTW(kT, dbeta) == CUDA(kT/2, 2 dbeta)
    bc = fo.BlochOracle(*args, kT=kT / 2, dbeta=0.01, convention="cuda", **kw)
    assert np.max(np.abs(Wf - bc.get_gibbs_state())) <= 1e-14 * np.max(np.abs(Wf))
    # Gibbs state nearly stationary under Moyal (wigner_bloch_fftw_1d.py:160-162)
    m = fo.MoyalOracle(*args, **kw)
    m.set_wignerfunction(W)
    m.propagate(200)
    assert rel_l2(m.W, W) < 1e-4


def test_dbeta_rounding():
    assert fo.dbeta_num_beta_steps(0.2, 0.005) == (0.005, 1000)
    db, n = fo.dbeta_num_beta_steps(0.3, 0.005)
    assert n == 667 and abs(db - 1.0 / (0.3 * 667)) < 1e-18


# --------------------------------------------------------------------------- random collision
def test_random_collision_limits():
    args = (128, 8., 128, 8., 0.01)
    kw = dict(V=lambda X, t: 0.5 * X * X + 0.05 * X ** 4, K=lambda P, t: 0.5 * P * P)
    rc = fo.RandomCollisionOracle(*args, gamma=0.0, kT=0.5, **kw)
    m = fo.MoyalOracle(*args, **kw)
    W0 = np.exp(-(m.X - 1) ** 2 - (m.P - 0.5) ** 2)
    rc.set_wignerfunction(W0)
    m.set_wignerfunction(W0)
    rc.propagate(30)
    m.propagate(30)
    assert np.max(np.abs(rc.W - m.W)) == 0.0
    # the dissipator keeps the Gibbs state fixed: W -> w W + (1-w) W_G
    rc = fo.RandomCollisionOracle(*args, gamma=0.08, kT=0.5, **kw)
    rc.set_wignerfunction(rc.gibbs)
    W = rc.W * rc.w_coeff + rc.scaled_gibbs_state
    assert np.max(np.abs(W - rc.W)) <= 1e-15 * np.max(np.abs(W))
    # relaxation towards the Gibbs state: the Moyal flow is an L2 isometry that
    # leaves W_G (nearly) fixed, so ||W - W_G|| decays exactly as exp(-gamma t)
    rc.set_wignerfunction(W0)
    d0 = np.linalg.norm(rc.W - rc.gibbs)
    rc.propagate(400)
    ratio = np.linalg.norm(rc.W - rc.gibbs) / d0
    assert abs(ratio - np.exp(-0.08 * 400 * 0.01)) < 1e-3


# --------------------------------------------------------------------------- density matrix
def gaussian_rho(r, x0=1.0, p0=0.5, s=1.0):
    psi = np.exp(-(r.Xv - x0) ** 2 / (2 * s * s) + 1j * p0 * r.Xv)
    return np.outer(psi.conj(), psi)


def test_rho_wigner_transform_of_gaussian():
    r = fo.RhoOracle(256, 10., 0.01, V=lambda X, t: 0.5 * X * X, K=lambda P, t: 0.5 * P * P)
    r.set_rho(gaussian_rho(r))
    W = r.get_wignerfunction()
    Wa = np.exp(-(r.X_wigner - 1.0) ** 2 - (r.P_wigner - 0.5) ** 2) / np.pi
    Wa /= Wa.sum() * r.wigner_dxdp
    assert rel_l2(W.real, Wa) < 1e-12
    assert np.max(np.abs(W.imag)) < 1e-12


def test_rho_invariants_and_moyal_agreement():
    V = lambda X, t: 0.5 * X * X + 0.05 * X ** 4
    K = lambda P, t: 0.5 * P * P
    r = fo.RhoOracle(256, 10., 0.01, V=V, K=K)
    r.set_rho(gaussian_rho(r))
    r.propagate(150)
    assert abs(np.trace(r.rho) * r.dX - 1.0) < 1e-13
    assert np.max(np.abs(r.rho - r.rho.conj().T)) < 1e-13
    assert abs(r.get_purity() - 1.0) < 1e-12
    m = fo.MoyalOracle(256, 10., 256, 10., 0.01, V=V, K=K, phase_time_shift=0.0)
    m.set_wignerfunction(np.exp(-(m.X - 1.0) ** 2 - (m.P - 0.5) ** 2))
    m.propagate(150)
    # two different discretisations of the same dynamics (identical only for
    # quadratic V): agreement to the discretisation error
    assert abs(r.get_average((lambda X, t: X,)) - m.series[-1]["X_average"]) < 1e-7
    assert abs(r.get_average((None, lambda P, t: P)) - m.series[-1]["P_average"]) < 1e-7


def test_rho_bloch_harmonic():
    kw = dict(V=lambda X, t: 0.5 * X * X, K=lambda P, t: 0.5 * P * P)
    rb = fo.RhoBlochOracle(128, 10., 0.01, kT=0.5, dbeta=0.005, **kw)
    rb.get_gibbs_state()
    W = rb.get_wignerfunction().real
    th = np.tanh(1.0)
    Wa = th / np.pi * np.exp(-2 * th * 0.5 * (rb.P_wigner ** 2 + rb.X_wigner ** 2))
    assert np.max(np.abs(W - Wa)) * rb.wigner_dxdp < 1e-8


# --------------------------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught(monkeypatch):
    """Plant slips in the oracle (on the CPU) and show the KATs reject them."""
    om2 = 1.3

    def run():
        o = fo.MoyalOracle(128, 8., 128, 8., 0.01, V=lambda X, t: 0.5 * om2 * X * X, K=lambda P, t: 0.5 * P * P)
        o.set_wignerfunction(np.exp(-(o.X - 1) ** 2 - (o.P - 0.5) ** 2))
        o.propagate(100)
        return rel_l2(o.W, leapfrog_gaussian(o, 100, om2))

    assert run() < 1e-12
    good_v, good_k = fo.MoyalOracle.expV, fo.MoyalOracle.expK
    # wrong sign of the potential phase
    monkeypatch.setattr(fo.MoyalOracle, "expV", lambda self, t: np.conj(good_v(self, t)))
    assert run() > 1e-3
    monkeypatch.setattr(fo.MoyalOracle, "expV", good_v)
    # dropped factor 1/2 of the half step (full V step twice)
    monkeypatch.setattr(fo.MoyalOracle, "expV", lambda self, t: good_v(self, t) ** 2)
    assert run() > 1e-3
    monkeypatch.setattr(fo.MoyalOracle, "expV", good_v)
    # transposed kinetic operand (lambda and p swapped)
    monkeypatch.setattr(fo.MoyalOracle, "expK", lambda self, t: good_k(self, t) * np.exp(1e-2j * self.Lambda))
    assert run() > 1e-3
