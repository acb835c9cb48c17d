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


# --------------------------------------------------------------------------- observables in closed form
def leapfrog_moments(o, n, om2, x0=1.0, p0=0.5, var=0.5):
    """Mean and covariance of the Gaussian under n exact leapfrog (kick-drift-kick) maps."""
    a = 0.5 * o.dt * om2
    kick = np.array([[1, 0], [-a, 1]])
    M = kick @ np.array([[1, o.dt], [0, 1]]) @ kick
    Mn = np.linalg.matrix_power(M, n)
    return Mn @ np.array([x0, p0]), Mn @ (var * np.eye(2)) @ Mn.T


def harmonic_run(n=400, om2=1.3):
    o = fo.MoyalOracle(256, 8., 256, 8., 0.01, V=lambda X, t: 0.5 * om2 * X * X, K=lambda P, t: 0.5 * P * P,
                       diff_V=lambda X, t: om2 * X, diff_K=lambda P, t: P)
    o.set_wignerfunction(np.exp(-(o.X - 1) ** 2 - (o.P - 0.5) ** 2))
    o.propagate(n)
    return o


def test_moyal_observables_of_the_leapfrog_gaussian():
    """Every per-step observable of the harmonic run against the moments of the
    closed-form Gaussian (get_Ehrenfest, get_purity, get_sigma_x_sigma_p,
    wigner_moyal_cuda_1d.py:398-415,542-562): <x>, <p>, <x^2>, <p^2>, <K'> = <p>,
    -<V'> = -om2 <x>, <H> = (<p^2> + om2 <x^2>)/2, purity 1 (pure state) and
    sigma_x sigma_p = sqrt(S_xx S_pp) of the covariance M^n (I/2) M^nT."""
    om2, n = 1.3, 400
    o = harmonic_run(n, om2)
    mu, S = leapfrog_moments(o, n, om2)
    s = o.series[-1]
    sm = o.sums(o.t)
    x2, p2 = S[0, 0] + mu[0] ** 2, S[1, 1] + mu[1] ** 2
    assert abs(s["X_average"] - mu[0]) < 1e-12
    assert abs(s["P_average"] - mu[1]) < 1e-12
    assert abs(sm[3] / sm[0] - x2) < 2e-12 * x2
    assert abs(sm[4] / sm[0] - p2) < 2e-12 * p2
    assert abs(s["X_average_RHS"] - mu[1]) < 1e-12
    assert abs(s["P_average_RHS"] + om2 * mu[0]) < 1e-12
    assert abs(s["hamiltonian_average"] - 0.5 * (p2 + om2 * x2)) < 2e-12
    assert abs(s["purity"] - 1.0) < 1e-12
    sxsp = np.sqrt(S[0, 0] * S[1, 1])
    assert abs(s["sigma_x_sigma_p"] - sxsp) < 5e-12 * sxsp
    # get_average of the unnormalised observable string's role: <x> and <x p>
    assert abs(o.get_average(lambda X, P, t: X) - mu[0]) < 1e-12
    assert abs(o.get_average(lambda X, P, t: X * P) - (S[0, 1] + mu[0] * mu[1])) < 2e-12


@pytest.mark.parametrize("conv,kT_eff", [("cuda", 0.5), ("fftw", 0.25)])
def test_gibbs_sigma_and_purity_closed_form(conv, kT_eff):
    """Harmonic Gibbs state: sigma_x sigma_p = coth(om / 2kT)/2 and purity
    tanh(om / 2kT) (om = 1), with the FFTW convention at kT/2; to the O(dbeta^2)
    splitting error."""
    b = fo.BlochOracle(256, 10., 256, 10., 0.01, kT=0.5, dbeta=0.005, convention=conv,
                       V=lambda X, t: 0.5 * X * X, K=lambda P, t: 0.5 * P * P)
    b.get_gibbs_state()
    ob = b.observables(b.sums(b.t))
    assert abs(ob["sigma_x_sigma_p"] - 0.5 / np.tanh(0.5 / kT_eff)) < 1e-6
    assert abs(b.get_purity() - np.tanh(0.5 / kT_eff)) < 1e-6
    assert abs(ob["purity"] - np.tanh(0.5 / kT_eff)) < 1e-6


def test_random_collision_ehrenfest_with_dissipation():
    """d<x>/dt and d<p>/dt of the dissipative run equal the amended right-hand
    sides <K'> + gamma(<x>_G - <x>) and -<V'> + gamma(<p>_G - <p>)
    (wigner_random_collision_model_fftw_1d.py:78-94), to the O(dt^2) of the
    finite differences; without the gamma terms they differ by ~1e-1."""
    rc = fo.RandomCollisionOracle(128, 8., 128, 8., 0.01, gamma=0.08, kT=0.5,
                                  V=lambda X, t: 0.5 * X * X + 0.05 * X ** 4, K=lambda P, t: 0.5 * P * P,
                                  diff_V=lambda X, t: X + 0.2 * X ** 3, diff_K=lambda P, t: P)
    rc.set_wignerfunction(np.exp(-(rc.X - 1) ** 2 - (rc.P - 0.5) ** 2))
    rc.propagate(300)
    ser = {k: np.array([s[k] for s in rc.series]) for k in ("X_average", "X_average_RHS", "P_average",
                                                            "P_average_RHS")}
    dx = np.gradient(ser["X_average"], rc.dt)[2:-2]
    dp = np.gradient(ser["P_average"], rc.dt)[2:-2]
    assert np.max(np.abs(dx - ser["X_average_RHS"][2:-2])) < 1e-4
    assert np.max(np.abs(dp - ser["P_average_RHS"][2:-2])) < 1e-3
    assert np.max(np.abs(dx - ser["P_average"][2:-2])) > 1e-2


@pytest.mark.parametrize("c", [0.0, 0.3])
def test_rho_second_moments_closed_form(c):
    """Chirped Gaussian psi = exp(-(x-x0)^2/(2 s^2) + i c (x-x0)^2 + i p0 x):
    <x^2> = x0^2 + s^2/2, <p^2> = p0^2 + 1/(2 s^2) + 2 c^2 s^2 and
    sigma_x sigma_p = sqrt(1 + 4 c^2 s^4)/2, through get_average tuples and
    get_sigma_x_sigma_p (rho_vneumann_cuda_1d.py:431-500)."""
    r = fo.RhoOracle(256, 10., 0.01, V=lambda X, t: 0.5 * X * X, K=lambda P, t: 0.5 * P * P)
    x0, p0, s = 1.0, 0.5, 1.0
    psi = np.exp(-(r.Xv - x0) ** 2 / (2 * s * s) + 1j * c * (r.Xv - x0) ** 2 + 1j * p0 * r.Xv)
    r.set_rho(np.outer(psi.conj(), psi))
    assert abs(r.get_average((lambda X, t: X,)) - x0) < 1e-12
    assert abs(r.get_average((None, lambda P, t: P)) - p0) < 1e-12
    assert abs(r.get_average((lambda X, t: X * X,)) - (x0 ** 2 + s * s / 2)) < 1e-12
    assert abs(r.get_average((None, lambda P, t: P * P)) - (p0 ** 2 + 1 / (2 * s * s) + 2 * c * c * s * s)) < 1e-12
    assert abs(r.get_sigma_x_sigma_p() - 0.5 * np.sqrt(1 + 4 * c * c * s ** 4)) < 1e-12


def wigner_grid(r):
    """The Wigner grid of the density-matrix classes (rho_vneumann_cuda_1d.py:103-110):
    x_j = dX (j - N/2)/sqrt(2) along columns, p_i = dP (i - N/2)/sqrt(2) along rows."""
    j = np.arange(r.N)
    return ((r.dX * (j - 0.5 * r.N) / np.sqrt(2.0))[np.newaxis, :],
            (r.dP * (j - 0.5 * r.N) / np.sqrt(2.0))[:, np.newaxis])


def test_wigner_time_on_the_grid():
    """get_wigner_time's indicator 0.5 (1 + signbit(H + Ip)) (rho_vneumann_cuda_1d.py:594-620,854)
    exactly on the grid: 1 for a state well inside H < -Ip, the reference's 0.5
    for one well outside, 0 when W = W_init, 0.5 everywhere for Ip > 0 (H >= 0)."""
    r = fo.RhoOracle(256, 10., 0.01, V=lambda X, t: 0.5 * X * X, K=lambda P, t: 0.5 * P * P)
    X, P = wigner_grid(r)
    K, V = (lambda P, t: 0.5 * P * P), (lambda X, t: 0.5 * X * X)
    inside = np.exp(-(X * X + P * P) / 0.18)
    outside = np.exp(-((X - 6.0) ** 2 + P * P) / 0.18)
    tot_in, tot_out = inside.sum() * r.wigner_dxdp, outside.sum() * r.wigner_dxdp
    assert abs(r.get_wigner_time(inside, 0 * inside, 0.0, K, V, -8.0) - tot_in) < 1e-12 * tot_in
    assert abs(r.get_wigner_time(outside, 0 * outside, 0.0, K, V, -1.0) - 0.5 * tot_out) < 1e-12 * tot_out
    assert r.get_wigner_time(outside, outside, 0.0, K, V, -1.0) == 0.0
    assert abs(r.get_wigner_time(inside, 0 * inside, 0.0, K, V, 0.5) - 0.5 * tot_in) < 1e-12 * tot_in


def wigner_time_error(N, A, om=2.0):
    """|get_wigner_time - (1 - exp(2 Ip/om)/2)| for W = exp(-(om x^2 + p^2/om))/pi,
    W_init = 0, H = (p^2 + om^2 x^2)/2, Ip < 0: the integral of W over H < -Ip is
    1 - exp(2 Ip/om) and the reference's indicator adds half the rest."""
    r = fo.RhoOracle(N, A, 0.01, V=lambda X, t: 0.5 * om * om * X * X, K=lambda P, t: 0.5 * P * P)
    X, P = wigner_grid(r)
    W = np.exp(-(om * X * X + P * P / om)) / np.pi
    errs = []
    for Ip in (-0.5, -1.0, -2.0):
        v = r.get_wigner_time(W, 0 * W, 0.0, lambda P, t: 0.5 * P * P, lambda X, t: 0.5 * om * om * X * X, Ip)
        errs.append(abs(v - (1.0 - 0.5 * np.exp(2.0 * Ip / om))))
    return max(errs)


def test_wigner_time_closed_form_and_refinement():
    """The closed form above to the quadrature error of the indicator's edge,
    which shrinks as both grid spacings are refined (x_amp grows as sqrt(N));
    a transposed state or a dropped 1/2 misses by 3e-2 or more."""
    e256 = wigner_time_error(256, 10.)
    e1024 = wigner_time_error(1024, 20.)
    assert e256 < 5e-3
    assert e1024 < 1.5e-3
    assert e1024 < 0.5 * e256


def test_planted_observable_mistakes_are_caught(monkeypatch):
    """Slips planted in the oracle's observables fail the closed-form checks above."""
    good_obs, good_sums = fo.MoyalOracle.observables, fo.MoyalOracle.sums

    def sxsp_err():
        om2, n = 1.3, 50
        o = harmonic_run(n, om2)
        mu, S = leapfrog_moments(o, n, om2)
        return abs(o.series[-1]["sigma_x_sigma_p"] - np.sqrt(S[0, 0] * S[1, 1]))

    assert sxsp_err() < 1e-11
    # <p^2> sum weighted by P instead of P^2
    def bad_sums(self, t=None):
        s = good_sums(self, t)
        s[4] = s[2]
        return s
    monkeypatch.setattr(fo.MoyalOracle, "sums", bad_sums)
    assert not sxsp_err() < 1e-3          # a wrong value or a NaN both fail the bar
    monkeypatch.setattr(fo.MoyalOracle, "sums", good_sums)

    # variance without the squared mean
    def bad_obs(self, s):
        o = good_obs(self, s)
        o["sigma_x_sigma_p"] = np.sqrt((s[3] / s[0]) * (s[4] / s[0]))
        return o
    monkeypatch.setattr(fo.MoyalOracle, "observables", bad_obs)
    assert not sxsp_err() < 1e-3
    monkeypatch.setattr(fo.MoyalOracle, "observables", good_obs)

    # Wigner time: indicator 0 outside the region (the reference keeps 0.5), and
    # the Wigner grid without its 1/sqrt(2)
    good_wt = fo.RhoOracle.get_wigner_time

    def no_quirk(self, wc, wi, t, K, V, Ip):
        return good_wt(self, wc, wi, t, K, V, Ip) - 0.5 * good_wt(self, wc, wi, t, K, V, 1.0)
    monkeypatch.setattr(fo.RhoOracle, "get_wigner_time", no_quirk)
    assert wigner_time_error(256, 10.) > 3e-2
    monkeypatch.setattr(fo.RhoOracle, "get_wigner_time", good_wt)

    def no_sqrt2(self, wc, wi, t, K, V, Ip):
        return good_wt(self, wc, wi, t, lambda P, t: K(P * np.sqrt(2.0), t), lambda X, t: V(X * np.sqrt(2.0), t), Ip)
    monkeypatch.setattr(fo.RhoOracle, "get_wigner_time", no_sqrt2)
    assert wigner_time_error(256, 10.) > 3e-2
    monkeypatch.setattr(fo.RhoOracle, "get_wigner_time", good_wt)
    # a second-moment slip of the density matrix: <p^2> through one FFT pair per P
    good_avg = fo.RhoOracle.get_average

    def bad_avg(self, obs):
        if len(obs) == 2 and obs[0] is None:
            g = obs[1]
            return good_avg(self, (None, lambda P, t: g(P, t) / np.maximum(np.abs(P), 1e-300) * P))
        return good_avg(self, obs)
    monkeypatch.setattr(fo.RhoOracle, "get_average", bad_avg)
    with pytest.raises(AssertionError):
        test_rho_second_moments_closed_form(0.3)
