"""numpy/scipy restatement of the reference split-operator propagators.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the checker for the HIP
product path and the CPU baseline timed by ``bench.py``. Never shipped, never
called by ``fastwigner_amd``.

Transforms are ``scipy.fft.rfft/irfft`` (pocketfft, complex128), the in-repo
stand-in for pyFFTW's r2c/c2r plans (wigner_moyal_fftw_1d.py:332-366); the
reference's own FFTW path cannot run here (Python 2, no pyFFTW, no libfftw3).
Kernel formulas (phase arguments, grid coordinates, mask placement, time
argument) follow the CUDA templates, as the HIP product is the drop-in for the
CUDA classes.

Conventions fixed here and used identically by the HIP kernels
---------------------------------------------------------------
* Phase time argument. The CUDA Wigner kernels evaluate V and K at
  ``t + 0.5*dt`` where ``t`` is the time *after* the increment
  (wigner_moyal_cuda_1d.py:529,602,643), the FFTW class at ``t``
  (wigner_moyal_fftw_1d.py:402,409) and the CUDA density-matrix kernels at ``t``
  (rho_vneumann_cuda_1d.py:660,681). ``phase_time_shift`` s selects
  tau = t_{n+1} + s*dt: 0.5 for the CUDA Wigner classes (default), 0.0 for the
  FFTW classes and the density-matrix classes, -0.5 for the midpoint rule.
* Grid coordinates use the CUDA kernel formulas X = dX*(j - N_x/2),
  P = dP*(i - N_p/2), Lambda = dLambda*j, Theta = dTheta*i
  (wigner_moyal_cuda_1d.py:598-599,637-638). They equal the FFTW grids
  (np.linspace / fftfreq, wigner_moyal_fftw_1d.py:101-127) to the last bit for
  dyadic spacings; the Nyquist sign (FFTW -N/2, CUDA +N/2) is irrelevant because
  c2r drops the imaginary part there.
* Inverse transforms are normalised by 1/N like pyFFTW. cuFFT's are not
  (wigner_moyal_cuda_1d.py:439); the per-step renormalisation cancels the
  difference for Moyal/Bloch, and the random-collision dissipator, which mixes W
  with the Gibbs state, needs the normalised form (wigner_random_collision_model_fftw_1d.py:63-74).
"""
import numpy as np
import scipy.fft as sfft

NSER = 9  # per-step sums, in this order:
SER_NAMES = ("S0", "Sx", "Sp", "Sxx", "Spp", "SdK", "SdV", "SH", "SWW")


def _ones(*a):
    return np.ones(np.broadcast(*[np.asarray(v) for v in a]).shape)


class Grid:
    """Phase-space grids, wigner_moyal_cuda_1d.py:112-143 / wigner_moyal_fftw_1d.py:101-127."""

    def __init__(self, X_gridDIM, X_amplitude, P_gridDIM, P_amplitude):
        self.Nx, self.Np = int(X_gridDIM), int(P_gridDIM)
        self.X_amplitude, self.P_amplitude = float(X_amplitude), float(P_amplitude)
        self.dX = 2.0 * self.X_amplitude / self.Nx
        self.dP = 2.0 * self.P_amplitude / self.Np
        lam = np.fft.fftfreq(self.Nx, self.dX / (2 * np.pi))
        th = np.fft.fftfreq(self.Np, self.dP / (2 * np.pi))
        self.dLambda = lam[1] - lam[0]
        self.dTheta = th[1] - th[0]
        j = np.arange(self.Nx)
        i = np.arange(self.Np)
        self.X = (self.dX * (j - 0.5 * self.Nx))[np.newaxis, :]
        self.P = (self.dP * (i - 0.5 * self.Np))[:, np.newaxis]
        self.Lambda = (self.dLambda * np.arange(self.Nx // 2 + 1))[np.newaxis, :]
        self.Theta = (self.dTheta * np.arange(self.Np // 2 + 1))[:, np.newaxis]
        self.dXdP = self.dX * self.dP


class MoyalOracle(Grid):
    """WignerMoyal{FTTW,CUDA}1D restated.

    Step: wigner_moyal_fftw_1d.py:397-416 (transforms) with the CUDA phases
    wigner_moyal_cuda_1d.py:592-663 and the x-p mask :516-517; propagate
    :418-440 / :521-540 (t += dt, step, W /= sum(W) dXdP, observables).
    ``V(X, t)``, ``K(P, t)``, ``diff_V``, ``diff_K`` are numpy callables; masks are
    callables of (X, Theta, t), (Lambda, P, t) and (X, P, t).
    """

    def __init__(self, X_gridDIM, X_amplitude, P_gridDIM, P_amplitude, dt, V, K,
                 diff_V=None, diff_K=None, t=0.0, phase_time_shift=0.5,
                 mask_x_theta=None, mask_lambda_p=None, mask_x_p=None, workers=1):
        Grid.__init__(self, X_gridDIM, X_amplitude, P_gridDIM, P_amplitude)
        assert self.Nx % 2 == 0 and self.Np % 2 == 0  # wigner_moyal_fftw_1d.py:55,62
        self.dt, self.t = float(dt), float(t)
        self.V, self.K = V, K
        self.diff_V = diff_V or (lambda X, t: 0.0 * X)
        self.diff_K = diff_K or (lambda P, t: 0.0 * P)
        self.shift = float(phase_time_shift)
        self.mask_x_theta = mask_x_theta or _ones
        self.mask_lambda_p = mask_lambda_p or _ones
        self.mask_x_p = mask_x_p
        self.workers = workers
        self.W = np.zeros((self.Np, self.Nx))
        self.series = []

    # -- state ------------------------------------------------------------------
    def set_wignerfunction(self, W):
        """wigner_moyal_fftw_1d.py:368-395: copy, then W /= sum(W) dXdP."""
        W = np.asarray(W, dtype=np.float64)
        assert W.shape == (self.Np, self.Nx)
        self.W = W / (W.sum() * self.dXdP)
        return self

    # -- factors ------------------------------------------------------------------
    def expV(self, t):
        """expV_bulk/_boundary, wigner_moyal_cuda_1d.py:631-663 (theta rows 0..Np/2)."""
        tau = t + self.shift * self.dt
        X, Th = self.X, self.Theta
        ph = -0.5 * self.dt * (self.V(X - 0.5 * Th, tau) - self.V(X + 0.5 * Th, tau))
        return (np.cos(ph) + 1j * np.sin(ph)) * self.mask_x_theta(X, Th, t)

    def expK(self, t):
        """expK_bulk/_boundary, wigner_moyal_cuda_1d.py:592-622 (lambda cols 0..Nx/2)."""
        tau = t + self.shift * self.dt
        P, L = self.P, self.Lambda
        ph = -self.dt * (self.K(P + 0.5 * L, tau) - self.K(P - 0.5 * L, tau))
        return (np.cos(ph) + 1j * np.sin(ph)) * self.mask_lambda_p(L, P, t)

    # -- transforms -------------------------------------------------------------------
    def _vpass(self, W, f):
        Z = sfft.rfft(W, axis=0, workers=self.workers)
        Z *= f
        return sfft.irfft(Z, n=self.Np, axis=0, workers=self.workers)

    def _kpass(self, W, f):
        Z = sfft.rfft(W, axis=1, workers=self.workers)
        Z *= f
        return sfft.irfft(Z, n=self.Nx, axis=1, workers=self.workers)

    def single_step_propagation(self):
        """Strang V/2-K-V/2, wigner_moyal_cuda_1d.py:496-519 (W not normalised)."""
        fv = self.expV(self.t)
        W = self._vpass(self.W, fv)
        W = self._kpass(W, self.expK(self.t))
        W = self._vpass(W, fv)
        if self.mask_x_p is not None:
            W *= self.mask_x_p(self.X, self.P, self.t)
        self.W = W
        return W

    def sums(self, t=None):
        """The nine per-step sums of the HIP epilogue over the *unnormalised* W."""
        t = self.t if t is None else t
        W, X, P = self.W, self.X, self.P
        cx = W.sum(axis=0)      # marginal over p for each x
        cp = W.sum(axis=1)      # marginal over x for each p
        return np.array([
            W.sum(), (cx * X[0]).sum(), (cp * P[:, 0]).sum(),
            (cx * X[0] ** 2).sum(), (cp * P[:, 0] ** 2).sum(),
            (cp * np.broadcast_to(self.diff_K(P, t), P.shape)[:, 0]).sum(),
            (cx * np.broadcast_to(self.diff_V(X, t), X.shape)[0]).sum(),
            (cp * np.broadcast_to(self.K(P, t), P.shape)[:, 0]).sum()
            + (cx * np.broadcast_to(self.V(X, t), X.shape)[0]).sum(),
            (W * W).sum(),
        ])

    def observables(self, s):
        """From the sums: <x>, <diff_K>, <p>, -<diff_V>, <H>, purity, sigma_x sigma_p.
        get_Ehrenfest wigner_moyal_cuda_1d.py:542-562; get_purity :398-403;
        get_sigma_x_sigma_p :405-415."""
        S0 = s[0]
        x, p = s[1] / S0, s[2] / S0
        with np.errstate(invalid="ignore"):
            # a negative variance product (coarse grids) gives NaN, as np.sqrt does in the reference
            sxsp = np.sqrt((s[3] / S0 - x * x) * (s[4] / S0 - p * p))
        return dict(
            X_average=x, X_average_RHS=s[5] / S0, P_average=p, P_average_RHS=-s[6] / S0,
            hamiltonian_average=s[7] / S0,
            purity=2.0 * np.pi * s[8] / (S0 * S0 * self.dXdP),
            sigma_x_sigma_p=sxsp,
        )

    def propagate(self, steps=1):
        """wigner_moyal_cuda_1d.py:521-540."""
        for _ in range(steps):
            self.t += self.dt
            self.single_step_propagation()
            s = self.sums(self.t)
            self.series.append(self.observables(s))
            self.W /= s[0] * self.dXdP
        return self.W

    def get_average(self, f):
        """wigner_moyal_cuda_1d.py:378-396 with f a callable of (X, P, t)."""
        return (self.W * f(self.X, self.P, self.t)).sum() * self.dXdP

    def get_purity(self):
        return 2.0 * np.pi * (self.W * self.W).sum() * self.dXdP


def absorbing_alpha_mask(g, alpha, dt):
    """FFTW absorbing boundary, wigner_moyal_fftw_1d.py:135-156, as a (X, Theta) mask."""
    Xm, Xp = g.X - 0.5 * g.Theta, g.X + 0.5 * g.Theta
    # the reference takes the min/max over its own Theta grid, whose Nyquist row is
    # negative (fftfreq); the CUDA grid's is positive. Use the FFTW extent.
    th_fftw = np.fft.fftfreq(g.Np, g.dP / (2 * np.pi))[: g.Np // 2 + 1][:, None]
    xmin = min((g.X - 0.5 * th_fftw).min(), (g.X + 0.5 * th_fftw).min())
    xmax = max((g.X - 0.5 * th_fftw).max(), (g.X + 0.5 * th_fftw).max())
    m = np.sin(np.pi * (Xp - xmin) / (xmax - xmin)) * np.sin(np.pi * (Xm - xmin) / (xmax - xmin))
    return np.abs(m) ** abs(alpha * dt), xmin, xmax


# ------------------------------------------------------------------------------------
# Bloch (thermal state)
# ------------------------------------------------------------------------------------

def dbeta_num_beta_steps(kT, dbeta):
    """wigner_bloch_cuda_1d.py:104-127 (same rule in wigner_bloch_fftw_1d.py:34-42)."""
    if kT > 0:
        n = 1.0 / (kT * dbeta)
        if round(n) != n:
            n = round(n)
            dbeta = 1.0 / (kT * n)
        return dbeta, int(n)
    return dbeta, np.inf


class BlochOracle(MoyalOracle):
    """Imaginary-time split operator.

    convention="cuda": exp(-dbeta/4 [V(X-Th/2)+V(X+Th/2)-V_min]) and
        exp(-dbeta/2 [K(P+L/2)+K(P-L/2)-K_min]) at t_initial, masks off
        (wigner_bloch_cuda_1d.py:290-359, mins :72-102).
    convention="fftw": exp(-dbeta/2 [..]) * absorbing mask and exp(-dbeta [..])
        (wigner_bloch_fftw_1d.py:65-100), i.e. FFTW(kT, dbeta) == CUDA(kT/2, 2 dbeta).
    Each step: V, K, V factors, then W /= sum(W) dXdP (wigner_bloch_cuda_1d.py:229-254).
    """

    def __init__(self, *a, kT=0.0, dbeta=0.01, t_initial=0.0, convention="cuda", **kw):
        MoyalOracle.__init__(self, *a, **kw)
        self.kT, self.dbeta, self.t_initial = kT, dbeta, t_initial
        self.convention = convention
        X, P, Th, L = self.X, self.P, self.Theta, self.Lambda
        self.Vsum = self.V(X - 0.5 * Th, t_initial) + self.V(X + 0.5 * Th, t_initial)
        self.Ksum = self.K(P + 0.5 * L, t_initial) + self.K(P - 0.5 * L, t_initial)
        self.V_min, self.K_min = self.Vsum.min(), self.Ksum.min()

    def factors(self, dbeta):
        cv, ck = (0.25, 0.5) if self.convention == "cuda" else (0.5, 1.0)
        fv = np.exp(-cv * dbeta * (self.Vsum - self.V_min))
        fk = np.exp(-ck * dbeta * (self.Ksum - self.K_min))
        if self.convention == "fftw":
            fv = fv * self.mask_x_theta(self.X, self.Theta, self.t)
        return fv, fk

    def bloch_single_step_propagation(self, dbeta):
        fv, fk = self.factors(dbeta)
        W = self._vpass(self.W, fv)
        W = self._kpass(W, fk)
        W = self._vpass(W, fv)
        self.W = W / (W.sum() * self.dXdP)
        return self.W

    def get_gibbs_state(self, kT=None, dbeta=None):
        """wigner_bloch_cuda_1d.py:129-165: uniform start, n steps, purity guard."""
        kT = self.kT if kT is None else kT
        dbeta = self.dbeta if dbeta is None else dbeta
        dbeta, n = dbeta_num_beta_steps(kT, dbeta)
        self.set_wignerfunction(np.full((self.Np, self.Nx), 1.0 / (self.Np * self.Nx)))
        self.steps_done = 0
        for k in range(n):
            self.bloch_single_step_propagation(dbeta)
            self.steps_done = k + 1
            if not self.get_purity() < 1.0:
                break
        return self.W


def harmonic_gibbs(g, omega, kT):
    """Closed-form HO Gibbs Wigner function, wigner_bloch_fftw_1d.py:152-154."""
    th = np.tanh(0.5 * omega / kT)
    H = 0.5 * (g.P ** 2 + omega ** 2 * g.X ** 2)
    return th / np.pi * np.exp(-2.0 * th * H / omega)


class RandomCollisionOracle(MoyalOracle):
    """WignerRandomCollisionModelFFTW1D restated
    (wigner_random_collision_model_fftw_1d.py:16-94): w = exp(-gamma dt/2);
    W0s = (1-w) W_Gibbs with W_Gibbs from the Bloch oracle (FFTW convention by
    default, as the reference builds it with WignerBlochFFTW1D :35-38);
    step = dissipator, Moyal step, dissipator (:56-76); Ehrenfest RHS amended by
    gamma(<.>_Gibbs - <.>) (:78-94)."""

    def __init__(self, *a, gamma, kT, dbeta=0.01, bloch_convention="fftw", gibbs=None, **kw):
        MoyalOracle.__init__(self, *a, **kw)
        self.gamma = float(gamma)
        self.w_coeff = np.exp(-self.gamma * 0.5 * self.dt)
        if gibbs is None:
            b = BlochOracle(*a, kT=kT, dbeta=dbeta, convention=bloch_convention, **kw)
            gibbs = b.get_gibbs_state()
        self.gibbs = np.array(gibbs)
        self.scaled_gibbs_state = self.gibbs * (1.0 - self.w_coeff)
        # the reference normalises the scaled state's marginals (:44-54); the
        # unscaled state gives the same averages and stays finite at gamma = 0
        cx = self.gibbs.sum(axis=0)
        cp = self.gibbs.sum(axis=1)
        self.X_average_gibbs = (cx / cx.sum() * self.X[0]).sum()
        self.P_average_gibbs = (cp / cp.sum() * self.P[:, 0]).sum()

    def single_step_propagation(self):
        self.W = self.W * self.w_coeff + self.scaled_gibbs_state
        MoyalOracle.single_step_propagation(self)
        self.W = self.W * self.w_coeff + self.scaled_gibbs_state
        return self.W

    def observables(self, s):
        o = MoyalOracle.observables(self, s)
        o["X_average_RHS"] += self.gamma * (self.X_average_gibbs - o["X_average"])
        o["P_average_RHS"] += self.gamma * (self.P_average_gibbs - o["P_average"])
        return o


# ------------------------------------------------------------------------------------
# Density matrix (von Neumann, Bloch, Wigner transform)
# ------------------------------------------------------------------------------------

class RhoOracle:
    """RhoVNeumannCUDA1D restated (rho_vneumann_cuda_1d.py).

    rho[i, j] = <X_j| rho |X'_i> (i = row = x', j = column = x), complex128 N x N.
    Grids :93-110; expV :672-684, expK :650-663 with FFT-ordered momenta;
    step :340-353 (expV, fft2, expK, ifft2, expV); propagate :355-374
    (rho /= Tr(rho) dX). Time argument tau = t_{n+1} + shift*dt, shift 0 as in
    the reference.
    """

    def __init__(self, X_gridDIM, X_amplitude, dt, V, K, diff_V=None, diff_K=None, t=0.0,
                 phase_time_shift=0.0, mask_x=None, mask_p=None):
        self.N = int(X_gridDIM)
        self.X_amplitude = float(X_amplitude)
        self.dX = 2.0 * self.X_amplitude / self.N
        P = np.fft.fftfreq(self.N, self.dX / (2 * np.pi))
        self.dP = P[1] - P[0]
        j = np.arange(self.N)
        self.Xv = self.dX * (j - 0.5 * self.N)
        self.Pv = self.dP * ((j + self.N // 2) % self.N - 0.5 * self.N)
        self.X = self.Xv[np.newaxis, :]        # column coordinate
        self.Xp = self.Xv[:, np.newaxis]       # row coordinate (X_prime)
        self.P = self.Pv[np.newaxis, :]
        self.Pp = self.Pv[:, np.newaxis]
        self.dt, self.t = float(dt), float(t)
        self.V, self.K = V, K
        self.diff_V = diff_V or (lambda X, t: 0.0 * X)
        self.diff_K = diff_K or (lambda P, t: 0.0 * P)
        self.shift = float(phase_time_shift)
        self.mask_x = mask_x or _ones
        self.mask_p = mask_p or _ones
        self.rho = np.zeros((self.N, self.N), np.complex128)
        # Wigner grids :103-110
        self.X_wigner = self.X / np.sqrt(2.0)
        self.P_wigner = np.fft.fftshift(self.Pv / np.sqrt(2.0))[:, np.newaxis]
        self.wigner_dxdp = (self.Xv[1] - self.Xv[0]) / np.sqrt(2.0) * (P[1] - P[0]) / np.sqrt(2.0)

    def trace(self, r=None):
        r = self.rho if r is None else r
        return np.trace(r)

    def set_rho(self, rho):
        """:301-338: copy, then rho /= Tr(rho) dX."""
        rho = np.asarray(rho, dtype=np.complex128)
        self.rho = rho / (np.trace(rho) * self.dX)
        return self

    def expV(self, t):
        tau = t + self.shift * self.dt
        ph = -0.5 * self.dt * (self.V(self.X, tau) - self.V(self.Xp, tau))
        return (np.cos(ph) + 1j * np.sin(ph)) * self.mask_x(self.X, self.Xp, t)

    def expK(self, t):
        tau = t + self.shift * self.dt
        ph = -self.dt * (self.K(self.P, tau) - self.K(self.Pp, tau))
        return (np.cos(ph) + 1j * np.sin(ph)) * self.mask_p(self.P, self.Pp, t)

    def single_step_propagation(self):
        fv = self.expV(self.t)
        r = self.rho * fv
        r = np.fft.fft2(r)
        r *= self.expK(self.t)
        r = np.fft.ifft2(r)
        r *= fv
        self.rho = r
        return r

    def propagate(self, steps=1):
        for _ in range(steps):
            self.t += self.dt
            self.single_step_propagation()
            self.rho /= np.trace(self.rho) * self.dX
        return self.rho

    # -- observables (:431-500) -----------------------------------------------------
    def get_average(self, observable):
        """observable = tuple of callables alternating x / p multipliers (None skips).
        x: rho[i, j] *= f(X_j); p: fft along axis 1, /N, *= g(P_j), ifft (:452-470)."""
        tmp = self.rho.copy()
        is_x = False
        for f in observable:
            is_x = not is_x
            if f is None:
                continue
            if is_x:
                tmp = tmp * f(self.X, self.t)
            else:
                tmp = np.fft.ifft(np.fft.fft(tmp, axis=1) * f(self.P, self.t), axis=1)
        return np.trace(tmp).real * self.dX

    def get_purity(self):
        """:474-488: sum |rho|^2 dX^2."""
        return (np.abs(self.rho) ** 2).sum() * self.dX ** 2

    def get_sigma_x_sigma_p(self):
        x2 = self.get_average((lambda X, t: X * X,))
        x1 = self.get_average((lambda X, t: X,))
        p2 = self.get_average((None, lambda P, t: P * P))
        p1 = self.get_average((None, lambda P, t: P))
        return np.sqrt((x2 - x1 ** 2) * (p2 - p1 ** 2))

    def get_wignerfunction(self):
        """Larkin 3-shear 45 degree rotation + sign-flipped FFT, :502-592."""
        a = np.tan(np.pi / 8.0)
        b = -np.sin(np.pi / 4.0)
        shearX = np.exp(1j * (-a * self.P * self.Xp))      # phase_shearX :753-767
        shearY = np.exp(1j * (-b * self.Pp * self.X))      # phase_shearY :769-783
        tmp = np.fft.ifft(np.fft.fft(self.rho, axis=1) * shearX, axis=1)
        tmp = np.fft.ifft(np.fft.fft(tmp, axis=0) * shearY, axis=0)
        tmp = np.fft.ifft(np.fft.fft(tmp, axis=1) * shearX, axis=1)
        flip = (1 - 2 * (np.arange(self.N) % 2))[:, np.newaxis]
        tmp = tmp * flip
        tmp = np.fft.ifft(tmp, axis=0) * self.N            # cuFFT inverse is unnormalised
        tmp = tmp * flip
        tmp /= tmp.sum() * self.wigner_dxdp
        return tmp

    def get_wigner_time(self, wigner_current, wigner_init, t, K, V, Ip):
        """:594-620 with the reference's indicator 0.5*(1+signbit(H+Ip)) (:854)."""
        N = self.N
        j = np.arange(N)
        X = (self.dX * (j - 0.5 * N) / np.sqrt(2.0))[np.newaxis, :]
        P = (self.dP * (j - 0.5 * N) / np.sqrt(2.0))[:, np.newaxis]
        H = K(P, t) + V(X, t)
        ind = 0.5 * (1 + np.signbit(H + Ip).astype(np.float64))
        return (ind * (np.real(wigner_current) - wigner_init)).sum() * self.wigner_dxdp


class RhoBlochOracle(RhoOracle):
    """RhoBlochCUDA1D restated (rho_bloch_cuda_1d.py): start from rho = I (:142),
    exp(-dbeta/4 [V(X)+V(X')-V_min]) and exp(-dbeta/2 [K(P)+K(P')-K_min]) at
    t_initial (:271-309), K_min == 0 because fill_K multiplies a zero array
    (:354-365), trace renormalisation each step (:224-241)."""

    def __init__(self, *a, kT=0.0, dbeta=0.01, t_initial=0.0, **kw):
        RhoOracle.__init__(self, *a, **kw)
        self.kT, self.dbeta, self.t_initial = kT, dbeta, t_initial
        self.V_min = (self.V(self.X, t_initial) + self.V(self.Xp, t_initial)).min()
        self.K_min = 0.0

    def bloch_single_step_propagation(self, dbeta):
        ti = self.t_initial
        fv = np.exp(-0.25 * dbeta * (self.V(self.X, ti) + self.V(self.Xp, ti) - self.V_min))
        fk = np.exp(-0.5 * dbeta * (self.K(self.P, ti) + self.K(self.Pp, ti) - self.K_min))
        r = np.fft.ifft2(np.fft.fft2(self.rho * fv) * fk) * fv
        self.rho = r / (np.trace(r) * self.dX)
        return self.rho

    def get_gibbs_state(self, kT=None, dbeta=None):
        kT = self.kT if kT is None else kT
        dbeta = self.dbeta if dbeta is None else dbeta
        dbeta, n = dbeta_num_beta_steps(kT, dbeta)
        self.set_rho(np.eye(self.N, dtype=np.complex128))
        for _ in range(n):
            self.bloch_single_step_propagation(dbeta)
            if not self.get_purity() < 1.0:
                break
        return self.rho
