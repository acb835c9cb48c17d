"""Moyal propagator of the Wigner function on MI355X.

``WignerMoyalHIP1D`` is the drop-in for ``WignerMoyalCUDA1D``
(wigner_moyal_cuda_1d.py:10-731): same keyword arguments (grid sizes and
amplitudes, ``dt``, ``t``, ``V``/``K``/``diff_V``/``diff_K`` as C-expression
strings, ``consts``, ``functions``, the three absorbing-boundary strings,
``max_thread_block``), same methods (``set_wignerfunction``, ``propagate``,
``single_step_propagation``, ``get_average``, ``get_purity``,
``get_sigma_x_sigma_p``, ``get_Ehrenfest``) and attributes (``X``, ``P``,
``Lambda``, ``Theta``, ``dX``, ``dP``, ``dLambda``, ``dTheta``, ``dXdP``, ``t``,
``wignerfunction`` and the Ehrenfest lists).

Additional keyword arguments:
    precision         "fp64" (default) or "fp32" (W and FFTs in FP32, the phase
                      argument with its mod-2pi reduction and every sum in FP64)
    phase_time_shift  V and K are evaluated at t_{n+1} + shift*dt; default 0.5
                      reproduces the CUDA kernels (wigner_moyal_cuda_1d.py:529,602,643),
                      0.0 the FFTW class, -0.5 the midpoint rule
    alpha             the FFTW class's absorbing boundary (wigner_moyal_fftw_1d.py:135-156),
                      translated into an abs_boundary_x_theta expression
    step_passes       0 (default) or 2: Moyal steps in two fused passes; 3: three passes
    phase_tables      1 / 0 / -1: cache the V and K factors in tables always /
                      when the strings do not reference t (default) / never
    device, tiles     HIP device ordinal; (col_e, col_p, row_e, row_p) FFT tiles
                      (default: the autotuned entry of fastwigner_amd/tuning.json)
    v_table           0: the V factors follow phase_tables; -1: a time-independent V is
                      evaluated in the column pass instead of read from a table
                      (default: the tuned entry, for a V free of t)
"""
import ctypes

import numpy as np

from . import _lib, tuning
from ._common import (DeviceArray, bind_kwargs, c_double, consts_source, precision_code)


class WignerMoyalHIP1D:
    """Second-order split-operator propagator of the Moyal equation for W(x, p, t)
    with H = K(p, t) + V(x, t) (R. Cabrera, D. I. Bondar, K. Jacobs, H. A. Rabitz,
    Phys. Rev. A 92, 042122 (2015)), on hand-written gfx950 kernels."""

    # kwargs that must not become C constants (besides the grid constants)
    _not_consts = ("t",)

    def __init__(self, **kwargs):
        bind_kwargs(self, kwargs)

        # Required attributes, wigner_moyal_cuda_1d.py:57-102
        try:
            self.X_gridDIM
        except AttributeError:
            raise AttributeError("Coordinate grid size (X_gridDIM) was not specified")
        assert self.X_gridDIM % 2 == 0, "Coordinate grid size (X_gridDIM) must be even"
        try:
            self.P_gridDIM
        except AttributeError:
            raise AttributeError("Momentum grid size (P_gridDIM) was not specified")
        assert self.P_gridDIM % 2 == 0, "Momentum grid size (P_gridDIM) must be even"
        try:
            self.X_amplitude
        except AttributeError:
            raise AttributeError("Coordinate grid range (X_amplitude) was not specified")
        try:
            self.P_amplitude
        except AttributeError:
            raise AttributeError("Momentum grid range (P_amplitude) was not specified")
        try:
            self.V
        except AttributeError:
            raise AttributeError("Potential energy (V) was not specified")
        try:
            self.K
        except AttributeError:
            raise AttributeError("Momentum dependence (K) was not specified")
        try:
            self.dt
        except AttributeError:
            raise AttributeError("Time-step (dt) was not specified")
        try:
            self.t
        except AttributeError:
            print("Warning: Initial time (t) was not specified, thus it is set to zero.")
            self.t = 0.
        self.t = np.float64(self.t)

        # Grids, wigner_moyal_cuda_1d.py:112-143
        self.dX = 2. * self.X_amplitude / self.X_gridDIM
        self.dP = 2. * self.P_amplitude / self.P_gridDIM
        self.X = np.linspace(-self.X_amplitude, self.X_amplitude - self.dX, self.X_gridDIM)[np.newaxis, :]
        self.Lambda = np.fft.fftfreq(self.X_gridDIM, self.dX / (2 * np.pi))
        self.dLambda = self.Lambda[1] - self.Lambda[0]
        self.Lambda = self.Lambda[:(1 + self.X_gridDIM // 2)][np.newaxis, :]
        self.P = np.linspace(-self.P_amplitude, self.P_amplitude - self.dP, self.P_gridDIM)[:, np.newaxis]
        self.Theta = np.fft.fftfreq(self.P_gridDIM, self.dP / (2 * np.pi))
        self.dTheta = self.Theta[1] - self.Theta[0]
        self.Theta = self.Theta[:(1 + self.P_gridDIM // 2)][:, np.newaxis]
        self.dXdP = self.dX * self.dP

        self.precision = getattr(self, "precision", "fp64")
        self._prec = precision_code(self.precision)
        self.phase_time_shift = float(getattr(self, "phase_time_shift", self._default_shift()))
        self.device = int(getattr(self, "device", 0))
        self.max_thread_block = getattr(self, "max_thread_block", 512)

        # C constants, wigner_moyal_cuda_1d.py:151-166 (+ consts and functions :163-166, :244-248)
        self.hip_consts = consts_source(kwargs, exclude=self._not_consts + self._extra_not_consts())
        self.hip_consts += getattr(self, "consts", "") or ""
        self.hip_consts += getattr(self, "functions", "") or ""
        self.cuda_consts = self.hip_consts

        # Absorbing boundaries, wigner_moyal_cuda_1d.py:174-182; FFTW alpha :135-156
        if not hasattr(self, "abs_boundary_x_theta"):
            self.abs_boundary_x_theta = self._alpha_mask() if hasattr(self, "alpha") else "1."
        if not hasattr(self, "abs_boundary_lambda_p"):
            self.abs_boundary_lambda_p = "1."
        self._mask_xp = getattr(self, "abs_boundary_x_p", None)
        if isinstance(self._mask_xp, str) is False:
            self._mask_xp = None

        self.hamiltonian = self.K + ' + ' + self.V

        self._handle = None
        tuned = tuning.entry("wigner", self.X_gridDIM, self.P_gridDIM, self._prec) or {}
        tiles = getattr(self, "tiles", None) or (tuple(tuned["tiles"]) if "tiles" in tuned else None)
        # the tuned V-factor policy was measured with a time-independent polynomial V; a V
        # that depends on t keeps the default (its table of V0 when it splits)
        static_v = not _lib.uses_token(self.V, "t")
        v_table = int(getattr(self, "v_table", tuned.get("v_table", 0) if static_v else 0))
        desc = _lib.wigner_desc(self.X_gridDIM, self.X_amplitude, self.P_gridDIM, self.P_amplitude, self.dt,
                                self.phase_time_shift, self._prec, self.device, tiles,
                                int(getattr(self, "phase_tables", 0)), int(getattr(self, "step_passes", 0)),
                                v_table)
        ex = _lib.wigner_exprs(
            V=self.V, K=self.K, consts=self.hip_consts,
            diff_V=getattr(self, "diff_V", None) if isinstance(getattr(self, "diff_V", None), str) else None,
            diff_K=getattr(self, "diff_K", None) if isinstance(getattr(self, "diff_K", None), str) else None,
            abs_boundary_x_theta=self.abs_boundary_x_theta, abs_boundary_lambda_p=self.abs_boundary_lambda_p,
            abs_boundary_x_p=self._mask_xp)
        self._desc, self._exprs = desc, ex
        h = ctypes.c_void_p()
        _lib.call("fw_wigner_create", ctypes.byref(desc), ctypes.byref(ex), ctypes.byref(h))
        self._handle = h

        self.wignerfunction = DeviceArray(self, (self.P_gridDIM, self.X_gridDIM),
                                          np.float32 if self._prec else np.float64,
                                          self._get_host_state, self._state_ptr)

        # Ehrenfest lists, wigner_moyal_cuda_1d.py:350-374
        try:
            self.diff_K
            self.diff_V
            self.X_average = []
            self.P_average = []
            self.X_average_RHS = []
            self.P_average_RHS = []
            self.hamiltonian_average = []
            self.purity_series = []
            self.isEhrenfest = True
        except AttributeError:
            self.isEhrenfest = False

    # -- configuration hooks for subclasses ----------------------------------------
    def _default_shift(self):
        return 0.5

    def _extra_not_consts(self):
        return ()

    def _alpha_mask(self):
        """FFTW absorbing boundary |sin(pi(x+ - xmin)/L) sin(pi(x- - xmin)/L)|^|alpha dt|
        (wigner_moyal_fftw_1d.py:135-156) as an x-theta mask expression."""
        th = np.fft.fftfreq(self.P_gridDIM, self.dP / (2 * np.pi))[:(1 + self.P_gridDIM // 2)][:, None]
        xm, xp = self.X - 0.5 * th, self.X + 0.5 * th
        xmin = min(xm.min(), xp.min())
        xmax = max(xm.max(), xp.max())
        L = xmax - xmin
        return ("pow(fabs(sin(M_PI * (X_plus - (%s)) / (%s)) * sin(M_PI * (X_minus - (%s)) / (%s))), %s)"
                % (c_double(xmin), c_double(L), c_double(xmin), c_double(L), c_double(abs(self.alpha * self.dt))))

    # -- device state ------------------------------------------------------------------
    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.load().fw_wigner_destroy(h)
            except Exception:
                pass
            self._handle = None

    def _get_host_state(self):
        out = np.empty((self.P_gridDIM, self.X_gridDIM))
        _lib.call("fw_wigner_get_state", self._handle, _lib.dptr(out))
        return out

    def _state_ptr(self):
        p, es = ctypes.c_void_p(), ctypes.c_int()
        _lib.call("fw_wigner_state_ptr", self._handle, ctypes.byref(p), ctypes.byref(es))
        return p.value

    def _moments(self, t=None):
        s = np.empty(_lib.FW_NSER)
        _lib.call("fw_wigner_moments", self._handle, float(self.t if t is None else t), _lib.dptr(s))
        return s

    # -- API ---------------------------------------------------------------------------
    def set_wignerfunction(self, new_wigner_func):
        """wigner_moyal_cuda_1d.py:454-494: ndarray / DeviceArray, C-expression
        string of (X, P), python function F(self, X, P) or a float; then
        W /= sum(W) dXdP."""
        if isinstance(new_wigner_func, (np.ndarray, DeviceArray)):
            assert new_wigner_func.shape == (self.P.size, self.X.size), \
                "The grid sizes does not match with the Wigner function"
            w = new_wigner_func.get() if isinstance(new_wigner_func, DeviceArray) else new_wigner_func
            assert np.isrealobj(w), "Supplied Wigner function must be real"
            w = np.ascontiguousarray(w, dtype=np.float64)
            _lib.call("fw_wigner_set_state", self._handle, _lib.dptr(w))
        elif callable(new_wigner_func) and not isinstance(new_wigner_func, str):
            w = np.ascontiguousarray(np.broadcast_to(new_wigner_func(self, self.X, self.P),
                                                     (self.P.size, self.X.size)), dtype=np.float64)
            _lib.call("fw_wigner_set_state", self._handle, _lib.dptr(w))
        elif isinstance(new_wigner_func, str):
            _lib.call("fw_wigner_set_state_expr", self._handle, new_wigner_func.encode())
        elif isinstance(new_wigner_func, float):
            _lib.call("fw_wigner_fill", self._handle, float(new_wigner_func))
        else:
            raise NotImplementedError("new_wigner_func must be either function or numpy.array")
        norm = ctypes.c_double()
        _lib.call("fw_wigner_normalize", self._handle, ctypes.byref(norm))
        return self

    def _times(self, steps):
        """t after each of the steps, accumulated as the reference does, t += dt
        (wigner_moyal_cuda_1d.py:529): np.add.accumulate adds in order, so the
        values are those of the sequential loop, bit for bit."""
        a = np.empty(steps + 1)
        a[0] = self.t
        a[1:] = self.dt
        times = np.add.accumulate(a)[1:]
        self.t = float(times[-1])
        return times

    def _record(self, series):
        """Append the Ehrenfest observables of each step (get_Ehrenfest,
        wigner_moyal_cuda_1d.py:542-562) computed from the in-kernel sums."""
        if not self.isEhrenfest:
            return
        S0 = series[:, 0]
        q = series[:, (1, 5, 2, 7)] / S0[:, np.newaxis]
        self.X_average.extend(q[:, 0].tolist())
        self.X_average_RHS.extend(q[:, 1].tolist())
        self.P_average.extend(q[:, 2].tolist())
        self.P_average_RHS.extend((-series[:, 6] / S0).tolist())
        self.hamiltonian_average.extend(q[:, 3].tolist())
        self.purity_series.extend((2. * np.pi * series[:, 8] / (S0 * S0 * self.dXdP)).tolist())

    def propagate(self, steps=1):
        """wigner_moyal_cuda_1d.py:521-540: ``steps`` x (t += dt, Strang step,
        W /= sum(W) dXdP, Ehrenfest observables), with no host synchronisation
        between steps; returns ``self.wignerfunction``."""
        steps = int(steps)
        if steps <= 0:
            return self.wignerfunction
        times = self._times(steps)
        series = np.empty((steps, _lib.FW_NSER))
        _lib.call("fw_wigner_propagate", self._handle, _lib.dptr(times), steps, _lib.dptr(series))
        self.last_series = series
        self._record(series)
        return self.wignerfunction

    def single_step_propagation(self):
        """One Strang step at the current t, t not advanced, no Ehrenfest record
        (wigner_moyal_cuda_1d.py:496-519; with the dissipator
        wigner_random_collision_model_fftw_1d.py:56-75). As in the reference the
        returned state is not normalised; a later ``propagate`` takes it as it is."""
        _lib.call("fw_wigner_single_step", self._handle, float(self.t), None)
        return self.wignerfunction

    def get_average(self, observable):
        """sum W * observable(X, P, t) dXdP (wigner_moyal_cuda_1d.py:378-396); the
        kernel is compiled once per string."""
        out = ctypes.c_double()
        _lib.call("fw_wigner_get_average", self._handle, observable.encode(), float(self.t), ctypes.byref(out))
        return out.value

    def get_purity(self):
        """2 pi sum W^2 dXdP (wigner_moyal_cuda_1d.py:398-403)."""
        s = self._moments()
        return 2. * np.pi * s[8] * self.dXdP

    def get_sigma_x_sigma_p(self):
        """sqrt((<x^2>-<x>^2)(<p^2>-<p>^2)) (wigner_moyal_cuda_1d.py:405-415)."""
        s = self._moments()
        S0 = s[0]
        x, p = s[1] / S0, s[2] / S0
        return np.sqrt((s[3] / S0 - x * x) * (s[4] / S0 - p * p))

    def get_Ehrenfest(self, t):
        """Append the observables of the current state at time t
        (wigner_moyal_cuda_1d.py:542-562)."""
        if self.isEhrenfest:
            self._record(self._moments(t)[np.newaxis, :])

    @classmethod
    def print_memory_info(cls):
        pass
