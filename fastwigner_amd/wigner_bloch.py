"""Bloch (imaginary-time) propagator of the Wigner function on MI355X.

``WignerBlochHIP1D`` is the drop-in for ``WignerBlochCUDA1D``
(wigner_bloch_cuda_1d.py:8-445): Gibbs state exp(-H/kT) from the uniform state
by ``n = round(1/(kT dbeta))`` steps with renormalisation and the purity guard
(``get_gibbs_state``, :129-165), the ground state with adaptive roll-back
(``get_ground_state``, :167-227), ``bloch_single_step_propagation`` (:229-254),
``get_dbeta_num_beta_steps`` (:104-127), ``get_V_min``/``get_K_min`` (:72-102).

``bloch_convention``:
    "cuda" (default)  exp(-dbeta/4 [V(x-th/2)+V(x+th/2)-V_min]),
                      exp(-dbeta/2 [K(p+l/2)+K(p-l/2)-K_min]), masks off
                      (wigner_bloch_cuda_1d.py:290-359) -- the Gibbs state at kT
    "fftw"            exp(-dbeta/2 [..]) times the absorbing mask, exp(-dbeta [..])
                      (wigner_bloch_fftw_1d.py:65-100) -- i.e. the Gibbs state at kT/2
"""
import ctypes

import numpy as np

from . import _lib
from .wigner_moyal import WignerMoyalHIP1D

_CONV = {"cuda": 0, "fftw": 1}


class WignerBlochHIP1D(WignerMoyalHIP1D):
    """Wigner function of the canonical state exp(-H/kT) by split-operator
    propagation of the Bloch equation in phase space (D. I. Bondar, A. G. Campos,
    R. Cabrera, H. A. Rabitz, Phys. Rev. E 93, 063304 (2016))."""

    def __init__(self, **kwargs):
        kwargs = dict(kwargs)
        if 't_initial' not in kwargs:
            kwargs.update(t_initial=0.)
            print("Warning: Initial time (t_initial) was not specified. So the default value was used t_initial = 0.")
        # wigner_bloch_cuda_1d.py:31-48: kT and dbeta do not enter the C constants
        self.kT = kwargs.pop('kT', 0.)
        self.dbeta = kwargs.pop('dbeta', 0.01)
        if 'dt' not in kwargs:
            kwargs.update(dt=self.dbeta)
        self.bloch_convention = kwargs.pop('bloch_convention', 'cuda')
        if self.bloch_convention not in _CONV:
            raise ValueError("bloch_convention must be 'cuda' or 'fftw'")
        WignerMoyalHIP1D.__init__(self, **kwargs)
        mins = (ctypes.c_double * 2)()
        _lib.call("fw_wigner_bloch_setup", self._handle, _CONV[self.bloch_convention], float(self.t_initial), mins)
        self.V_min, self.K_min = mins[0], mins[1]

    def get_V_min(self):
        """min over the theta-x grid of V(X-Theta/2)+V(X+Theta/2) at t_initial."""
        return self.V_min

    def get_K_min(self):
        """min over the p-lambda grid of K(P+Lambda/2)+K(P-Lambda/2) at t_initial."""
        return self.K_min

    @classmethod
    def get_dbeta_num_beta_steps(cls, kT, dbeta):
        """wigner_bloch_cuda_1d.py:104-127."""
        if kT > 0:
            num_beta_steps = 1. / (kT * dbeta)
            if round(num_beta_steps) != num_beta_steps:
                num_beta_steps = round(num_beta_steps)
                dbeta = 1. / (kT * num_beta_steps)
            num_beta_steps = int(num_beta_steps)
        else:
            num_beta_steps = np.inf
        return dbeta, num_beta_steps

    def _bloch(self, dbeta, steps, guard):
        series = np.empty((max(int(steps), 1), _lib.FW_NSER))
        done = ctypes.c_int()
        _lib.call("fw_wigner_bloch_propagate", self._handle, float(dbeta), int(steps), float(guard),
                  _lib.dptr(series), ctypes.byref(done))
        return done.value, series[:done.value]

    def bloch_single_step_propagation(self, dbeta):
        """One Bloch step of size dbeta, renormalised (wigner_bloch_cuda_1d.py:229-254)."""
        self._bloch(dbeta, 1, 0.)
        return self.wignerfunction

    def get_gibbs_state(self, kT=None, dbeta=None):
        """wigner_bloch_cuda_1d.py:129-165: uniform state, n Bloch steps, stop
        (with a warning) at the first step whose purity is not below 1. The steps
        run back to back on the device; the guard is evaluated there."""
        dbeta = self.dbeta if dbeta is None else dbeta
        kT = self.kT if kT is None else kT
        dbeta, num_beta_steps = self.get_dbeta_num_beta_steps(kT, dbeta)
        if not np.isfinite(num_beta_steps):
            raise ValueError("kT must be positive for the Gibbs state; use get_ground_state for kT = 0")
        self.set_wignerfunction(1. / (self.P_gridDIM * self.X_gridDIM))
        guard = 1.0 if self.bloch_convention == "cuda" else 0.0
        done, series = self._bloch(dbeta, num_beta_steps, guard)
        self.gibbs_steps = done
        if done < num_beta_steps or (guard and len(series) and not self._purity(series[-1]) < 1.):
            k = max(done - 1, 1)
            print("Warning: Gibbs state calculations interupted because purity = %.10f > 1."
                  "Current kT = %.6f" % (self.get_purity(), 1 / (k * dbeta)))
        return self.wignerfunction

    def _purity(self, s):
        return 2. * np.pi * s[8] / (s[0] * s[0] * self.dXdP)

    def get_ground_state(self, dbeta=None, abs_tol_purity=1e-12):
        """wigner_bloch_cuda_1d.py:167-227: Bloch steps from the uniform state until
        1 - purity < abs_tol_purity; a step that raises the energy, exceeds purity 1
        or violates sigma_x sigma_p >= 1/2 is rolled back and dbeta halved."""
        previous_energy = current_energy = np.inf
        previous_purity = current_purity = 0.
        dbeta = self.dbeta if dbeta is None else dbeta
        dbeta = np.float64(dbeta)
        self.set_wignerfunction(1. / (self.P_gridDIM * self.X_gridDIM))
        _lib.call("fw_wigner_snapshot", self._handle, 0)
        while current_purity < (1. - abs_tol_purity) and dbeta > 1e-12:
            _, series = self._bloch(dbeta, 1, 0.)
            s = series[-1]
            S0 = s[0]
            current_purity = self._purity(s)
            current_energy = s[7] / S0
            x, p = s[1] / S0, s[2] / S0
            sigma = np.sqrt((s[3] / S0 - x * x) * (s[4] / S0 - p * p))
            if current_purity <= 1. and current_energy < previous_energy and sigma >= 0.5:
                previous_energy = current_energy
                previous_purity = current_purity
                _lib.call("fw_wigner_snapshot", self._handle, 0)
            else:
                _lib.call("fw_wigner_snapshot", self._handle, 1)
                dbeta *= 0.5
                current_energy = previous_energy
                current_purity = previous_purity
        return self.wignerfunction


WignerBlochCUDA1D = WignerBlochHIP1D
