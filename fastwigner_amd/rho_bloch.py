"""Bloch (imaginary-time) propagator of the density matrix on MI355X.

``RhoBlochHIP1D`` is the drop-in for ``RhoBlochCUDA1D``
(rho_bloch_cuda_1d.py:9-366): the Gibbs state exp(-H/kT) from rho = I
("double(i == j)", :142) by n = round(1/(kT dbeta)) steps with trace
renormalisation and the purity guard (``get_gibbs_state``, :124-160), the ground
state with adaptive roll-back (``get_ground_state``, :162-222),
``bloch_single_step_propagation`` (:224-241), ``get_V_min`` / ``get_K_min``
(:69-97; K_min is 0 as in the reference, whose fill kernel multiplies a zero
array, :354-365). Factors exp(-dbeta/4 [V(X)+V(X')-V_min]) and
exp(-dbeta/2 [K(P)+K(P')-K_min]) at t_initial (:271-309).
"""
import ctypes

import numpy as np

from . import _lib
from .rho_vneumann import RhoVNeumannHIP1D
from .wigner_bloch import WignerBlochHIP1D


class RhoBlochHIP1D(RhoVNeumannHIP1D):
    """Density matrix of the canonical state exp(-H/kT) by split-operator
    propagation of the Bloch equation."""

    def __init__(self, **kwargs):
        kwargs = dict(kwargs)
        if 't_initial' not in kwargs:
            kwargs.update(t_initial=0.)
            print("Warning: Initial time (t_initial) was not specified. So the default value was used t_initial = 0.")
        self.kT = kwargs.pop('kT', 0.)
        self.dbeta = kwargs.pop('dbeta', 0.01)
        if 'dt' not in kwargs:
            kwargs.update(dt=self.dbeta)
        RhoVNeumannHIP1D.__init__(self, **kwargs)
        mins = (ctypes.c_double * 2)()
        _lib.call("fw_rho_bloch_setup", self._handle, float(self.t_initial), mins)
        self.V_min, self.K_min = mins[0], mins[1]

    def get_V_min(self):
        return self.V_min

    def get_K_min(self):
        return self.K_min

    get_dbeta_num_beta_steps = WignerBlochHIP1D.get_dbeta_num_beta_steps

    def _bloch(self, dbeta, steps, guard):
        series = np.empty((max(int(steps), 1), _lib.FW_RSER))
        done = ctypes.c_int()
        _lib.call("fw_rho_bloch_propagate", self._handle, float(dbeta), int(steps), float(guard),
                  _lib.dptr(series), ctypes.byref(done))
        return done.value, series[:done.value]

    def bloch_single_step_propagation(self, dbeta):
        self._bloch(dbeta, 1, 0.)
        return self.rho

    def get_gibbs_state(self, kT=None, dbeta=None):
        """rho_bloch_cuda_1d.py:124-160."""
        dbeta = self.dbeta if dbeta is None else dbeta
        kT = self.kT if kT is None else kT
        dbeta, num_beta_steps = self.get_dbeta_num_beta_steps(kT, dbeta)
        if not np.isfinite(num_beta_steps):
            raise ValueError("kT must be positive for the Gibbs state; use get_ground_state for kT = 0")
        self.set_rho("double(i == j)")
        done, series = self._bloch(dbeta, num_beta_steps, 1.0)
        self.gibbs_steps = done
        if done < num_beta_steps:
            print("Warning: Gibbs state calculations interupted because purity = %.10f > 1."
                  "Current kT = %.6f" % (self.get_purity(), 1 / (max(done - 1, 1) * dbeta)))
        return self.rho

    def get_ground_state(self, dbeta=None, abs_tol_purity=1e-12):
        """rho_bloch_cuda_1d.py:162-222 (the uncertainty check is commented out there)."""
        previous_energy = current_energy = np.inf
        previous_purity = current_purity = 0.
        dbeta = np.float64(self.dbeta if dbeta is None else dbeta)
        self.set_rho("double(i == j)")
        _lib.call("fw_rho_snapshot", self._handle, 0)
        while current_purity < (1. - abs_tol_purity) and dbeta > 1e-12:
            self._bloch(dbeta, 1, 0.)
            current_purity = self.get_purity()
            current_energy = self.get_average((None, self.K)) + self.get_average((self.V,))
            if current_purity <= 1. and current_energy < previous_energy:
                previous_energy = current_energy
                previous_purity = current_purity
                _lib.call("fw_rho_snapshot", self._handle, 0)
            else:
                _lib.call("fw_rho_snapshot", self._handle, 1)
                dbeta *= 0.5
                current_energy = previous_energy
                current_purity = previous_purity
        return self.rho


RhoBlochCUDA1D = RhoBlochHIP1D
