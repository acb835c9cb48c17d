"""Random-collision (decoherence) model of open-system dynamics on MI355X.

    dW/dt = {{H, W}} + gamma (W_0 - W)

``WignerRandomCollisionHIP1D`` restates ``WignerRandomCollisionModelFFTW1D``
(wigner_random_collision_model_fftw_1d.py:6-94), which exists in the reference
only on the CPU, with the C-expression kwargs of the GPU classes:
w = exp(-gamma dt/2) (:32); W_0 is the Gibbs state of the same Hamiltonian from
the Bloch propagator, scaled by (1 - w) (:35-38); each step is
dissipator -> Moyal step -> dissipator (:56-76), with the dissipator
W <- w W + (1-w) W_0 fused into the first and last passes of the step; the
Ehrenfest right-hand sides are amended by gamma(<.>_Gibbs - <.>) (:78-94).

The reference builds W_0 with WignerBlochFFTW1D, whose exponents are twice the
CUDA ones (the Gibbs state at kT/2); ``bloch_convention`` defaults to "fftw"
to reproduce it, "cuda" gives the Gibbs state at kT. The phase time shift
defaults to 0 (the FFTW class evaluates V, K at t_{n+1}).
"""
import numpy as np

from . import _lib
from .wigner_bloch import WignerBlochHIP1D
from .wigner_moyal import WignerMoyalHIP1D


class WignerRandomCollisionHIP1D(WignerMoyalHIP1D):
    """Second-order split operator for the random collision model."""

    def __init__(self, **kwargs):
        kwargs = dict(kwargs)
        conv = kwargs.pop("bloch_convention", "fftw")
        WignerMoyalHIP1D.__init__(self, **kwargs)
        try:
            self.gamma
        except AttributeError:
            raise AttributeError("Collision rate (gamma) was not specified")
        try:
            self.kT
        except AttributeError:
            raise AttributeError("Temperature (kT) was not specified")
        self.w_coeff = np.exp(-self.gamma * 0.5 * self.dt)

        bkw = {k: v for k, v in kwargs.items() if k not in ("phase_time_shift",)}
        bkw.setdefault("t_initial", getattr(self, "t", 0.))
        bloch = WignerBlochHIP1D(bloch_convention=conv, **bkw)
        self.gibbs_state = bloch.get_gibbs_state().get()
        del bloch
        self.scaled_gibbs_state = self.gibbs_state * (1. - self.w_coeff)
        w0s = np.ascontiguousarray(self.scaled_gibbs_state, dtype=np.float64)
        _lib.call("fw_wigner_set_dissipator", self._handle, _lib.dptr(w0s), float(self.w_coeff))

        if self.isEhrenfest:
            # marginals of the Gibbs state (:40-54)
            cx = self.gibbs_state.sum(axis=0)
            cp = self.gibbs_state.sum(axis=1)
            self.X_average_gibbs = np.dot(cx / cx.sum(), self.X.reshape(-1))
            self.P_average_gibbs = np.dot(cp / cp.sum(), self.P.reshape(-1))

    def _default_shift(self):
        return 0.0

    def _extra_not_consts(self):
        return ("bloch_convention",)

    def _record(self, series):
        if not self.isEhrenfest:
            return
        n0 = len(self.X_average)
        WignerMoyalHIP1D._record(self, series)
        for i in range(n0, len(self.X_average)):
            self.X_average_RHS[i] += self.gamma * (self.X_average_gibbs - self.X_average[i])
            self.P_average_RHS[i] += self.gamma * (self.P_average_gibbs - self.P_average[i])


WignerRandomCollisionModel1D = WignerRandomCollisionHIP1D
