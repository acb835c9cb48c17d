"""Batched propagation of many independent Wigner functions on one MI355X.

The reference advances one grid per Python object (WignerMoyalCUDA1D,
wigner_moyal_cuda_1d.py:10-731; the random-collision model
wigner_random_collision_model_fftw_1d.py:16-94 with its Gibbs state from
WignerBlochFFTW1D, wigner_bloch_fftw_1d.py:27-109). ``WignerBatchHIP1D`` holds
hundreds of such propagators with different square grids, potential
parameters, collision rates and temperatures, and advances all of them with
one launch per pass (mode 1, the default) or one persistent launch per step
(mode 0), replayed from a hipGraph per 64 steps.

Each instance is exactly the single-grid step: three fused passes (V columns,
K rows, V columns), the random-collision dissipator W <- w W + (1-w) W_Gibbs
around it, renormalisation and the nine per-step sums. Expressions are C
strings as in the single-grid classes; per-instance parameters are named in
``param_names`` and bound per instance (e.g. V = "a2*X*X + a3*X*X*X +
a4*X*X*X*X" with param_names = ("a2", "a3", "a4")).
"""
import ctypes

import numpy as np

from . import _lib, tuning
from ._common import precision_code

SIZES = (64, 128, 256, 512, 1024)


def dbeta_num_beta_steps(kT, dbeta):
    """Step count n = round(1/(kT dbeta)) and the re-set dbeta = 1/(kT n)
    (wigner_bloch_cuda_1d.py:104-127, wigner_bloch_fftw_1d.py:35-42)."""
    n = 1.0 / (kT * dbeta)
    if round(n) != n:
        n = round(n)
        dbeta = 1.0 / (kT * n)
    return dbeta, int(n)


class WignerBatchHIP1D:
    """Many independent Moyal / random-collision propagators in one launch per step.

    instances: list of dicts with keys
        X_gridDIM      N (the grid is N x N, N in 64..1024, a power of two)
        X_amplitude, P_amplitude
        params         values of ``param_names`` (at most 8)
        gamma          collision rate (optional, default 0: plain Moyal)
        kT             temperature of the Gibbs state W_0 (needed if gamma > 0)
    """

    def __init__(self, instances, V, K, dt, param_names=(), diff_V=None, diff_K=None, consts=None, t=0.,
                 precision="fp32", mode=1, phase_time_shift=0.0, device=0, e_for_n=None, blocks_per_cu=0):
        self.instances = [dict(d) for d in instances]
        self.n = len(self.instances)
        self.dt, self.t = float(dt), float(t)
        self.V, self.K, self.diff_V, self.diff_K = V, K, diff_V, diff_K
        self.param_names = tuple(param_names)
        self.precision = precision
        self.mode = int(mode)
        specs = []
        for d in self.instances:
            n = int(d["X_gridDIM"])
            if n not in SIZES:
                raise ValueError("X_gridDIM must be one of %s" % (SIZES,))
            d.setdefault("P_amplitude", d.get("X_amplitude"))
            specs.append((n, d["X_amplitude"], d["P_amplitude"], d.get("params", ())))
        if e_for_n is None:
            tuned = tuning.batch_entry(precision_code(precision))
            e_for_n = tuned["e_for_n"] if tuned else None
        self.e_for_n = e_for_n
        self._desc = _lib.batch_desc(self.n, dt, precision_code(precision), device, self.mode, phase_time_shift,
                                     e_for_n, blocks_per_cu)
        self._inst = _lib.batch_instances(specs)
        self._exprs = _lib.batch_exprs(V, K, ",".join(self.param_names), consts, diff_V, diff_K)
        h = ctypes.c_void_p()
        _lib.call("fw_batch_create", ctypes.byref(self._desc), self._inst, ctypes.byref(self._exprs),
                  ctypes.byref(h))
        self._handle = h
        # grids per instance (wigner_moyal_cuda_1d.py:112-143)
        self.grids = []
        for d in self.instances:
            n, xa, pa = int(d["X_gridDIM"]), float(d["X_amplitude"]), float(d["P_amplitude"])
            dX, dP = 2 * xa / n, 2 * pa / n
            j = np.arange(n)
            self.grids.append(dict(N=n, dX=dX, dP=dP, X=(dX * (j - n / 2))[np.newaxis, :],
                                   P=(dP * (j - n / 2))[:, np.newaxis]))
        pts = ctypes.c_longlong()
        tasks, grid = ctypes.c_int(), ctypes.c_int()
        _lib.call("fw_batch_info", self._handle, ctypes.byref(pts), ctypes.byref(tasks), ctypes.byref(grid))
        self.total_points, self.tasks_per_step, self.grid = pts.value, tasks.value, grid.value
        self.series = np.zeros((0, self.n, _lib.FW_NSER))

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h:
            try:
                _lib.load().fw_batch_destroy(h)
            except Exception:
                pass
            self._handle = None

    # -- state ---------------------------------------------------------------------
    def set_wignerfunction(self, i, W, normalise=True):
        g = self.grids[i]
        W = np.ascontiguousarray(np.broadcast_to(np.asarray(W, dtype=np.float64), (g["N"], g["N"])))
        if normalise:
            W = W / (W.sum() * g["dX"] * g["dP"])
        _lib.call("fw_batch_set_state", self._handle, int(i), _lib.dptr(W))

    def get_wignerfunction(self, i):
        g = self.grids[i]
        out = np.empty((g["N"], g["N"]))
        _lib.call("fw_batch_get_state", self._handle, int(i), _lib.dptr(out))
        return out

    def copy_states_from(self, other):
        """Every instance's W <- that of `other` (same instances and precision), on the device."""
        _lib.call("fw_batch_copy_states", self._handle, other._handle)

    # -- thermal states and the dissipator ----------------------------------------------
    def prepare_random_collision(self, dbeta=0.01, bloch_convention="fftw", t_initial=0.):
        """Gibbs state of every instance with gamma > 0 at its kT by the batched
        Bloch propagator (the reference builds W_0 with WignerBlochFFTW1D,
        wigner_random_collision_model_fftw_1d.py:35-38), then the source
        (1 - w) W_Gibbs with w = exp(-gamma dt/2) (:32). The state W is left
        holding the Gibbs states: set the initial states afterwards."""
        db = np.zeros(self.n)
        steps = np.zeros(self.n, dtype=np.int32)
        w = np.ones(self.n)
        for i, d in enumerate(self.instances):
            gamma = float(d.get("gamma", 0.0))
            if gamma > 0:
                db[i], steps[i] = dbeta_num_beta_steps(float(d["kT"]), dbeta)
                w[i] = np.exp(-gamma * 0.5 * self.dt)
        uni = [np.full((g["N"], g["N"]), 1.0) for g in self.grids]
        for i, u in enumerate(uni):
            self.set_wignerfunction(i, u)
        conv = 1 if bloch_convention == "fftw" else 0
        _lib.call("fw_batch_bloch_propagate", self._handle, conv, float(t_initial), _lib.dptr(db),
                  steps.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        _lib.call("fw_batch_gibbs_dissipators", self._handle, _lib.dptr(w))
        self.bloch_steps, self.bloch_dbeta, self.w_coeff = steps, db, w

    def set_dissipator(self, i, W0s, w_coeff):
        W0s = np.ascontiguousarray(W0s, dtype=np.float64)
        _lib.call("fw_batch_set_dissipator", self._handle, int(i), _lib.dptr(W0s), float(w_coeff))

    # -- propagation -----------------------------------------------------------------
    def propagate(self, steps=1, record=True):
        """`steps` steps of every instance; the nine sums of every step of every
        instance are appended to ``series`` (steps x instances x 9)."""
        steps = int(steps)
        if steps <= 0:
            return
        times = self.t + self.dt * np.arange(1, steps + 1)
        ser = np.empty((steps, self.n, _lib.FW_NSER)) if record else None
        _lib.call("fw_batch_propagate", self._handle, _lib.dptr(times), steps,
                  _lib.dptr(ser) if record else None)
        self.t = float(times[-1])
        if record:
            self.series = np.concatenate([self.series, ser], axis=0)

    @property
    def passes_per_step(self):
        """Launches per real-time step (1 persistent, 3, or 2 for the fused two-pass steps)."""
        n = ctypes.c_int()
        _lib.call("fw_batch_passes", self._handle, ctypes.byref(n))
        return n.value

    def step_time_us(self, steps=640):
        """Device time per step from graph replays between HIP events (advances the state)."""
        us = ctypes.c_double()
        _lib.call("fw_batch_step_time", self._handle, int(steps), ctypes.byref(us))
        self.t += self.dt * (steps // 64) * 64
        return us.value

    # -- observables from the per-step sums (get_Ehrenfest, wigner_moyal_cuda_1d.py:542-562) --
    def observables(self, i):
        s = self.series[:, i, :]
        S0 = s[:, 0]
        g = self.grids[i]
        return dict(X_average=s[:, 1] / S0, P_average=s[:, 2] / S0, X_average_RHS=s[:, 5] / S0,
                    P_average_RHS=-s[:, 6] / S0, hamiltonian_average=s[:, 7] / S0,
                    purity=2 * np.pi * s[:, 8] / (S0 * S0 * g["dX"] * g["dP"]))


QUARTIC_V = "a2*X*X + a3*X*X*X + a4*X*X*X*X"
QUARTIC_DV = "2.*a2*X + 3.*a3*X*X + 4.*a4*X*X*X"


def random_quartic_instances(count=256, seed=4, sizes=(128, 256, 512, 1024), amplitude=8.):
    """Instances of the batched random-collision workload: grids drawn
    uniformly from `sizes`; V = a2 x^2 + a3 x^3 + a4 x^4 with a2 ~ U[0.2, 1],
    a3 ~ U[-0.1, 0.1], a4 ~ U[0, 0.2]; Gaussian initial states of unit width
    centred at x0, p0 ~ U[-2, 2]; gamma ~ U[0, 0.1]; kT ~ U[0.1, 1];
    x, p in [-amplitude, amplitude). Returns (instances, initial states)."""
    rng = np.random.default_rng(seed)
    out, w0 = [], []
    for _ in range(count):
        n = int(rng.choice(sizes))
        a2, a3, a4 = rng.uniform(0.2, 1.0), rng.uniform(-0.1, 0.1), rng.uniform(0.0, 0.2)
        x0, p0 = rng.uniform(-2, 2), rng.uniform(-2, 2)
        gamma, kT = rng.uniform(0.0, 0.1), rng.uniform(0.1, 1.0)
        out.append(dict(X_gridDIM=n, X_amplitude=amplitude, P_amplitude=amplitude, params=(a2, a3, a4),
                        gamma=gamma, kT=kT, x0=x0, p0=p0))
        w0.append((x0, p0))
    return out, w0


def gaussian_state(n, amplitude, x0, p0):
    dX = 2 * amplitude / n
    x = dX * (np.arange(n) - n / 2)
    return np.exp(-(x[np.newaxis, :] - x0) ** 2 - (x[:, np.newaxis] - p0) ** 2) / np.pi
