"""von Neumann propagator of the density matrix on MI355X.

``RhoVNeumannHIP1D`` is the drop-in for ``RhoVNeumannCUDA1D``
(rho_vneumann_cuda_1d.py:15-858): same keyword arguments (``X_gridDIM``,
``X_amplitude``, ``dt``, ``t``, ``V``/``K``/``diff_V``/``diff_K`` strings,
``consts``, ``functions``, ``abs_boundary_x``, ``abs_boundary_p``,
``max_thread_block``), methods (``set_rho``, ``propagate``,
``single_step_propagation``, ``get_Ehrenfest``, ``get_average`` with its tuple of
alternating x / p observables, ``get_purity``, ``get_sigma_x_sigma_p``,
``get_wignerfunction``, ``get_wigner_time``) and attributes (``X``, ``P``,
``dX``, ``dP``, ``X_wigner``, ``P_wigner``, ``wigner_dxdp``, ``rho``, ``t`` and
the Ehrenfest lists).

Per-step Ehrenfest observables come from sums fused into the step kernels
(trace, <x>, <V'>, <V>, purity) and one diagonal-sum pass for the momentum
family (<p>, <K'>, <K>). The reference also records the Wigner time each step,
which needs the user attributes ``wigner_initial`` and ``Ip``
(rho_vneumann_cuda_1d.py:408-412); here it is recorded when both are set.

Additional keyword arguments: ``precision`` ("fp64" or "fp32": complex64
storage and FFTs, FP64 phases and sums), ``phase_time_shift`` (V, K at
t_{n+1} + shift dt; default 0 as the reference's density-matrix kernels),
``device``, ``tiles`` = (row_e, row_p, col_e, col_p).
"""
import ctypes

import numpy as np

from . import _lib, tuning
from ._common import DeviceArray, bind_kwargs, consts_source, precision_code

_GRID_NAMES = ("X_gridDIM", "X_amplitude", "dt", "dX", "dP")


class RhoVNeumannHIP1D:
    """Second-order split-operator propagator for the von Neumann equation for
    rho(x, x', t) with H = K(p, t) + V(x, t), on hand-written gfx950 kernels."""

    _not_consts = ("t",)

    def __init__(self, **kwargs):
        bind_kwargs(self, kwargs)
        for name, msg in (("X_gridDIM", "Coordinate grid size (X_gridDIM) was not specified"),
                          ("X_amplitude", "Coordinate grid range (X_amplitude) was not specified"),
                          ("V", "Potential energy (V) was not specified"),
                          ("K", "Momentum dependence (K) was not specified"),
                          ("dt", "Time-step (dt) was not specified")):
            if not hasattr(self, name):
                raise AttributeError(msg)
        if not hasattr(self, "t"):
            print("Warning: Initial time (t) was not specified, thus it is set to zero.")
            self.t = 0.
        self.t = np.float64(self.t)

        # grids, rho_vneumann_cuda_1d.py:93-110
        self.dX = 2. * self.X_amplitude / self.X_gridDIM
        self.X = np.linspace(-self.X_amplitude, self.X_amplitude - self.dX, self.X_gridDIM)
        self.P = np.fft.fftfreq(self.X_gridDIM, self.dX / (2 * np.pi))
        self.dP = self.P[1] - self.P[0]
        self.P_wigner = self.P / np.sqrt(2.)
        self.X_wigner = self.X / np.sqrt(2.)
        self.wigner_dxdp = (self.X_wigner[1] - self.X_wigner[0]) * (self.P_wigner[1] - self.P_wigner[0])
        self.P_wigner = np.fft.fftshift(self.P_wigner)[:, np.newaxis]
        self.X_wigner = self.X_wigner[np.newaxis, :]

        self.precision = getattr(self, "precision", "fp64")
        self._prec = precision_code(self.precision)
        self.phase_time_shift = float(getattr(self, "phase_time_shift", 0.0))
        self.device = int(getattr(self, "device", 0))
        self.max_thread_block = getattr(self, "max_thread_block", 512)

        # C constants, rho_vneumann_cuda_1d.py:118-133 (+ consts, functions :130-133,188-192)
        self.hip_consts = consts_source(kwargs, exclude=self._not_consts + self._extra_not_consts() + _GRID_NAMES)
        self.hip_consts += getattr(self, "consts", "") or ""
        self.hip_consts += getattr(self, "functions", "") or ""
        self.cuda_consts = self.hip_consts
        mx = getattr(self, "abs_boundary_x", None)
        mp = getattr(self, "abs_boundary_p", None)
        self.abs_boundary_x = mx if isinstance(mx, str) else "1."
        self.abs_boundary_p = mp if isinstance(mp, str) else "1."

        self._handle = None
        tiles = getattr(self, "tiles", None) or tuning.lookup("rho", self.X_gridDIM, self.X_gridDIM, self._prec)
        desc = _lib.rho_desc(self.X_gridDIM, self.X_amplitude, self.dt, self.phase_time_shift, self._prec,
                             self.device, tiles)
        dv = getattr(self, "diff_V", None)
        dk = getattr(self, "diff_K", None)
        ex = _lib.rho_exprs(V=self.V, K=self.K, consts=self.hip_consts,
                            diff_V=dv if isinstance(dv, str) else None, diff_K=dk if isinstance(dk, str) else None,
                            abs_boundary_x=mx if isinstance(mx, str) else None,
                            abs_boundary_p=mp if isinstance(mp, str) else None)
        self._desc, self._exprs = desc, ex
        h = ctypes.c_void_p()
        _lib.call("fw_rho_create", ctypes.byref(desc), ctypes.byref(ex), ctypes.byref(h))
        self._handle = h
        n = self.X_gridDIM
        self.rho = DeviceArray(self, (n, n), np.complex64 if self._prec else np.complex128,
                               self._get_host_state, self._state_ptr)
        self._wigner = DeviceArray(self, (n, n), np.complex64 if self._prec else np.complex128,
                                   self._get_wigner, lambda: None)

        try:
            self.diff_K
            self.diff_V
            self.X_average = []
            self.P_average = []
            self.X_average_RHS = []
            self.P_average_RHS = []
            self.hamiltonian_average = []
            self.wigner_time = []
            self.purity_series = []
            self.isEhrenfest = True
        except AttributeError:
            self.isEhrenfest = False

    def _extra_not_consts(self):
        return ()

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                _lib.load().fw_rho_destroy(h)
            except Exception:
                pass
            self._handle = None

    # -- device state ----------------------------------------------------------------
    def _get_host_state(self):
        n = self.X_gridDIM
        out = np.empty((n, n), np.complex128)
        _lib.call("fw_rho_get_state", self._handle, out.ctypes.data_as(_lib._PD))
        return out

    def _state_ptr(self):
        p, es = ctypes.c_void_p(), ctypes.c_int()
        _lib.call("fw_rho_state_ptr", self._handle, ctypes.byref(p), ctypes.byref(es))
        return p.value

    def _get_wigner(self):
        n = self.X_gridDIM
        out = np.empty((n, n), np.complex128)
        _lib.call("fw_rho_wigner", self._handle, out.ctypes.data_as(_lib._PD))
        return out

    # -- API ----------------------------------------------------------------------------
    def set_rho(self, new_rho):
        """rho_vneumann_cuda_1d.py:301-338: ndarray / DeviceArray, C string of
        (X, X_prime, i, j), python function F(self, X, X_prime) or a float;
        then rho /= Tr(rho) dX."""
        n = self.X_gridDIM
        if isinstance(new_rho, (np.ndarray, DeviceArray)):
            assert new_rho.shape == (n, n), "The grid sizes does not match with the density matrix"
            r = new_rho.get() if isinstance(new_rho, DeviceArray) else new_rho
            r = np.ascontiguousarray(r, dtype=np.complex128)
            _lib.call("fw_rho_set_state", self._handle, r.ctypes.data_as(_lib._PD))
        elif isinstance(new_rho, str):
            _lib.call("fw_rho_set_state_expr", self._handle, new_rho.encode())
        elif isinstance(new_rho, float):
            r = np.full((n, n), new_rho, dtype=np.complex128)
            _lib.call("fw_rho_set_state", self._handle, r.ctypes.data_as(_lib._PD))
        elif callable(new_rho):
            r = np.ascontiguousarray(np.broadcast_to(new_rho(self, self.X[:, np.newaxis], self.X[np.newaxis, :]),
                                                     (n, n)), dtype=np.complex128)
            _lib.call("fw_rho_set_state", self._handle, r.ctypes.data_as(_lib._PD))
        else:
            raise NotImplementedError("new_rho must be either function or numpy.array")
        tr = (ctypes.c_double * 2)()
        _lib.call("fw_rho_normalize", self._handle, tr)
        return self

    def _times(self, steps):
        times = np.empty(steps)
        for s in range(steps):
            self.t += self.dt
            times[s] = self.t
        return times

    def _observables(self, s):
        """Normalised Ehrenfest observables from the raw sums of one step."""
        tr = complex(s[0], s[1])
        c = 1.0 / (tr * self.dX)

        def tr_of(k):
            return (c * complex(s[k], s[k + 1])).real * self.dX

        return dict(X_average=tr_of(2), P_average=tr_of(11), X_average_RHS=tr_of(15), P_average_RHS=-tr_of(6),
                    hamiltonian_average=tr_of(17) + tr_of(8), purity=s[10] * abs(c) ** 2 * self.dX ** 2)

    def _record(self, series):
        """_observables over the rows of `series` at once (numpy, no per-step loop)."""
        s = np.asarray(series)
        if s.size == 0:
            return
        c = 1.0 / ((s[:, 0] + 1j * s[:, 1]) * self.dX)

        def tr_of(k):
            return (c * (s[:, k] + 1j * s[:, k + 1])).real * self.dX

        self.X_average.extend(tr_of(2).tolist())
        self.X_average_RHS.extend(tr_of(15).tolist())
        self.P_average.extend(tr_of(11).tolist())
        self.P_average_RHS.extend((-tr_of(6)).tolist())
        self.hamiltonian_average.extend((tr_of(17) + tr_of(8)).tolist())
        self.purity_series.extend((s[:, 10] * np.abs(c) ** 2 * self.dX ** 2).tolist())

    def _wigner_time_enabled(self):
        return self.isEhrenfest and hasattr(self, "wigner_initial") and hasattr(self, "Ip")

    def propagate(self, steps=1):
        """rho_vneumann_cuda_1d.py:355-374: steps x (t += dt, step, rho /= Tr(rho) dX,
        Ehrenfest observables); returns ``self.rho``. With ``wigner_initial`` and ``Ip``
        set, every step also computes the Wigner function and its Wigner time
        (rho_vneumann_cuda_1d.py:408-412), on the device inside the step graph;
        ``wigner_current`` is then the last step's Wigner function."""
        steps = int(steps)
        if steps <= 0:
            return self.rho
        times = self._times(steps)
        series = np.empty((steps, _lib.FW_RSER))
        mom = 1 if self.isEhrenfest else 0
        if self._wigner_time_enabled():
            self._upload_wigner_initial(self.wigner_initial)
            wt = np.empty(steps)
            _lib.call("fw_rho_propagate_wigner_time", self._handle, _lib.dptr(times), steps, mom, float(self.Ip),
                      _lib.dptr(series), _lib.dptr(wt))
            self.wigner_current = self._wigner
            self.wigner_time.extend(wt.tolist())
        else:
            _lib.call("fw_rho_propagate", self._handle, _lib.dptr(times), steps, mom, _lib.dptr(series))
        self.last_series = series
        if self.isEhrenfest:
            self._record(series)
        return self.rho

    def _upload_wigner_initial(self, w_init):
        """W_init goes to the device once per array object (re-assigning the attribute,
        or passing another array, uploads the new one)."""
        if getattr(self, "_winit_src", None) is w_init:
            return
        w0 = w_init.get() if isinstance(w_init, DeviceArray) else np.asarray(w_init)
        n = self.X_gridDIM
        assert w0.shape == (n, n), "wigner_initial must be N x N on the Wigner grid"
        w0 = np.ascontiguousarray(np.real(w0), dtype=np.float64)
        _lib.call("fw_rho_set_wigner_initial", self._handle, _lib.dptr(w0))
        self._winit_src = w_init

    def single_step_propagation(self):
        """One step at the current t, t not advanced (rho_vneumann_cuda_1d.py:340-353).
        As in the reference the returned rho is not normalised."""
        _lib.call("fw_rho_single_step", self._handle, float(self.t), None)
        return self.rho

    def get_Ehrenfest(self, t):
        if self.isEhrenfest:
            s = np.empty(_lib.FW_RSER)
            _lib.call("fw_rho_moments", self._handle, float(t), _lib.dptr(s))
            self._record([s])

    def get_average(self, observable):
        """Tr[rho F1(x) g1(p) F2(x) ...] dX for observable = ("F1(x)", "g1(p)", ...),
        None skipping a slot (rho_vneumann_cuda_1d.py:431-472)."""
        obs = (ctypes.c_char_p * len(observable))(*[o.encode() if o else None for o in observable])
        out = ctypes.c_double()
        _lib.call("fw_rho_get_average", self._handle, obs, len(observable), float(self.t), ctypes.byref(out))
        return out.value

    def get_purity(self):
        """Tr rho^2 = sum |rho|^2 dX^2 (rho_vneumann_cuda_1d.py:474-488)."""
        s = np.empty(_lib.FW_RSER)
        _lib.call("fw_rho_moments", self._handle, float(self.t), _lib.dptr(s))
        return self._observables(s)["purity"]

    def get_sigma_x_sigma_p(self):
        """rho_vneumann_cuda_1d.py:490-500."""
        return np.sqrt((self.get_average(("X * X",)) - self.get_average(("X",)) ** 2) *
                       (self.get_average((None, "P * P")) - self.get_average((None, "P")) ** 2))

    def get_wignerfunction(self):
        """Wigner function on the (X/sqrt2, P/sqrt2) grid by the Larkin three-shear
        rotation (rho_vneumann_cuda_1d.py:502-592); a complex DeviceArray."""
        _lib.call("fw_rho_wigner", self._handle, None)
        return self._wigner

    def get_wigner_time(self, wigner_current, wigner_init, t):
        """int 0.5 (1 + signbit(H + Ip)) (W - W_init) dxdp, the reference's indicator
        (rho_vneumann_cuda_1d.py:594-620,830-858), over the device's Wigner function:
        wigner_current must be the array get_wignerfunction returned (the device
        computes on its own copy; any other array is refused rather than silently
        ignored). Ip is the attribute (or constructor keyword) ``Ip``."""
        if wigner_current is not self._wigner:
            raise ValueError("get_wigner_time: wigner_current must be the array returned by get_wignerfunction() "
                             "(the Wigner function is kept on the device)")
        self._upload_wigner_initial(wigner_init)
        out = ctypes.c_double()
        _lib.call("fw_rho_wigner_time", self._handle, float(t), float(self.Ip), ctypes.byref(out))
        return out.value

    @classmethod
    def print_memory_info(cls):
        pass


RhoVNeumannCUDA1D = RhoVNeumannHIP1D
