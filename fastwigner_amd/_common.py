"""Shared host helpers of the drop-in propagator classes."""
import numbers
from types import FunctionType, MethodType

import numpy as np

from . import _lib

# Names the native library emits as C constants itself (include/fastwigner.h,
# fw_wigner.cpp make_prelude); the Python side must not emit them again.
GRID_CONST_NAMES = frozenset(
    ["X_gridDIM", "P_gridDIM", "X_amplitude", "P_amplitude", "dt", "dX", "dP", "dLambda", "dTheta"])

# Keyword arguments of this implementation that are not physics constants.
CONTROL_KWARGS = frozenset(["precision", "phase_time_shift", "device", "tiles", "bloch_convention",
                            "phase_tables", "step_passes"])


def bind_kwargs(obj, kwargs):
    """Store kwargs as attributes; python functions become bound methods
    (wigner_moyal_cuda_1d.py:48-55)."""
    for name, value in kwargs.items():
        if isinstance(value, FunctionType):
            setattr(obj, name, MethodType(value, obj))
        else:
            setattr(obj, name, value)


def consts_source(kwargs, exclude=()):
    """int/float kwargs become C constants (wigner_moyal_cuda_1d.py:151-166).

    Doubles are printed with %.17g, which round-trips exactly; the reference's
    %.15e drops one significant digit."""
    lines = []
    for name, value in kwargs.items():
        if name in GRID_CONST_NAMES or name in CONTROL_KWARGS or name in exclude:
            continue
        if isinstance(value, (bool, numbers.Integral)) and not isinstance(value, np.bool_):
            lines.append("    const int %s = %d;\n" % (name, int(value)))
        elif isinstance(value, numbers.Real):
            lines.append("    const double %s = %s;\n" % (name, c_double(float(value))))
    return "".join(lines)


def c_double(v):
    s = "%.17g" % v
    if not any(ch in s for ch in ".eEn"):
        s += ".0"
    return s


def precision_code(p):
    if p in (None, "fp64", "float64", np.float64, 64):
        return 0
    if p in ("fp32", "float32", np.float32, 32):
        return 1
    raise ValueError("precision must be 'fp64' or 'fp32'")


class DeviceArray:
    """Handle to a state held in device memory (the role of pycuda's GPUArray in
    the reference): ``propagate()`` returns it and callers do ``.get()``
    (wigner_moyal_cuda_1d.py:854)."""

    def __init__(self, owner, shape, dtype, getter, ptr_getter):
        self._owner = owner
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self._getter = getter
        self._ptr_getter = ptr_getter

    @property
    def size(self):
        return int(np.prod(self.shape))

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def ptr(self):
        """Device address of the storage."""
        return self._ptr_getter()

    def get(self):
        """Copy to a new host numpy array."""
        return self._getter()

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def __repr__(self):
        return "DeviceArray(shape=%s, dtype=%s)" % (self.shape, self.dtype)
