"""Evaluate the C-expression strings of the propagators with numpy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference splices strings such as ``V="0.5*X*X+0.1*X*X*X*X"`` into CUDA
kernels (wigner_moyal_cuda_1d.py:564-583). The oracle evaluates the same string
with numpy so a single expression drives both the HIP kernel and the checker.
Only the common C math subset is supported (exp, sin, cos, tan, sqrt, pow, fabs,
abs, log, tanh, cosh, sinh, atan, M_PI); C integer division is not emulated, so
test expressions use floating-point literals.
"""
import numpy as np

_NS = {
    "exp": np.exp, "sin": np.sin, "cos": np.cos, "tan": np.tan, "sqrt": np.sqrt,
    "pow": np.power, "fabs": np.abs, "abs": np.abs, "log": np.log, "tanh": np.tanh,
    "cosh": np.cosh, "sinh": np.sinh, "atan": np.arctan, "atan2": np.arctan2,
    "M_PI": np.pi, "fmin": np.minimum, "fmax": np.maximum,
}


def make(expr, consts=None, args=("X", "t"), derived=None):
    """Return a numpy callable f(*args) evaluating the C expression ``expr``.

    ``consts`` maps extra names (kwargs the reference turns into C constants,
    wigner_moyal_cuda_1d.py:151-166) to values or to python callables standing
    in for ``functions`` (e.g. ``E(t)``). ``derived(*vals)`` returns further names
    the kernel defines from its arguments (e.g. X_minus, X_plus of the x-theta
    mask, wigner_moyal_cuda_1d.py:640-641).
    """
    if expr is None:
        return None
    code = compile(expr, "<c-expr>", "eval")
    base = dict(_NS)
    if consts:
        base.update(consts)

    def f(*vals):
        ns = dict(base)
        ns.update(zip(args, vals))
        if derived is not None:
            ns.update(derived(*vals))
        out = eval(code, {"__builtins__": {}}, ns)
        # broadcast scalars (e.g. "1.") to the shape of the inputs
        shape = np.broadcast(*[np.asarray(v) for v in vals]).shape
        return np.broadcast_to(np.asarray(out, dtype=np.result_type(out, np.float64)), shape).copy()

    return f


def x_theta_names(X, Theta, t):
    return {"X_minus": X - 0.5 * Theta, "X_plus": X + 0.5 * Theta}
