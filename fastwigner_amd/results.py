"""Result files of long propagations and the HHG spectrum of their dipole series.

The reference saves a run into an HDF5 file with h5py:
  * group ``settings``: the numeric keyword arguments of the propagator (values
    that h5py cannot store are skipped) plus ``X_wigner`` and ``P_wigner``
    (strong_field_physics_cuda_1d.py:110-122);
  * group ``frames/<k>`` per saved frame: ``wigner``, ``t`` and ``prob`` (the
    diagonal of rho) (:227-237);
  * group ``ehrenfest``: the per-step series ``X_average``, ``P_average``,
    ``X_average_RHS``, ``P_average_RHS``, ``hamiltonian_average``,
    ``wigner_time`` (:348-354);
and its viewer reads the frames back in increasing k (file_viwer.py:37-67,153-165).

h5py is not a dependency here. ``ResultsWriter`` keeps the same groups and names
as keys ``"settings/<name>"``, ``"frames/<k>/<name>"``, ``"ehrenfest/<name>"`` of
one NumPy ``.npz`` archive, written as each frame arrives (a frame is copied from
the device once and not held in host memory afterwards), so ``numpy.load`` opens
it without pickles. ``load_results`` returns the groups as dictionaries.

``hhg_spectrum`` is the viewer's high-harmonic spectrum (file_viwer.py:237-260):
the squared modulus of the FFT of the Blackman-windowed series, shifted to
centred frequencies in units of the laser frequency and normalised to its
maximum.
"""
import numbers
import os
import zipfile

import numpy as np

EHRENFEST_SERIES = ("X_average", "P_average", "X_average_RHS", "P_average_RHS", "hamiltonian_average",
                    "wigner_time", "purity_series")


def _settable(value):
    """The values the reference's h5py settings group accepts (others raise
    TypeError there and are skipped, strong_field_physics_cuda_1d.py:113-117)."""
    if isinstance(value, (bool, numbers.Number, np.number)):
        return True
    if isinstance(value, str):
        return True
    if isinstance(value, np.ndarray) and value.dtype != object:
        return True
    return False


class ResultsWriter:
    """Writes settings, frames and Ehrenfest series of a run into ``path`` (.npz).

    Use as a context manager, or call ``close()``; the archive is complete only
    after closing. Frames are numbered 0, 1, ... in the order of ``add_frame``.
    """

    def __init__(self, path, settings=None, propagator=None):
        self.path = path
        self._tmp = path + ".part"
        self._zip = zipfile.ZipFile(self._tmp, "w", compression=zipfile.ZIP_STORED, allowZip64=True)
        self._frames = 0
        self._closed = False
        items = dict(settings or {})
        if propagator is not None:
            for name in ("X_wigner", "P_wigner"):
                if hasattr(propagator, name):
                    items.setdefault(name, np.asarray(getattr(propagator, name)))
        for key, val in items.items():
            if _settable(val):
                self._put("settings/%s" % key, np.asarray(val))

    def _put(self, key, arr):
        arr = np.asarray(arr)
        if arr.dtype == object:
            raise TypeError("%s: object arrays are not stored" % key)
        if not arr.flags.c_contiguous:
            arr = np.ascontiguousarray(arr)
        with self._zip.open(key + ".npy", "w", force_zip64=True) as f:
            np.lib.format.write_array(f, arr, allow_pickle=False)

    def add_frame(self, wigner, t, prob=None):
        """One frame: the Wigner function (a host array, or a device array with
        ``.get()`` as the propagators return), its time and optionally the
        probability density (the diagonal of rho). Returns the frame number."""
        if self._closed:
            raise ValueError("the results file is closed")
        k = self._frames
        w = wigner.get() if hasattr(wigner, "get") and not isinstance(wigner, np.ndarray) else wigner
        self._put("frames/%d/wigner" % k, np.asarray(w))
        self._put("frames/%d/t" % k, np.asarray(float(t)))
        if prob is not None:
            self._put("frames/%d/prob" % k, np.asarray(prob))
        self._frames += 1
        return k

    def add_propagator_frame(self, propagator):
        """Frame from a propagator: its current Wigner function (``get_wignerfunction()``
        for the density-matrix classes) and time; ``prob`` is the diagonal of rho
        when the propagator has one (strong_field_physics_cuda_1d.py:230-237)."""
        if hasattr(propagator, "rho"):
            rho = propagator.rho.get() if hasattr(propagator.rho, "get") else np.asarray(propagator.rho)
            return self.add_frame(propagator.get_wignerfunction(), propagator.t, np.diagonal(rho).copy())
        return self.add_frame(propagator.wignerfunction, propagator.t)

    def set_ehrenfest(self, propagator=None, **series):
        """The per-step series of ``propagator`` (those of EHRENFEST_SERIES it has)
        and any extra named series."""
        out = {}
        if propagator is not None:
            for name in EHRENFEST_SERIES:
                val = getattr(propagator, name, None)
                if val is not None and len(val):
                    out[name] = np.asarray(val, dtype=np.float64)
        out.update({k: np.asarray(v) for k, v in series.items()})
        for name, arr in out.items():
            self._put("ehrenfest/%s" % name, arr)

    def close(self):
        if not self._closed:
            self._zip.close()
            os.replace(self._tmp, self.path)
            self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc, tb):
        if exc_type is None:
            self.close()
        else:
            self._zip.close()
            os.remove(self._tmp)
            self._closed = True
        return False


def load_results(path):
    """{"settings": {...}, "frames": [{"wigner", "t", "prob"?}, ...] (in frame
    order, as file_viwer.py:61-63 sorts them), "ehrenfest": {...}}."""
    out = {"settings": {}, "frames": {}, "ehrenfest": {}}
    with np.load(path, allow_pickle=False) as z:
        for key in z.files:
            parts = key.split("/")
            if parts[0] == "frames":
                out["frames"].setdefault(int(parts[1]), {})[parts[2]] = z[key]
            elif parts[0] in ("settings", "ehrenfest"):
                val = z[key]
                out[parts[0]][parts[1]] = val[()] if val.ndim == 0 else val
    frames = []
    for k in sorted(out["frames"]):
        f = out["frames"][k]
        f["t"] = float(f["t"])
        frames.append(f)
    out["frames"] = frames
    return out


def blackman_window(N):
    """Symmetric Blackman window of N points (scipy.signal.blackman, the
    window of file_viwer.py:252): 0.42 - 0.5 cos(2 pi n/(N-1)) + 0.08 cos(4 pi n/(N-1))."""
    n = np.arange(N)
    return 0.42 - 0.5 * np.cos(2.0 * np.pi * n / (N - 1)) + 0.08 * np.cos(4.0 * np.pi * n / (N - 1))


def hhg_spectrum(series, dt, omega):
    """(frequencies / omega, spectrum) of file_viwer.py:237-260: spectrum =
    |fftshift(fft(blackman(N) J))|^2 / max, frequencies fftshift(fftfreq(N,
    dt / 2 pi)) / omega. J is the dipole series (the reference plots <p>; the
    dipole acceleration P_average_RHS is the other usual choice)."""
    J = np.asarray(series, dtype=np.float64)
    N = J.size
    if N < 2:
        raise ValueError("the series needs at least two samples")
    spectrum = np.abs(np.fft.fftshift(np.fft.fft(blackman_window(N) * J))) ** 2
    omegas = np.fft.fftshift(np.fft.fftfreq(N, dt / (2.0 * np.pi))) / omega
    peak = spectrum.max()
    return omegas, (spectrum / peak if peak > 0 else spectrum)
