"""Direct O(N^2) discrete Fourier transforms used to pin the FFT conventions.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Forward:  X_k = sum_n x_n exp(-2 pi i n k / N)          (FFTW_FORWARD / cufftExec*2Z)
Inverse:  x_n = (1/N) sum_k X_k exp(+2 pi i n k / N)    (FFTW_BACKWARD as normalised by
          pyFFTW, wigner_moyal_fftw_1d.py:341-366; cuFFT's inverse is unnormalised,
          cufft.py:155, and the per-step renormalisation cancels the factor).
The twiddles are evaluated exactly on the integer phase (n*k mod N) so the
matrices carry no accumulated phase error.
"""
import numpy as np


def dft_matrix(n, inverse=False):
    k = np.arange(n)
    nk = np.outer(k, k) % n
    sign = 1.0 if inverse else -1.0
    m = np.exp(sign * 2j * np.pi * nk / n)
    return m / n if inverse else m


def dft(x, axis=-1, inverse=False):
    x = np.moveaxis(np.asarray(x, dtype=np.complex128), axis, -1)
    y = x @ dft_matrix(x.shape[-1], inverse).T
    return np.moveaxis(y, -1, axis)


def rdft(x, axis=-1):
    """Half spectrum (N/2+1 bins) of a real signal, like r2c."""
    n = np.asarray(x).shape[axis]
    y = dft(x, axis=axis)
    return np.take(y, np.arange(n // 2 + 1), axis=axis)


def irdft(y, n, axis=-1):
    """c2r: Hermitian continuation of the half spectrum, imaginary parts of the
    DC and Nyquist bins discarded, normalised inverse DFT, real part."""
    y = np.moveaxis(np.asarray(y, dtype=np.complex128), axis, -1).copy()
    y[..., 0] = y[..., 0].real
    if n % 2 == 0:
        y[..., n // 2] = y[..., n // 2].real
    full = np.zeros(y.shape[:-1] + (n,), dtype=np.complex128)
    full[..., : n // 2 + 1] = y
    for k in range(n // 2 + 1, n):
        full[..., k] = np.conj(y[..., n - k])
    x = dft(full, inverse=True).real
    return np.moveaxis(x, -1, axis)
