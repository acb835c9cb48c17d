"""CPU oracle for the FastWigner split-operator propagators.

TEST INFRASTRUCTURE ONLY. This package is the checker, never the product: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it. The product path (``fastwigner_amd``) runs the HIP
kernels and fails loudly when its native library is missing.

Contents
--------
``fw_oracle``  numpy/scipy (pocketfft, complex128) restatement of the reference
               propagators: Moyal, Bloch (both exponent conventions), random
               collision, von Neumann, density-matrix Bloch, Larkin-shear Wigner
               transform. Each function cites the reference file:line it follows.
``dft``        direct O(N^2) DFT matrices, used to pin the FFT conventions.
``cexpr``      evaluates the C-expression strings the propagators take (``V``,
               ``K``, masks) with numpy, so one expression drives both sides.

How the oracle is pinned (tests/test_oracle_*.py)
-------------------------------------------------
* r2c/c2r transforms against the direct DFT (``dft``), including the c2r rule
  that drops the imaginary part at DC and Nyquist.
* Moyal step against an independent full-spectrum DFT-matrix formulation on
  8x8 and 16x16 grids (brute force).
* Harmonic oscillator: the Strang step is exactly the leapfrog map, so a
  Gaussian stays Gaussian with mean/covariance propagated by the 2x2 leapfrog
  matrix (closed form, ~1e-13).
* Bloch: the closed-form harmonic Gibbs state, and the identity
  FFTW(kT, dbeta) == CUDA(kT/2, 2 dbeta) between the two exponent conventions.
* Random collision: gamma=0 reduces to Moyal exactly; the Gibbs state is a
  fixed point of the dissipator.
* von Neumann: trace, hermiticity and purity conservation; the Wigner transform
  of a Gaussian density matrix against its analytic Wigner function; Moyal and
  von Neumann evolutions of the same state agree on <x>, <p>.
"""
