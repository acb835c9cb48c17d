"""FastWigner on MI355X: split-operator phase-space propagators (Moyal, Bloch,
random collision, von Neumann, density-matrix Bloch) on hand-written gfx950
kernels behind a C ABI (include/fastwigner.h, libfastwigner.so).

The classes keep the reference's constructor arguments (grid sizes and
amplitudes, dt, V/K/masks as C-expression strings spliced into the kernels),
their propagate() entry point and their observable getters. Importing the
package does not touch the GPU; constructing a propagator loads the native
library and fails loudly if it is missing or no GPU is visible.
"""
__version__ = "0.1.0"

_EXPORTS = {
    "WignerMoyalHIP1D": "wigner_moyal", "WignerMoyalCUDA1D": "wigner_moyal_cuda_1d",
    "WignerBlochHIP1D": "wigner_bloch", "WignerBlochCUDA1D": "wigner_bloch_cuda_1d",
    "WignerRandomCollisionHIP1D": "random_collision",
    "RhoVNeumannHIP1D": "rho_vneumann", "RhoVNeumannCUDA1D": "rho_vneumann_cuda_1d",
    "RhoBlochHIP1D": "rho_bloch", "RhoBlochCUDA1D": "rho_bloch_cuda_1d",
    "WignerBatchHIP1D": "batch",
}


def __getattr__(name):
    mod = _EXPORTS.get(name)
    if mod is None:
        raise AttributeError(name)
    import importlib
    return getattr(importlib.import_module("." + mod, __name__), name)
