"""Drop-in module name of the reference's rho_bloch_cuda_1d.py."""
from .rho_bloch import RhoBlochHIP1D
from .rho_vneumann import RhoVNeumannHIP1D

RhoBlochCUDA1D = RhoBlochHIP1D
RhoVNeumannCUDA1D = RhoVNeumannHIP1D

__all__ = ["RhoBlochCUDA1D", "RhoBlochHIP1D", "RhoVNeumannCUDA1D"]
