"""Drop-in module name of the reference's rho_vneumann_cuda_1d.py."""
from .rho_vneumann import RhoVNeumannHIP1D

RhoVNeumannCUDA1D = RhoVNeumannHIP1D

__all__ = ["RhoVNeumannCUDA1D", "RhoVNeumannHIP1D"]
