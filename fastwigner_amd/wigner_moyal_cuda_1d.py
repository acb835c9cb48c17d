"""Drop-in module name of the reference's wigner_moyal_cuda_1d.py."""
from .wigner_moyal import WignerMoyalHIP1D

WignerMoyalCUDA1D = WignerMoyalHIP1D

__all__ = ["WignerMoyalCUDA1D", "WignerMoyalHIP1D"]
