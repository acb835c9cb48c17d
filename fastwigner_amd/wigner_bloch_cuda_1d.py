"""Drop-in module name of the reference's wigner_bloch_cuda_1d.py."""
from .wigner_bloch import WignerBlochHIP1D
from .wigner_moyal import WignerMoyalHIP1D

WignerBlochCUDA1D = WignerBlochHIP1D
WignerMoyalCUDA1D = WignerMoyalHIP1D

__all__ = ["WignerBlochCUDA1D", "WignerBlochHIP1D", "WignerMoyalCUDA1D"]
