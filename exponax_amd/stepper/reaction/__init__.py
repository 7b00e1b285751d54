"""Reaction-diffusion steppers (``exponax/stepper/reaction/``)."""
from .._reaction import AllenCahn, CahnHilliard, FisherKPP, SwiftHohenberg

__all__ = ["AllenCahn", "CahnHilliard", "FisherKPP", "SwiftHohenberg"]
