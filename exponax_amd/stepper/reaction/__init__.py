"""Reaction-diffusion steppers (``exponax/stepper/reaction/``)."""
from .._reaction import AllenCahn, BelousovZhabotinsky, CahnHilliard, FisherKPP, GrayScott, SwiftHohenberg

__all__ = ["AllenCahn", "BelousovZhabotinsky", "CahnHilliard", "FisherKPP", "GrayScott", "SwiftHohenberg"]
