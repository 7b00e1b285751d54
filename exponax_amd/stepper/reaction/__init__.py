"""Reaction-diffusion steppers (``exponax/stepper/reaction/``) whose reaction
is a polynomial of the state. Cahn-Hilliard, Gray-Scott and
Belousov-Zhabotinsky need channel-coupled or post-filtered reactions that
these kernels do not evaluate."""
from .._reaction import AllenCahn, FisherKPP, SwiftHohenberg

__all__ = ["AllenCahn", "FisherKPP", "SwiftHohenberg"]
