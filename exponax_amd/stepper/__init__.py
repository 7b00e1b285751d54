"""Concrete steppers with exponax's constructor signatures."""
from ._burgers import Burgers
from ._kuramoto_sivashinsky import KuramotoSivashinsky, KuramotoSivashinskyConservative
from ._navier_stokes import KolmogorovFlowVorticity, NavierStokesVorticity

__all__ = ["Burgers", "KuramotoSivashinsky", "KuramotoSivashinskyConservative", "KolmogorovFlowVorticity",
           "NavierStokesVorticity"]
