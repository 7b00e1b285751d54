"""Concrete steppers with exponax's constructor signatures."""
from . import generic
from ._burgers import Burgers
from ._korteweg_de_vries import KortewegDeVries
from ._kuramoto_sivashinsky import KuramotoSivashinsky, KuramotoSivashinskyConservative
from ._linear import Advection, AdvectionDiffusion, Diffusion, Dispersion, HyperDiffusion
from ._navier_stokes import KolmogorovFlowVelocity, KolmogorovFlowVorticity, NavierStokesVelocity, NavierStokesVorticity
from ._reaction import AllenCahn, FisherKPP, SwiftHohenberg
from . import reaction
from ._wave import Wave

__all__ = ["generic", "reaction", "AllenCahn", "FisherKPP", "SwiftHohenberg", "Advection", "AdvectionDiffusion", "Burgers", "Diffusion", "Dispersion", "HyperDiffusion",
           "KortewegDeVries", "KuramotoSivashinsky", "KuramotoSivashinskyConservative", "KolmogorovFlowVorticity",
           "NavierStokesVorticity", "NavierStokesVelocity", "KolmogorovFlowVelocity", "Wave"]
