"""rustpde_amd: MI355X (gfx950) implementation of rustpde's 2-D Rayleigh-Benard step.

Host mirror of the reference's API over librustpde_amd.so (HIP kernels + C ABI):
Navier2D.new / Navier2D.new_periodic, the Integrate methods, integrate(), the
funspace Space2 transforms, the SolverField solvers, ensembles of confined instances,
Statistics and checkpoint/restart (Navier2D.write / read), and the adjoint descent for
steady states (Navier2DAdjoint).
"""
from ._lib import RbcError, load  # noqa: F401
from .adjoint import Navier2DAdjoint  # noqa: F401
from .funspace import (Space2, cheb_dirichlet, cheb_dirichlet_bc, cheb_neumann, chebyshev,  # noqa: F401
                       fourier_r2c)
from .ensemble import Ensemble  # noqa: F401
from .integrate import integrate  # noqa: F401
from .navier import Navier2D, get_ka, get_nu  # noqa: F401
from .solver import HHOLTZ, HHOLTZ_ADI, POISSON, Solver  # noqa: F401
from .statistics import Statistics  # noqa: F401
