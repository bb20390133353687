"""MI355X-native semi-implicit dynamical-core step of PinCFlow.jl.

Host-side mirror of the reference interface: ``Namelists(...)`` and
``integrate(namelists)`` (src/Integration/integrate.jl:68-415) over the C ABI
of libpincflow_hip.so (include/pincflow_hip.h), whose HIP kernels keep every
field resident in HBM.
"""
from . import _abi
from .namelists import (AgnesiHill, AtmosphereNamelist, CosineBubble, CosineHill, DiscretizationNamelist,
                        DomainNamelist, GaussianHills, GridNamelist, Namelists, OutputNamelist, PoissonNamelist,
                        Sin2Sponge, Sin2TopSponge, SinhStretching, SpongeNamelist, TracerDipole, TracerNamelist,
                        Vortex, WKBNamelist, to_config)
from . import examples, learned, output
from .ensemble import EnsembleError, SolveBatch, ensemble
from .state import PinCFlowError, PoissonNotConverged, State, integrate

__all__ = [
    "State", "integrate", "ensemble", "EnsembleError", "SolveBatch", "PinCFlowError", "PoissonNotConverged",
    "AgnesiHill", "AtmosphereNamelist", "CosineBubble", "CosineHill", "DiscretizationNamelist", "DomainNamelist",
    "GaussianHills", "GridNamelist", "Namelists", "OutputNamelist", "PoissonNamelist", "Sin2Sponge",
    "Sin2TopSponge", "SinhStretching", "SpongeNamelist", "TracerDipole", "TracerNamelist", "Vortex", "WKBNamelist",
    "to_config",
    "examples", "learned", "output",
]
