"""Namelists mirroring src/Types/NamelistTypes of PinCFlow.jl.

Each dataclass keeps the keyword names and defaults of the reference
constructor (DomainNamelist.jl:72-102, DiscretizationNamelist.jl:44-60,
PoissonNamelist.jl:48-66, AtmosphereNamelist.jl:108-156, GridNamelist.jl:35-45,
SpongeNamelist.jl:48-66, OutputNamelist.jl, Namelists.jl).  Symbol options
are strings (``model="Boussinesq"`` for ``model = :Boussinesq``) and an
unknown option raises ``ValueError`` as the reference's ``@dispatch`` macros
raise an error.

The reference takes Julia closures for topography, stretching, initial fields
and sponges.  Closures cannot run on the GPU, so those fields take one of the
parametric forms below (or a number for a uniform field); the forms cover the
reference examples (periodic_hill.jl, mountain_wave.jl, cold_bubble.jl,
vortex.jl) and the production configurations.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple, Union

from . import _abi


# ---------------------------------------------------------------------------
# parametric forms standing in for the reference's function-valued entries
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class CosineHill:
    """h(x, y) = h0 / 2 (1 + cos(pi x / l0)) (periodic_hill.jl:29)."""
    h0: float
    l0: float


@dataclass(frozen=True)
class AgnesiHill:
    """h(x, y) = h0 / (1 + (x^2 + y^2) / l0^2) (mountain_wave.jl:34)."""
    h0: float
    l0: float


@dataclass(frozen=True)
class GaussianHills:
    """h(x, y) = sum_n h_n exp(-((x - x_n)^2 + (y - y_n)^2) / w_n^2); hills = ((h, x0, y0, w), ...)."""
    hills: Tuple[Tuple[float, float, float, float], ...]


@dataclass(frozen=True)
class SinhStretching:
    """zhat -> lz sinh(b zhat / lz) / sinh(b): monotone, fixes 0 and lz."""
    b: float


@dataclass(frozen=True)
class CosineBubble:
    """A (1 + cos(pi r)) for r <= 1, r^2 = sum ((x_i - c_i) / r_i)^2 over r_i > 0 (cold_bubble.jl:21-28)."""
    amplitude: float
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0
    rx: float = 0.0
    ry: float = 0.0
    rz: float = 0.0


@dataclass(frozen=True)
class Vortex:
    """Winds of vortex.jl:23-39: u = -A y/ry (1 + cos(pi r))/2, v = A x/rx (1 + cos(pi r))/2."""
    amplitude: float
    rx: float
    ry: float


@dataclass(frozen=True)
class TracerDipole:
    """initial_tracer of vortex.jl:50-58: sign(x) (1 + cos(pi r)) / 2 for
    r = sqrt(((|x| - rx) / rx)^2 + (y / ry)^2) <= 1, else 0."""
    rx: float
    ry: float


@dataclass(frozen=True)
class Sin2Sponge:
    """lhs_sponge of mountain_wave.jl:40-53: alpharmax (ax + ay + az) / divisor."""
    alpharmax: float
    dxr: float = 0.0
    dyr: float = 0.0
    dzr: float = 0.0
    divisor: float = 3.0


@dataclass(frozen=True)
class Sin2TopSponge:
    """rhs_sponge of periodic_hill.jl:33-36: A sin^2(pi/2 (z - zr)/(lz - zr)) / dt for z >= zr."""
    zr: float
    amplitude: float = 1.0


# ---------------------------------------------------------------------------
# namelists
# ---------------------------------------------------------------------------
@dataclass
class DomainNamelist:
    x_size: int = 1
    y_size: int = 1
    z_size: int = 1
    nbx: int = 3
    nby: int = 3
    nbz: int = 3
    lx: float = 1.0e3
    ly: float = 1.0e3
    lz: float = 1.0e3
    npx: int = 1
    npy: int = 1
    npz: int = 1


@dataclass
class OutputNamelist:
    output_variables: Sequence[str] = ()
    save_ray_volumes: bool = False
    prepare_restart: bool = False
    restart: bool = False
    iin: int = -1
    output_steps: bool = False
    nout: int = 1
    iterations: int = 1
    output_interval: float = 3.6e3
    tmax: float = 3.6e3
    input_file: str = "./pincflow_input.h5"
    output_file: str = "./pincflow_output.h5"


@dataclass
class DiscretizationNamelist:
    cfl_number: float = 5.0e-1
    wkb_cfl_number: float = 3.0e-1
    dtmin: float = 1.0e-6
    dtmax: float = 1.0e3
    # addition of this build: arithmetic of the RK stages' fluxes, "fp64"
    # (reference) or "mixed" (reconstruction, fluxes and flux divergences in
    # FP32, updates in FP64)
    flux_precision: str = "fp64"
    adaptive_time_step: bool = True
    limiter_type: str = "MCVariant"


@dataclass
class PoissonNamelist:
    tolerance: float = 1.0e-8
    poisson_iterations: int = 1000
    preconditioner: bool = True
    dtau: float = 1.0
    preconditioner_iterations: int = 2
    initial_cleaning: bool = True
    tolerance_is_relative: bool = False
    # additions of this build: "fp64" (reference arithmetic) or "mixed" (the
    # FP64 BiCGSTAB with the preconditioner evaluated in FP32)
    precision: str = "fp64"
    # "zero" (as the reference), "previous" (the same stage's solution of the
    # previous time step), "extrapolated" (the linear extrapolation in time of
    # the same stage's last two solutions) or "learned" (the attached
    # network); mixed solve only
    initial_guess: str = "zero"
    # "line_relaxation" (the reference's) or "multigrid"; mixed solve only
    preconditioner_type: str = "line_relaxation"
    # multigrid: pre- and post-smoothing steps per level, steps on the coarsest level
    mg_smoothing_steps: int = 2
    mg_coarse_steps: int = 20
    # "fp32" or "bf16": storage of the finest level's coefficients read by the
    # mixed-precision preconditioner
    preconditioner_storage: str = "fp32"


@dataclass
class AtmosphereNamelist:
    model: str = "PseudoIncompressible"
    specify_reynolds_number: bool = False
    inverse_reynolds_number: float = 0.0
    kinematic_viscosity: float = 1.5e-5
    thermal_conductivity: float = 3.0e-5
    kinematic_diffusivity: float = 0.0
    background: str = "Isothermal"
    buoyancy_frequency: float = 1.0e-2
    potential_temperature: float = 3.0e2
    temperature: float = 3.0e2
    ground_pressure: float = 1.0e5
    coriolis_frequency: float = 1.0e-4
    tropopause_height: float = 1.0e4
    troposphere_lapse_rate: float = 6.5e-3
    stratosphere_lapse_rate: float = -5.0e-3
    initial_rhop: Union[float, CosineBubble] = 0.0
    initial_u: Union[float, Vortex] = 0.0
    initial_v: Union[float, Vortex] = 0.0
    initial_w: float = 0.0
    initial_pip: float = 0.0
    initial_thetap: Union[float, CosineBubble] = 0.0
    buoyancy_initialization: str = "initial_rhop"


@dataclass
class GridNamelist:
    vertical_grid_stretching: Optional[SinhStretching] = None
    resolved_topography: Union[None, CosineHill, AgnesiHill, GaussianHills] = None


@dataclass
class SpongeNamelist:
    damp_horizontal_wind_on_rhs: bool = False
    relax_to_mean: bool = False
    lhs_sponge: Optional[Sin2Sponge] = None
    rhs_sponge: Optional[Sin2TopSponge] = None
    relaxed_u: float = 0.0
    relaxed_v: float = 0.0
    relaxed_w: float = 0.0


@dataclass
class WKBNamelist:
    wkb_mode: str = "NoWKB"


@dataclass
class TracerNamelist:
    tracer_setup: str = "NoTracer"
    leading_order_impact: bool = False
    initial_tracer: Union[float, TracerDipole] = 0.0


@dataclass
class Namelists:
    domain: DomainNamelist = field(default_factory=DomainNamelist)
    output: OutputNamelist = field(default_factory=OutputNamelist)
    discretization: DiscretizationNamelist = field(default_factory=DiscretizationNamelist)
    poisson: PoissonNamelist = field(default_factory=PoissonNamelist)
    atmosphere: AtmosphereNamelist = field(default_factory=AtmosphereNamelist)
    grid: GridNamelist = field(default_factory=GridNamelist)
    sponge: SpongeNamelist = field(default_factory=SpongeNamelist)
    wkb: WKBNamelist = field(default_factory=WKBNamelist)
    tracer: TracerNamelist = field(default_factory=TracerNamelist)


_MODELS = {"Boussinesq": _abi.BOUSSINESQ, "PseudoIncompressible": _abi.PSEUDO_INCOMPRESSIBLE,
           "Compressible": _abi.COMPRESSIBLE}
_BACKGROUNDS = {
    "NeutralStratification": _abi.NEUTRAL_STRATIFICATION,
    "StableStratification": _abi.STABLE_STRATIFICATION,
    "Isothermal": _abi.ISOTHERMAL,
    "Isentropic": _abi.ISENTROPIC,
    "Realistic": _abi.REALISTIC,
    "LapseRates": _abi.LAPSE_RATES,
}
_PRECISIONS = {"fp64": _abi.PRECISION_FP64, "mixed": _abi.PRECISION_MIXED}


def _option(table, value, what):
    if value not in table:
        raise ValueError(f"Invalid {what}: {value!r} (valid: {sorted(table)})")
    return table[value]


def to_config(nl: Namelists) -> _abi.PcfConfig:
    """Serialise Namelists into the POD pcf_config the C ABI takes."""
    d, o, di, po, a, g, s = nl.domain, nl.output, nl.discretization, nl.poisson, nl.atmosphere, nl.grid, nl.sponge
    if (d.npx, d.npy, d.npz) != (1, 1, 1):
        raise ValueError("one GPU holds the whole domain: npx = npy = npz = 1")
    if di.limiter_type != "MCVariant":
        raise ValueError(f"Invalid limiter_type: {di.limiter_type!r}")
    if nl.wkb.wkb_mode != "NoWKB":
        raise ValueError("MS-GWaM (wkb_mode != NoWKB) is not part of this library")
    if nl.tracer.tracer_setup not in ("NoTracer", "TracerOn"):
        raise ValueError(f"Invalid tracer_setup: {nl.tracer.tracer_setup!r}")
    if nl.tracer.leading_order_impact:
        raise ValueError("leading_order_impact needs MS-GWaM, which is not part of this library")
    if not isinstance(nl.tracer.initial_tracer, TracerDipole) and nl.tracer.initial_tracer != 0.0:
        raise ValueError("initial_tracer: 0 or a TracerDipole")
    if a.kinematic_diffusivity != 0.0:
        raise ValueError("kinematic_diffusivity != 0 is not part of this library")
    c = _abi.PcfConfig()
    c.abi_version = _abi.ABI_VERSION
    c.x_size, c.y_size, c.z_size = d.x_size, d.y_size, d.z_size
    c.nbx, c.nby, c.nbz = d.nbx, d.nby, d.nbz
    c.lx, c.ly, c.lz = d.lx, d.ly, d.lz
    c.cfl_number, c.dtmin, c.dtmax = di.cfl_number, di.dtmin, di.dtmax
    c.adaptive_time_step = int(di.adaptive_time_step)
    c.poisson_iterations = po.poisson_iterations
    c.tolerance, c.dtau = po.tolerance, po.dtau
    c.preconditioner = int(po.preconditioner)
    c.preconditioner_iterations = po.preconditioner_iterations
    c.initial_cleaning = int(po.initial_cleaning)
    c.tolerance_is_relative = int(po.tolerance_is_relative)
    c.precision = _option(_PRECISIONS, po.precision, "precision")
    c.flux_precision = _option(_PRECISIONS, di.flux_precision, "flux_precision")
    c.model = _option(_MODELS, a.model, "model")
    c.background = _option(_BACKGROUNDS, a.background, "background")
    c.specify_reynolds_number = int(a.specify_reynolds_number)
    c.inverse_reynolds_number = a.inverse_reynolds_number
    c.kinematic_viscosity = a.kinematic_viscosity
    c.thermal_conductivity = a.thermal_conductivity
    c.buoyancy_frequency = a.buoyancy_frequency
    c.potential_temperature = a.potential_temperature
    c.temperature = a.temperature
    c.ground_pressure = a.ground_pressure
    c.coriolis_frequency = a.coriolis_frequency
    c.tropopause_height = a.tropopause_height
    c.troposphere_lapse_rate = a.troposphere_lapse_rate
    c.stratosphere_lapse_rate = a.stratosphere_lapse_rate
    if a.initial_pip != 0.0:
        raise ValueError("initial_pip must be 0")
    if a.buoyancy_initialization not in ("initial_rhop", "initial_thetap"):
        raise ValueError("Unknown buoyancy-initialization mode!")
    c.buoyancy_initialization = _abi.INIT_RHOP if a.buoyancy_initialization == "initial_rhop" else _abi.INIT_THETAP
    bubble = a.initial_rhop if a.buoyancy_initialization == "initial_rhop" else a.initial_thetap
    c.perturbation = _abi.PERT_NONE
    if isinstance(bubble, CosineBubble):
        c.perturbation = _abi.PERT_COSINE_BUBBLE
        c.pert_amplitude = bubble.amplitude
        c.pert_x0, c.pert_y0, c.pert_z0 = bubble.x0, bubble.y0, bubble.z0
        c.pert_rx, c.pert_ry, c.pert_rz = bubble.rx, bubble.ry, bubble.rz
    elif bubble != 0.0:
        raise ValueError("a uniform initial_rhop / initial_thetap must be 0")
    c.initial_u = a.initial_u if not isinstance(a.initial_u, Vortex) else 0.0
    c.initial_v = a.initial_v if not isinstance(a.initial_v, Vortex) else 0.0
    c.initial_w = a.initial_w
    if isinstance(a.initial_u, Vortex) or isinstance(a.initial_v, Vortex):
        vx = a.initial_u if isinstance(a.initial_u, Vortex) else a.initial_v
        if c.perturbation != _abi.PERT_NONE:
            raise ValueError("a vortex cannot be combined with a buoyancy bubble")
        c.perturbation = _abi.PERT_VORTEX
        c.pert_amplitude, c.pert_rx, c.pert_ry = vx.amplitude, vx.rx, vx.ry
    c.stretching = _abi.STRETCH_NONE
    if isinstance(g.vertical_grid_stretching, SinhStretching):
        c.stretching = _abi.STRETCH_SINH
        c.stretching_parameter = g.vertical_grid_stretching.b
    elif g.vertical_grid_stretching is not None:
        raise ValueError("unsupported vertical_grid_stretching form")
    t = g.resolved_topography
    if t is None:
        c.topography = _abi.TOPO_FLAT
    elif isinstance(t, CosineHill):
        c.topography, c.topo_height, c.topo_width = _abi.TOPO_COSINE_HILL, t.h0, t.l0
    elif isinstance(t, AgnesiHill):
        c.topography, c.topo_height, c.topo_width = _abi.TOPO_AGNESI, t.h0, t.l0
    elif isinstance(t, GaussianHills):
        if len(t.hills) > _abi.MAX_HILLS:
            raise ValueError(f"at most {_abi.MAX_HILLS} hills")
        c.topography = _abi.TOPO_GAUSSIANS
        c.n_hills = len(t.hills)
        for n, (h, x0, y0, w) in enumerate(t.hills):
            c.hills[n].height, c.hills[n].x0, c.hills[n].y0, c.hills[n].width = h, x0, y0, w
    else:
        raise ValueError("unsupported resolved_topography form")
    c.damp_horizontal_wind_on_rhs = int(s.damp_horizontal_wind_on_rhs)
    c.relax_to_mean = int(s.relax_to_mean)
    c.lhs_sponge = _abi.LHS_SPONGE_NONE
    c.lhs_divisor = 3.0
    if isinstance(s.lhs_sponge, Sin2Sponge):
        c.lhs_sponge = _abi.LHS_SPONGE_SIN2
        c.lhs_alpharmax = s.lhs_sponge.alpharmax
        c.lhs_dxr, c.lhs_dyr, c.lhs_dzr = s.lhs_sponge.dxr, s.lhs_sponge.dyr, s.lhs_sponge.dzr
        c.lhs_divisor = s.lhs_sponge.divisor
    elif s.lhs_sponge is not None:
        raise ValueError("unsupported lhs_sponge form")
    c.rhs_sponge = _abi.RHS_SPONGE_NONE
    c.rhs_amplitude = 1.0
    if isinstance(s.rhs_sponge, Sin2TopSponge):
        c.rhs_sponge = _abi.RHS_SPONGE_SIN2_TOP
        c.rhs_zr, c.rhs_amplitude = s.rhs_sponge.zr, s.rhs_sponge.amplitude
    elif s.rhs_sponge is not None:
        raise ValueError("unsupported rhs_sponge form")
    c.relaxed_u, c.relaxed_v, c.relaxed_w = s.relaxed_u, s.relaxed_v, s.relaxed_w
    c.output_steps = int(o.output_steps)
    c.iterations = o.iterations
    c.nout = o.nout
    c.output_interval = o.output_interval
    c.tmax = o.tmax
    c.initial_guess = _option({"zero": _abi.GUESS_ZERO, "previous": _abi.GUESS_PREVIOUS,
                               "learned": _abi.GUESS_LEARNED, "extrapolated": _abi.GUESS_EXTRAPOLATED},
                              po.initial_guess, "initial_guess")
    c.preconditioner_type = _option({"line_relaxation": _abi.PC_LINE_RELAXATION, "multigrid": _abi.PC_MULTIGRID},
                                    po.preconditioner_type, "preconditioner_type")
    c.mg_smoothing_steps = int(po.mg_smoothing_steps)
    c.mg_coarse_steps = int(po.mg_coarse_steps)
    c.preconditioner_storage = _option({"fp32": _abi.STORAGE_FP32, "bf16": _abi.STORAGE_BF16},
                                       po.preconditioner_storage, "preconditioner_storage")
    tr = nl.tracer
    c.tracer_setup = _abi.TRACER_ON if tr.tracer_setup == "TracerOn" else _abi.NO_TRACER
    if isinstance(tr.initial_tracer, TracerDipole):
        c.initial_tracer = _abi.TRACER_DIPOLE
        c.tracer_rx, c.tracer_ry = tr.initial_tracer.rx, tr.initial_tracer.ry
    return c


def config_bytes(cfg: _abi.PcfConfig) -> bytes:
    return ctypes.string_at(ctypes.addressof(cfg), ctypes.sizeof(cfg))
