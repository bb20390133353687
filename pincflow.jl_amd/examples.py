"""Example configurations (src/Examples of PinCFlow.jl) as Namelists.

``periodic_hill``, ``mountain_wave``, ``cold_bubble``, ``hot_bubble`` and ``vortex`` return the
Namelists their reference namesake passes to ``integrate``; the keyword
arguments keep the reference names and defaults.  ``mountain_wave_3d`` and
``gaussian_hills_3d`` are the production configurations measured by bench.py;
``bubble_member`` builds the members of the seeded bubble ensemble.
"""
from __future__ import annotations

import math

from .namelists import (AgnesiHill, AtmosphereNamelist, CosineBubble, CosineHill, DiscretizationNamelist,
                        DomainNamelist, GaussianHills, GridNamelist, Namelists, OutputNamelist, PoissonNamelist,
                        Sin2Sponge, Sin2TopSponge, SinhStretching, SpongeNamelist, TracerDipole, TracerNamelist,
                        Vortex)


def periodic_hill(*, x_size: int = 40, z_size: int = 40, output_steps: bool = False,
                  iterations: int = 1, precision: str = "fp64") -> Namelists:
    """src/Examples/periodic_hill.jl:3-50 (Boussinesq, stable stratification)."""
    h0, l0, lz, zr = 500.0, 10000.0, 20000.0, 10000.0
    return Namelists(
        atmosphere=AtmosphereNamelist(model="Boussinesq", background="StableStratification",
                                      coriolis_frequency=0.0, initial_u=10.0),
        domain=DomainNamelist(x_size=x_size, z_size=z_size, lx=20000.0, lz=lz),
        grid=GridNamelist(resolved_topography=CosineHill(h0, l0)),
        output=OutputNamelist(output_variables=("w",), prepare_restart=True, output_steps=output_steps,
                              iterations=iterations),
        sponge=SpongeNamelist(rhs_sponge=Sin2TopSponge(zr=zr)),
        poisson=PoissonNamelist(precision=precision),
    )


def mountain_wave(*, x_size: int = 40, y_size: int = 40, z_size: int = 40) -> Namelists:
    """src/Examples/mountain_wave.jl:3-65 (pseudo-incompressible, isothermal 300 K)."""
    h0, l0, lx, ly, lz = 100.0, 1000.0, 20000.0, 20000.0, 20000.0
    return Namelists(
        atmosphere=AtmosphereNamelist(coriolis_frequency=0.0, initial_u=10.0),
        domain=DomainNamelist(x_size=x_size, y_size=y_size, z_size=z_size, lx=lx, ly=ly, lz=lz),
        grid=GridNamelist(resolved_topography=AgnesiHill(h0, l0)),
        output=OutputNamelist(output_variables=("w",), prepare_restart=True),
        sponge=SpongeNamelist(lhs_sponge=Sin2Sponge(alpharmax=0.0179, dxr=lx / 2, dyr=ly / 2, dzr=lz / 2),
                              relaxed_u=10.0),
    )


def cold_bubble(*, x_size: int = 40, z_size: int = 40) -> Namelists:
    """src/Examples/cold_bubble.jl:3-52 (pseudo-incompressible, isentropic)."""
    lx, lz = 20000.0, 20000.0
    rx, rz = lx / 8, lz / 8
    return Namelists(
        atmosphere=AtmosphereNamelist(background="Isentropic",
                                      initial_rhop=CosineBubble(0.005, x0=0.0, z0=3 * rz, rx=rx, rz=rz)),
        discretization=DiscretizationNamelist(dtmax=60.0),
        domain=DomainNamelist(x_size=x_size, z_size=z_size, lx=lx, lz=lz),
        output=OutputNamelist(output_variables=("thetap",), prepare_restart=True),
    )


def hot_bubble(*, x_size: int = 40, z_size: int = 40) -> Namelists:
    """src/Examples/hot_bubble.jl:3-52 (compressible, isentropic)."""
    lx, lz = 20000.0, 20000.0
    rx, rz = lx / 8, lz / 8
    return Namelists(
        atmosphere=AtmosphereNamelist(model="Compressible", background="Isentropic",
                                      initial_rhop=CosineBubble(-0.005, x0=0.0, z0=5 * rz, rx=rx, rz=rz)),
        discretization=DiscretizationNamelist(dtmax=60.0),
        domain=DomainNamelist(x_size=x_size, z_size=z_size, lx=lx, lz=lz),
        output=OutputNamelist(output_variables=("thetap",), prepare_restart=True),
    )


def vortex(*, x_size: int = 40, y_size: int = 40) -> Namelists:
    """src/Examples/vortex.jl:3-78 (Boussinesq, neutral) with its passive
    tracer dipole."""
    lx, ly = 20000.0, 20000.0
    v = Vortex(5.0, lx / 4, ly / 4)
    return Namelists(
        atmosphere=AtmosphereNamelist(model="Boussinesq", background="NeutralStratification",
                                      initial_u=v, initial_v=v),
        domain=DomainNamelist(x_size=x_size, y_size=y_size, lx=lx, ly=ly),
        output=OutputNamelist(prepare_restart=True),
        tracer=TracerNamelist(tracer_setup="TracerOn", initial_tracer=TracerDipole(lx / 4, ly / 4)),
    )


def mountain_wave_3d(*, x_size: int = 256, y_size: int = 256, z_size: int = 128, iterations: int = 100,
                     precision: str = "fp64", stretching: float = 1.5) -> Namelists:
    """Pseudo-incompressible mountain wave at production size: isothermal 250 K,
    lx = ly = 200 km, lz = 30 km with sinh stretching, one Gaussian hill
    (500 m high, 10 km half-width) at the centre, u0 = 10 m/s, f = 1e-4 s^-1,
    LHS sponge relaxing to u0 over the top third, fixed number of steps."""
    lx, ly, lz = 200000.0, 200000.0, 30000.0
    return Namelists(
        atmosphere=AtmosphereNamelist(model="PseudoIncompressible", background="Isothermal", temperature=250.0,
                                      coriolis_frequency=1.0e-4, initial_u=10.0),
        domain=DomainNamelist(x_size=x_size, y_size=y_size, z_size=z_size, lx=lx, ly=ly, lz=lz),
        grid=GridNamelist(vertical_grid_stretching=SinhStretching(stretching) if stretching else None,
                          resolved_topography=GaussianHills(((500.0, 0.0, 0.0, 10000.0),))),
        output=OutputNamelist(output_steps=True, iterations=iterations),
        sponge=SpongeNamelist(lhs_sponge=Sin2Sponge(alpharmax=0.0179, dzr=lz / 3, divisor=1.0), relaxed_u=10.0),
        poisson=PoissonNamelist(precision=precision),
    )


def gaussian_hills_3d(*, x_size: int = 512, y_size: int = 512, z_size: int = 128, iterations: int = 100,
                      seed: int = 2024, n_hills: int = 16, precision: str = "mixed") -> Namelists:
    """As mountain_wave_3d, with the sum of n seeded Gaussian hills (height
    U(100, 800) m, half-width U(5, 20) km, uniform centres); the hills come from
    numpy's PCG64 generator with the given seed."""
    import numpy as np

    lx, ly = 200000.0, 200000.0
    rng = np.random.Generator(np.random.PCG64(seed))
    hills = tuple((float(rng.uniform(100.0, 800.0)), float(rng.uniform(-lx / 2, lx / 2)),
                   float(rng.uniform(-ly / 2, ly / 2)), float(rng.uniform(5000.0, 20000.0)))
                  for _ in range(n_hills))
    nl = mountain_wave_3d(x_size=x_size, y_size=y_size, z_size=z_size, iterations=iterations, precision=precision)
    nl.grid.resolved_topography = GaussianHills(hills)
    return nl


def bubble_member(*, seed: int, x_size: int = 0, z_size: int = 0, iterations: int = 50,
                  precision: str = "fp64") -> Namelists:
    """One member of the seeded bubble ensemble: from numpy's PCG64 generator
    with the given seed, a horizontal grid of 64, 96, 128 or 192 cells squared
    and 64 or 96 levels (x_size / z_size override them when positive), the
    Boussinesq (stable stratification) or pseudo-incompressible (isentropic)
    model, a hot or cold cosine bubble of potential-temperature amplitude
    U(-2, 2) K and radius U(1, 3) km at a random position, over flat ground
    or one Gaussian hill (U(100, 500) m, half-width U(2, 5) km), in a 60 km x
    60 km x 10 km box (cells at least three times wider than high, which the
    reference's z-line preconditioner needs to converge); `iterations` output
    steps."""
    import numpy as np

    lx = ly = 60000.0
    lz = 10000.0
    rng = np.random.Generator(np.random.PCG64(seed))
    nxy = int(rng.choice([64, 96, 128, 192]))
    nz = int(rng.choice([64, 96]))
    boussinesq = bool(rng.uniform() < 0.5)
    amp = float(rng.uniform(-2.0, 2.0))
    r = float(rng.uniform(1000.0, 3000.0))
    x0 = float(rng.uniform(-lx / 2 + r, lx / 2 - r))
    y0 = float(rng.uniform(-ly / 2 + r, ly / 2 - r))
    z0 = float(rng.uniform(r + 500.0, lz - r))
    hill = None
    if rng.uniform() < 0.5:
        hill = GaussianHills(((float(rng.uniform(100.0, 500.0)), 0.0, 0.0, float(rng.uniform(2000.0, 5000.0))),))
    return Namelists(
        atmosphere=AtmosphereNamelist(
            model="Boussinesq" if boussinesq else "PseudoIncompressible",
            background="StableStratification" if boussinesq else "Isentropic",
            buoyancy_initialization="initial_thetap",
            # A (1 + cos(pi r)) peaks at 2 A
            initial_thetap=CosineBubble(amp / 2, x0=x0, y0=y0, z0=z0, rx=r, ry=r, rz=r)),
        # dtmax of the bubble examples (cold_bubble.jl, hot_bubble.jl): at rest the CFL limit allows any step
        discretization=DiscretizationNamelist(dtmax=60.0),
        domain=DomainNamelist(x_size=x_size or nxy, y_size=x_size or nxy, z_size=z_size or nz, lx=lx, ly=ly, lz=lz),
        grid=GridNamelist(resolved_topography=hill),
        output=OutputNamelist(output_variables=("thetap",), output_steps=True, iterations=iterations),
        poisson=PoissonNamelist(precision=precision),
    )
