"""The oracle's tracer transport (TracerOn) against what the mathematics
fixes, beside the reference's own vortex chi norms (test_oracle_golden.py).

* Mass: the flux-form update with periodic x, y and zero vertical wall
  fluxes (set_tracer_vertical_boundaries!.jl, BoundaryFluxes) conserves
  sum J chi over the interior to rounding, through RK3, the reset of step (3)
  and the second explicit stage.
* Consistency with the continuity equation: chi = c (rho + rhobar), a
  uniform mixing ratio, stays uniform in the Boussinesq vortex, because the
  tracer flux (reconstruction of chi / pbar with the carrier (J pbar) u) is
  c times the mass flux whose discrete divergence the pressure correction
  removes to the Poisson tolerance.  A missing J or pbar factor, a wrong face
  index or an upwind side swapped breaks either check.
"""
import numpy as np
import pytest

INTERIOR = (slice(3, -3),) * 3


def vortex_oracle(pcf, abi, oracle_mod):
    nl = pcf.examples.vortex(x_size=16, y_size=16)
    o = oracle_mod.Oracle(pcf.to_config(nl), abi)
    o.initial_cleaning()
    return o


def test_tracer_mass_is_conserved(pcf, abi, oracle_mod):
    o = vortex_oracle(pcf, abi, oracle_mod)
    jac = o.get("JAC")

    def mass():
        return float(np.sum((jac * o.get("CHI"))[INTERIOR]))

    m0 = mass()
    scale = float(np.sum(np.abs(jac * o.get("CHI"))[INTERIOR]))
    for _ in range(3):
        o.step()
        assert abs(mass() - m0) <= 1e-13 * scale
    # the tracer did move
    assert not np.allclose(o.get("CHI"), o.get("INITIALTRACER"))


def test_uniform_mixing_ratio_stays_uniform(pcf, abi, oracle_mod):
    o = vortex_oracle(pcf, abi, oracle_mod)
    dens = o.get("RHOBAR") + o.get("RHO")
    o.set("CHI", 0.7 * dens)
    o.set("INITIALTRACER", 0.7 * dens)
    for _ in range(3):
        o.step()
    ratio = (o.get("CHI") / (o.get("RHOBAR") + o.get("RHO")))[INTERIOR]
    assert float(np.max(np.abs(ratio - 0.7))) <= 1e-8


def test_tracer_off_by_default(pcf, abi, oracle_mod):
    nl = pcf.examples.periodic_hill(x_size=8, z_size=8)
    o = oracle_mod.Oracle(pcf.to_config(nl), abi)
    with pytest.raises(oracle_mod.OracleError):
        o.get("CHI")
