"""The compressible model's heat conduction in the oracle against closed forms.

The hot_bubble fixture pins the compressible model in aggregate, but its
thermal conductivity (3e-5 m^2/s) moves no norm of the fixture beyond its
tolerance, so the conduction fluxes (compute_fluxes!.jl:1592-1756) and the
heating (conductive_heating.jl:98-117) are checked on their own here.

With rhobar = 1 and rho = rho0 = 0, theta = pbar = f for a smooth f of the
Cartesian position.  In terrain-following coordinates

* the x flux is -mu J f_x (the metric term turns the derivative at fixed
  zhat into the Cartesian one), likewise the y flux,
* the z flux is -mu (f_z - z_x f_x - z_y f_y) with z_x = -G13 J,
* the heating -(rho + rhobar) / J div(flux) is mu times the Cartesian
  Laplacian of f,

each to second order in the grid spacing: the error against the closed form
must fall about fourfold when the grid is refined twice.  A dropped metric
term, a wrong sign, a wrong 1 / (2 dz) or a shifted index leaves an error
that does not shrink.
"""
import ctypes

import numpy as np
import pytest

LX = LY = 20000.0
LZ = 10000.0
KAPPA_T = 2.0  # m^2/s


def case(pcf, n):
    nl = pcf.Namelists()
    nl.atmosphere.model = "Compressible"
    nl.atmosphere.background = "Isentropic"
    nl.atmosphere.thermal_conductivity = KAPPA_T
    nl.domain = pcf.DomainNamelist(x_size=n, y_size=n, z_size=n, lx=LX, ly=LY, lz=LZ)
    nl.grid = pcf.GridNamelist(resolved_topography=pcf.GaussianHills(((400.0, 0.0, 0.0, 5000.0),)))
    return nl


def f_and_derivatives(x, y, z, lx, ly, lz):
    ax, ay, az = 2 * np.pi / lx, 2 * np.pi / ly, np.pi / lz
    s = np.sin(ax * x) * np.sin(ay * y)
    f = 2.0 + s * np.cos(az * z) + 0.3 * z / lz
    fx = ax * np.cos(ax * x) * np.sin(ay * y) * np.cos(az * z)
    fy = ay * np.sin(ax * x) * np.cos(ay * y) * np.cos(az * z)
    fz = -az * s * np.sin(az * z) + 0.3 / lz
    lap = -(ax * ax + ay * ay + az * az) * s * np.cos(az * z)
    return f, fx, fy, fz, lap


def run(pcf, abi, oracle_mod, n):
    nl = case(pcf, n)
    o = oracle_mod.Oracle(pcf.to_config(nl), abi)
    c = o.constants()
    lx, ly, lz = LX / c["lref"], LY / c["lref"], LZ / c["lref"]
    mu = KAPPA_T / c["uref"] / c["lref"]
    xs, ys = np.empty(o.nxx), np.empty(o.nyy)
    oracle_mod.lib().orc_grid_1d(o.h, xs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 ys.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    zc, jac, m13, m23 = o.get("ZC"), o.get("JAC"), o.get("MET13"), o.get("MET23")
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    X, Y = X[:, :, None] * np.ones_like(zc), Y[:, :, None] * np.ones_like(zc)
    f = f_and_derivatives(X, Y, zc, lx, ly, lz)[0]
    zero = np.zeros_like(zc)
    o.set("RHOBAR", np.ones_like(zc))
    o.set("PBAR", f)
    for name in ("RHO", "P0_RHO", "P0_U", "P0_V", "P0_W"):
        o.set(name, zero)
    o.compute_theta_fluxes()
    fluxes = [o.get(f"PHITHETA_{d}") for d in "XYZ"]
    # heating = DP / dt of one LHS P update with zero P fluxes (u0 = v0 = w0 = 0);
    # compute_fluxes zeroes the conduction fluxes at the walls first, as in a step
    o.compute_fluxes()
    o.update_lhs(1.0, 1, abi.VAR_P)
    heating = o.get("DP")
    nb, nz = 3, n
    ii, jj = slice(nb, nb + n), slice(nb, nb + n)
    kin = slice(nb + 2, nb + nz - 2)  # two layers from each wall
    errs = {}
    # x faces (i + 1/2): Cartesian point halfway, J averaged
    xe = 0.5 * (X[ii, jj, kin] + X[nb + 1:nb + n + 1, jj, kin])
    ze = 0.5 * (zc[ii, jj, kin] + zc[nb + 1:nb + n + 1, jj, kin])
    je = 0.5 * (jac[ii, jj, kin] + jac[nb + 1:nb + n + 1, jj, kin])
    _, fx, _, _, _ = f_and_derivatives(xe, Y[ii, jj, kin], ze, lx, ly, lz)
    want = -mu * je * fx
    errs["x"] = np.max(np.abs(fluxes[0][ii, jj, kin] - want)) / np.max(np.abs(want))
    ye = 0.5 * (Y[ii, jj, kin] + Y[ii, nb + 1:nb + n + 1, kin])
    ze = 0.5 * (zc[ii, jj, kin] + zc[ii, nb + 1:nb + n + 1, kin])
    je = 0.5 * (jac[ii, jj, kin] + jac[ii, nb + 1:nb + n + 1, kin])
    _, _, fy, _, _ = f_and_derivatives(X[ii, jj, kin], ye, ze, lx, ly, lz)
    want = -mu * je * fy
    errs["y"] = np.max(np.abs(fluxes[1][ii, jj, kin] - want)) / np.max(np.abs(want))
    # z faces (k + 1/2) of the interior, k0 .. k1 - 1
    k_lo, k_hi = slice(nb, nb + nz - 1), slice(nb + 1, nb + nz)
    ze = 0.5 * (zc[ii, jj, k_lo] + zc[ii, jj, k_hi])
    zx = -0.5 * (m13[ii, jj, k_lo] * jac[ii, jj, k_lo] + m13[ii, jj, k_hi] * jac[ii, jj, k_hi])
    zy = -0.5 * (m23[ii, jj, k_lo] * jac[ii, jj, k_lo] + m23[ii, jj, k_hi] * jac[ii, jj, k_hi])
    _, fx, fy, fz, _ = f_and_derivatives(X[ii, jj, k_lo], Y[ii, jj, k_lo], ze, lx, ly, lz)
    want = -mu * (fz - zx * fx - zy * fy)
    errs["z"] = np.max(np.abs(fluxes[2][ii, jj, k_lo] - want)) / np.max(np.abs(want))
    _, _, _, _, lap = f_and_derivatives(X[ii, jj, kin], Y[ii, jj, kin], zc[ii, jj, kin], lx, ly, lz)
    want = mu * lap
    errs["heating"] = np.max(np.abs(heating[ii, jj, kin] - want)) / np.max(np.abs(want))
    # the heating is a flux divergence: sum(heating J / (rho + rhobar)) = 0 over the interior
    full = slice(nb, nb + nz)
    cons = np.sum(heating[ii, jj, full] * jac[ii, jj, full])
    errs["conservation"] = abs(cons) / np.sum(np.abs(heating[ii, jj, full] * jac[ii, jj, full]))
    return errs


def test_conduction_matches_closed_forms_to_second_order(pcf, abi, oracle_mod):
    coarse = run(pcf, abi, oracle_mod, 16)
    fine = run(pcf, abi, oracle_mod, 32)
    print("coarse", coarse)
    print("fine", fine)
    for key in ("x", "y", "z", "heating"):
        assert fine[key] < 0.05, (key, fine[key])
        assert coarse[key] / fine[key] > 3.0, (key, coarse[key], fine[key])
    assert coarse["conservation"] < 1e-12 and fine["conservation"] < 1e-12


def test_conduction_off_without_conductivity(pcf, abi, oracle_mod):
    nl = case(pcf, 8)
    nl.atmosphere.thermal_conductivity = 0.0
    o = oracle_mod.Oracle(pcf.to_config(nl), abi)
    o.set("P0_RHO", o.get("P0_RHO") + 0.01)
    o.compute_theta_fluxes()
    for d in "XYZ":
        assert not np.any(o.get(f"PHITHETA_{d}"))
