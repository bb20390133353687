"""The oracle's grid and background atmosphere against what fixes them
independently of how the reference writes them (Grid.jl:20-60, 175-312;
Atmosphere.jl:297-360; compute_n2!.jl:43-82; Constants.jl:103-144).

Geometry.  The terrain-following coordinate is z = h + (lz - h) s(zhat) / lz
with s the stretching of the flat column.  Every column is therefore an
affine image of a far-field column, the Jacobian dz/dzhat integrates to
lz - h over a column, and the contravariant metric follows from the chain
rule: G13 = dzhat/dx = -(dz/dx at fixed zhat) / J (likewise G23), and
G33 = |grad zhat|^2 = 1 / J^2 + G13^2 + G23^2.  The topography recovered
from the levels is the namelist's Gaussian hill, and the levels of the flat
column are the midpoints of the sinh-stretched edges.

Backgrounds.  The isothermal atmosphere (T0 constant) is the hydrostatic
ideal gas p = p_g exp(-g z / (R T0)), theta = T0 (p_g / p)^kappa, rho = p /
(R T0), P = rho theta; the isentropic one has theta = theta0 and (p /
p_g)^kappa = 1 - g z / (c_p theta0).  Both are made dimensionless with the
reference scales the reference documents (rho_ref = 1.184 kg m^-3, p_ref =
101325 Pa, R = 287 J kg^-1 K^-1, g = 9.81 m s^-2; a_ref^2 = p_ref / rho_ref,
L_ref = p_ref / (rho_ref g), theta_ref = a_ref^2 / R, t_ref = L_ref /
a_ref).  The isothermal buoyancy frequency is N^2 = g^2 / (c_p T0); the
oracle's centred difference on the stretched grid agrees to its
discretisation error.

A dropped factor, a sign, a swapped index or a wrong scale in any of these
fields fails one of the checks below by far more than its tolerance.
"""
import numpy as np
import pytest

NB = 3  # halo width of the oracle's fields
GAMMA, R, G = 1.4, 287.0, 9.81
KAPPA = (GAMMA - 1.0) / GAMMA
RHOREF, PREF = 1.184, 101325.0
AREF2 = PREF / RHOREF
LREF = PREF / RHOREF / G
THETAREF = AREF2 / R
TREF = LREF / np.sqrt(AREF2)


@pytest.fixture(scope="module")
def hill(pcf, oracle_mod, abi):
    nl = pcf.examples.mountain_wave_3d(x_size=48, y_size=40, z_size=24)
    cfg = pcf.to_config(nl)
    o = oracle_mod.Oracle(cfg, abi)
    f = {n: o.get(n) for n in ("ZC", "JAC", "MET13", "MET23", "MET33", "PBAR", "THETABAR", "RHOBAR", "N2")}
    return nl, cfg, f


def interior(a):
    return a[NB:-NB, NB:-NB, NB:-NB]


def test_columns_are_affine_images_of_the_flat_column(hill):
    nl, cfg, f = hill
    lz = cfg.lz / LREF
    zc, jac = f["ZC"], f["JAC"]
    flat_z, flat_j = zc[NB, NB, :], jac[NB, NB, :]  # a corner column, 140 km from the hill
    ks = slice(NB, -NB)
    # h from every level of every column: the same value on each level
    h = (zc[:, :, ks] - flat_z[ks]) / (1.0 - flat_z[ks] / lz)
    assert np.max(np.abs(h - h[:, :, :1])) <= 1e-12 * lz
    h = h[:, :, 0]
    assert np.max(np.abs(jac[:, :, ks] / flat_j[ks] - ((lz - h) / lz)[:, :, None])) <= 1e-13
    # dz/dzhat integrates to lz - h over the interior levels
    dzhat = lz / (zc.shape[2] - 2 * NB)
    col = np.sum(interior(jac), axis=2) * dzhat
    assert np.max(np.abs(col - (lz - h[NB:-NB, NB:-NB]))) <= 1e-12 * lz


def test_topography_is_the_namelist_hill(hill):
    nl, cfg, f = hill
    lz = cfg.lz / LREF
    zc = f["ZC"]
    flat_z = zc[NB, NB, :]
    h = (zc[:, :, NB] - flat_z[NB]) / (1.0 - flat_z[NB] / lz) * LREF
    nx, ny = zc.shape[0] - 2 * NB, zc.shape[1] - 2 * NB
    dx, dy = cfg.lx / nx, cfg.ly / ny
    x = -cfg.lx / 2 + (np.arange(nx) + 0.5) * dx
    y = -cfg.ly / 2 + (np.arange(ny) + 0.5) * dy
    (h0, x0, y0, w), = nl.grid.resolved_topography.hills
    expect = h0 * np.exp(-((x[:, None] - x0) ** 2 + (y[None, :] - y0) ** 2) / w ** 2)
    got = h[NB:-NB, NB:-NB]
    assert np.max(np.abs(got - expect)) <= 1e-9 * h0
    # the hill is centred: the levels are mirror-symmetric in x and in y
    assert np.max(np.abs(got - got[::-1, :])) <= 1e-9 * h0
    assert np.max(np.abs(got - got[:, ::-1])) <= 1e-9 * h0


def test_flat_column_levels_are_midpoints_of_sinh_stretched_edges(hill):
    nl, cfg, f = hill
    lz = cfg.lz / LREF
    nz = f["ZC"].shape[2] - 2 * NB
    b = nl.grid.vertical_grid_stretching.b
    edges = lz * np.sinh(b * np.arange(nz + 1) / nz) / np.sinh(b)
    mid = 0.5 * (edges[1:] + edges[:-1])
    assert np.max(np.abs(f["ZC"][NB, NB, NB:-NB] - mid)) <= 1e-13 * lz
    # J = (edge above - edge below) / dzhat on the flat column
    assert np.max(np.abs(f["JAC"][NB, NB, NB:-NB] - np.diff(edges) / (lz / nz))) <= 1e-12


def test_metric_is_the_chain_rule_of_the_levels(hill):
    nl, cfg, f = hill
    zc, jac = f["ZC"], f["JAC"]
    nx, ny = zc.shape[0] - 2 * NB, zc.shape[1] - 2 * NB
    dx, dy = cfg.lx / nx / LREF, cfg.ly / ny / LREF
    c = (slice(NB, -NB), slice(NB, -NB), slice(NB, -NB))
    dzdx = (zc[NB + 1:zc.shape[0] - NB + 1, NB:-NB, NB:-NB] - zc[NB - 1:-NB - 1, NB:-NB, NB:-NB]) / (2 * dx)
    dzdy = (zc[NB:-NB, NB + 1:zc.shape[1] - NB + 1, NB:-NB] - zc[NB:-NB, NB - 1:-NB - 1, NB:-NB]) / (2 * dy)
    g13, g23, g33, j = f["MET13"][c], f["MET23"][c], f["MET33"][c], jac[c]
    scale = np.max(np.abs(g13))
    assert scale > 1e-3  # the hill tilts the levels
    assert np.max(np.abs(g13 + dzdx / j)) <= 1e-11 * scale
    assert np.max(np.abs(g23 + dzdy / j)) <= 1e-11 * scale
    assert np.max(np.abs(g33 - (1.0 / j ** 2 + g13 ** 2 + g23 ** 2))) <= 1e-13 * np.max(g33)
    # G13 is odd about the hill's axis in x, G23 in y
    assert np.max(np.abs(g13 + g13[::-1, :, :])) <= 1e-11 * scale
    assert np.max(np.abs(g23 + g23[:, ::-1, :])) <= 1e-11 * scale


def test_isothermal_background_is_the_hydrostatic_ideal_gas(hill):
    nl, cfg, f = hill
    t0, pg = cfg.temperature, cfg.ground_pressure
    z = f["ZC"] * LREF
    p = pg * np.exp(-G * z / (R * t0))
    theta = t0 * (pg / p) ** KAPPA
    rho = p / (R * t0)
    for name, expect in (("PBAR", rho * theta / (RHOREF * THETAREF)), ("THETABAR", theta / THETAREF),
                         ("RHOBAR", rho / RHOREF)):
        got = f[name]
        assert np.max(np.abs(got - expect) / np.abs(expect)) <= 1e-12, name


def test_isothermal_buoyancy_frequency(hill):
    nl, cfg, f = hill
    cp = R / KAPPA
    n2 = G ** 2 / (cp * cfg.temperature) * TREF ** 2
    got = interior(f["N2"])
    # the centred difference of an exponential on the stretched levels
    assert np.max(np.abs(got / n2 - 1.0)) <= 2e-2
    assert abs(float(np.mean(got)) / n2 - 1.0) <= 1e-2


@pytest.mark.parametrize("example", ["cold_bubble", "hot_bubble"])
def test_isentropic_background_is_the_hydrostatic_adiabat(pcf, oracle_mod, abi, example):
    nl = getattr(pcf.examples, example)(x_size=16, z_size=20)
    cfg = pcf.to_config(nl)
    o = oracle_mod.Oracle(cfg, abi)
    th0, pg = cfg.potential_temperature, cfg.ground_pressure
    cp = R / KAPPA
    z = o.get("ZC") * LREF
    p = pg * (1.0 - G * z / (cp * th0)) ** (1.0 / KAPPA)
    rho = (p / pg) ** (1.0 / GAMMA) * pg / (R * th0)  # rho theta = P = (p_g / R) (p / p_g)^(1/gamma)
    for name, expect in (("PBAR", rho * th0 / (RHOREF * THETAREF)), ("THETABAR", np.full_like(z, th0 / THETAREF)),
                         ("RHOBAR", rho / RHOREF)):
        got = o.get(name)
        assert np.max(np.abs(got - expect) / np.abs(expect)) <= 1e-12, name
    assert np.max(np.abs(interior(o.get("N2")))) == 0.0


@pytest.mark.parametrize("example,stable", [("periodic_hill", True), ("vortex", False)])
def test_boussinesq_background_is_uniform(pcf, oracle_mod, abi, example, stable):
    """the Boussinesq reference state: rho_ref, theta0, P = rho theta, and N^2
    the namelist's buoyancy frequency squared in units of t_ref^-2 (zero for
    the neutral stratification)"""
    nl = getattr(pcf.examples, example)(x_size=16)
    cfg = pcf.to_config(nl)
    o = oracle_mod.Oracle(cfg, abi)
    th0 = cfg.potential_temperature / THETAREF
    assert np.array_equal(o.get("RHOBAR"), np.ones_like(o.get("RHOBAR")))
    assert np.max(np.abs(o.get("THETABAR") / th0 - 1.0)) <= 1e-15
    assert np.max(np.abs(o.get("PBAR") / th0 - 1.0)) <= 1e-15
    n2 = (cfg.buoyancy_frequency * TREF) ** 2 if stable else 0.0
    if stable:
        assert n2 > 0.0
    assert np.max(np.abs(o.get("N2") - n2)) <= 1e-14 * max(n2, 1.0)
