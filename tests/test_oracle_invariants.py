"""Pins of the CPU oracle beyond the reference's norm fixtures.

The reference's own fixtures (test_oracle_golden.py) check the whole FP64 path
at the 3.45e-4 level.  The checks below pin individual oracle functions much
more tightly. Each uses mathematics that fixes the result independently of the
oracle's code:

* compute_operator! (compute_operator!.jl:40-1606) against the short
  corrector: for any s, compute_lhs!(u) - compute_lhs!(correct!(u; dpip =
  s / (sqrt(pbar^2/rhobar) dt))) = A s.  This ties every coefficient,
  including the special levels next to the walls, to compute_lhs!.jl:68-160
  and correct!.jl:164-385;
* the null space of A (constant pressure has no gradient) and its left null
  space (the J sqrt(pbar^2/rhobar)-weighted sum of any divergence vanishes);
* apply_operator! Total = Horizontal + the (ac, ad, au) column part;
* apply_preconditioner! (apply_preconditioner!.jl:40-124) against a banded
  LAPACK solve of (1 - deta L_v) and the horizontal apply;
* apply_bicgstab! (apply_bicgstab!.jl:24-150) against a textbook
  right-preconditioned BiCGSTAB (van der Vorst 1992) on the densely assembled
  A and M^-1;
* apply_1d_muscl! (apply_1d_muscl!.jl:51-93) against the properties of the
  kappa = 1/3 MUSCL scheme it implements;
* compute_stress_tensor (compute_stress_tensor.jl:55-164) against the
  deviatoric strain rate 2 S - 2/3 div I of linear flows on a flat grid;
* compute_lhs! against the discrete Gauss theorem, the LHS flux divergence
  against finite-volume conservation and free-stream preservation.
"""
import numpy as np
import pytest
from scipy.linalg import solve_banded


def make(pcf, oracle_mod, nl, **over):
    cfg = pcf.to_config(nl)
    for k, v in over.items():
        setattr(cfg, k, v)
    return oracle_mod.Oracle(cfg, pcf._abi)


def inner(o, a):
    return a[3:3 + o.nx, 3:3 + o.ny, 3:3 + o.nz]


def cases(pcf):
    return {
        "periodic_hill": pcf.examples.periodic_hill(x_size=12, z_size=10),
        "mountain_wave": pcf.examples.mountain_wave(x_size=8, y_size=6, z_size=9),
        "mountain_wave_3d": pcf.examples.mountain_wave_3d(x_size=10, y_size=8, z_size=12, iterations=1),
        "cold_bubble": pcf.examples.cold_bubble(x_size=10, z_size=8),
    }


def flat_boussinesq(pcf, u0=10.0):
    N = pcf.namelists
    return N.Namelists(
        atmosphere=N.AtmosphereNamelist(model="Boussinesq", background="StableStratification",
                                        coriolis_frequency=0.0, initial_u=u0),
        domain=N.DomainNamelist(x_size=8, y_size=6, z_size=10, lx=8000.0, ly=9000.0, lz=12000.0))


CASES = ["periodic_hill", "mountain_wave", "mountain_wave_3d", "cold_bubble"]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("damp", [0, 1])
def test_operator_equals_corrector_divergence(pcf, oracle_mod, name, damp):
    o = make(pcf, oracle_mod, cases(pcf)[name], damp_horizontal_wind_on_rhs=damp)
    dt, rf = 0.37, 1.3
    o.compute_sponges(dt, 0.1)
    o.compute_operator(dt, rf)
    rng = np.random.default_rng(11)
    s = rng.standard_normal((o.nx, o.ny, o.nz))
    As = o.apply_operator(s)
    o.compute_lhs()
    b0 = o.get_vector("RHS")
    o.set_vector("SOLUTION", s)
    o.solution_to_dpip(dt)
    o.correct(dt, rf)
    o.compute_lhs()
    b1 = o.get_vector("RHS")
    err = np.abs((b0 - b1) - As).max() / np.abs(As).max()
    assert err < 1e-13, err


@pytest.mark.parametrize("name", CASES)
def test_operator_null_spaces(pcf, oracle_mod, name):
    o = make(pcf, oracle_mod, cases(pcf)[name])
    o.compute_sponges(0.2, 0.1)
    o.compute_operator(0.2, 1.0)
    fcs = np.sqrt(inner(o, o.get("PBAR")) ** 2 / inner(o, o.get("RHOBAR")))
    jac = inner(o, o.get("JAC"))
    rng = np.random.default_rng(5)
    s = rng.standard_normal((o.nx, o.ny, o.nz))
    As = o.apply_operator(s)
    gain = np.abs(As).max() / np.abs(s).max()
    # constant pressure correction: zero gradient, zero divergence
    assert np.abs(o.apply_operator(fcs)).max() / (gain * np.abs(fcs).max()) < 1e-13
    # the weighted volume sum of any discrete divergence is zero
    w = jac * fcs
    assert abs((w * As).sum()) / np.abs(w * As).sum() < 1e-13


@pytest.mark.parametrize("name", CASES)
def test_total_is_horizontal_plus_column(pcf, oracle_mod, name):
    o = make(pcf, oracle_mod, cases(pcf)[name])
    o.compute_sponges(0.2, 0.1)
    o.compute_operator(0.2, 1.0)
    s = np.random.default_rng(2).standard_normal((o.nx, o.ny, o.nz))
    zero = np.zeros_like(s[:, :, :1])
    s_up = np.concatenate([s[:, :, 1:], zero], axis=2)     # zero beyond the top wall
    s_dn = np.concatenate([zero, s[:, :, :-1]], axis=2)    # zero below the ground
    col = o.get_vector("AC") * s + o.get_vector("AD") * s_dn + o.get_vector("AU") * s_up
    tot = o.apply_operator(s, total=True)
    hor = o.apply_operator(s, total=False)
    assert np.abs(tot - (hor + col)).max() / np.abs(tot).max() < 1e-14


@pytest.mark.parametrize("iterations", [1, 2, 3])
def test_preconditioner_is_pseudo_time_line_solves(pcf, oracle_mod, iterations):
    nl = cases(pcf)["mountain_wave_3d"]
    o = make(pcf, oracle_mod, nl, preconditioner_iterations=iterations, dtau=0.8)
    o.compute_sponges(0.2, 0.1)
    o.compute_operator(0.37, 1.0)
    c = o.constants()
    deta = 0.8 / (2 * (1 / c["dx"] ** 2 + 1 / c["dy"] ** 2))
    ac, ad, au = o.get_vector("AC"), o.get_vector("AD"), o.get_vector("AU")
    b = np.random.default_rng(4).standard_normal((o.nx, o.ny, o.nz))
    s = np.zeros_like(b)
    for _ in range(iterations):
        rhs = s + deta * (o.apply_operator(s, total=False) - b)
        s = np.empty_like(b)
        for i in range(o.nx):
            for j in range(o.ny):
                band = np.zeros((3, o.nz))
                band[0, 1:] = -deta * au[i, j, :-1]
                band[1, :] = 1 - deta * ac[i, j, :]
                band[2, :-1] = -deta * ad[i, j, 1:]
                s[i, j] = solve_banded((1, 1), band, rhs[i, j])
    out = o.apply_preconditioner(b)
    assert np.abs(out - s).max() / np.abs(s).max() < 1e-13


def _dense(o, fn):
    shape = (o.nx, o.ny, o.nz)
    n = o.nx * o.ny * o.nz
    cols = []
    for q in range(n):
        e = np.zeros(n)
        e[q] = 1.0
        cols.append(fn(e.reshape(shape, order="F")).ravel(order="F"))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("name,size", [("mountain_wave_3d", (5, 4, 8)), ("periodic_hill", (6, 1, 8))])
def test_bicgstab_matches_textbook(pcf, oracle_mod, name, size):
    if name == "periodic_hill":
        nl = pcf.examples.periodic_hill(x_size=size[0], z_size=size[2])
    else:
        nl = pcf.examples.mountain_wave_3d(x_size=size[0], y_size=size[1], z_size=size[2], iterations=1)
    o = make(pcf, oracle_mod, nl)
    o.compute_sponges(0.2, 0.1)
    o.compute_operator(0.37, 1.0)
    A = _dense(o, o.apply_operator)
    M = _dense(o, o.apply_preconditioner)
    n = A.shape[0]
    # A is singular with the one-dimensional null space of constant pressure
    sv = np.linalg.svd(A, compute_uv=False)
    assert sv[-1] / sv[0] < 1e-13 and sv[-2] / sv[0] > 1e-6
    b = A @ np.random.default_rng(3).standard_normal(n)
    o.set_vector("RHS", b.reshape((o.nx, o.ny, o.nz), order="F"))
    rc, niter = o.apply_bicgstab(1.0)  # tolref = 1: criterion ||r|| / ||b|| <= 1e-8
    assert rc == 0
    x_o = o.get_vector("SOLUTION").ravel(order="F")

    def rms(a):
        return np.sqrt(a @ a / n)

    x = np.zeros(n)
    r = b.copy()
    r_hat = b.copy()
    p = np.zeros(n)
    v = np.zeros(n)
    rho_old = alpha = omega = 1.0
    k = 0
    while k < 1000:
        k += 1
        rho = r_hat @ r
        p = r + (rho / rho_old) * (alpha / omega) * (p - omega * v)
        y = M @ p
        v = A @ y
        alpha = rho / (r_hat @ v)
        s = r - alpha * v
        z = M @ s
        t = A @ z
        omega = (t @ s) / (t @ t)
        x = x + alpha * y + omega * z
        r = s - omega * t
        rho_old = rho
        if rms(r) / rms(b) <= 1e-8:
            break
    assert niter == k
    assert np.abs(x_o - x).max() / np.abs(x).max() < 1e-6
    assert rms(A @ x_o - b) / rms(b) <= 1.05e-8


def test_muscl_kappa_one_third_properties(oracle_mod):
    n = 12
    # constant data stays constant; the two end cells hold the placeholder 1000
    left, right = oracle_mod.muscl_line(np.full(n, 3.7))
    assert np.all(left[1:-1] == 3.7) and np.all(right[1:-1] == 3.7)
    assert left[0] == right[0] == left[-1] == right[-1] == 1000.0
    # linear data: exact face values
    xc = np.arange(n, dtype=float)
    left, right = oracle_mod.muscl_line(2.0 + 0.5 * xc)
    assert np.allclose(right[1:-1], 2.0 + 0.5 * (xc[1:-1] + 0.5), rtol=0, atol=1e-14)
    assert np.allclose(left[1:-1], 2.0 + 0.5 * (xc[1:-1] - 0.5), rtol=0, atol=1e-14)
    # cell averages of an increasing quadratic: the kappa = 1/3 scheme is exact
    # where the limiter is inactive (slope ratio between 0.4 and 2.5; here 0.82-0.97)
    xq = 5.0 + xc
    avg = xq ** 2 + 1.0 / 12.0
    left, right = oracle_mod.muscl_line(avg)
    assert np.allclose(right[1:-1], (xq[1:-1] + 0.5) ** 2, rtol=1e-14)
    assert np.allclose(left[1:-1], (xq[1:-1] - 0.5) ** 2, rtol=1e-14)
    # a step: no new extrema (TVD); at a local extremum the scheme is first order
    step = np.where(xc < 6, 1.0, 2.0)
    left, right = oracle_mod.muscl_line(step)
    assert np.all((left[1:-1] >= 1.0) & (left[1:-1] <= 2.0) & (right[1:-1] >= 1.0) & (right[1:-1] <= 2.0))
    bump = np.zeros(n)
    bump[5] = 1.0
    left, right = oracle_mod.muscl_line(bump)
    assert left[5] == right[5] == 1.0
    # slope ratio 6: the right-face slope saturates at 2 (no overshoot of the
    # downwind neighbour), the left-face slope at 2/ratio (TVD bound 2 theta)
    left, right = oracle_mod.muscl_line(np.array([0.0, 6.0, 7.0, 8.0]))
    assert right[1] == 7.0
    assert abs(left[1] - 5.0) < 1e-14


@pytest.mark.parametrize("flow,expect", [
    ("u=a*z", {(1, 3): 1.0}),
    ("u=a*y", {(1, 2): 1.0}),
    ("v=a*x", {(1, 2): 1.0}),
    ("v=a*z", {(2, 3): 1.0}),
    ("w=a*x", {(1, 3): 1.0}),
    ("u=a*x", {(1, 1): 4.0 / 3.0, (2, 2): -2.0 / 3.0, (3, 3): -2.0 / 3.0}),
    ("w=a*z", {(1, 1): -2.0 / 3.0, (2, 2): -2.0 / 3.0, (3, 3): 4.0 / 3.0}),
])
def test_stress_tensor_of_linear_flows(pcf, oracle_mod, flow, expect):
    o = make(pcf, oracle_mod, flat_boussinesq(pcf, u0=0.0))
    c = o.constants()
    assert np.abs(o.get("MET13")).max() == 0.0 and np.allclose(o.get("JAC"), 1.0, rtol=0, atol=1e-13)
    xs = np.empty(o.nxx)
    ys = np.empty(o.nyy)
    oracle_mod.lib().orc_grid_1d(o.h, oracle_mod._ptr(xs), oracle_mod._ptr(ys))
    zc = o.get("ZC")
    shape = (o.nxx, o.nyy, o.nzz)
    X = np.broadcast_to(xs[:, None, None], shape)
    Y = np.broadcast_to(ys[None, :, None], shape)
    a = 0.3
    fields = {"U": np.zeros(shape), "V": np.zeros(shape), "W": np.zeros(shape)}
    var, rhs = flow.split("=")
    coord = {"x": X, "y": Y, "z": zc}[rhs[-1]]
    # staggering: u sits at x + dx/2, v at y + dy/2, w at z + dz/2 (flat grid)
    off = {("u", "x"): c["dx"] / 2, ("v", "y"): c["dy"] / 2, ("w", "z"): c["dz"] / 2}.get((var, rhs[-1]), 0.0)
    fields[var.upper()] = a * (coord + off)
    for k_, f in fields.items():
        o.set(k_, np.asfortranarray(f))
    for mu in (1, 2, 3):
        for nu in range(mu, 4):
            want = a * expect.get((mu, nu), 0.0)
            got = o.stress_tensor(6, 5, 7, mu, nu)
            assert abs(got - want) < 1e-12 * a, (mu, nu, got, want)


@pytest.mark.parametrize("name", CASES)
def test_discrete_gauss(pcf, oracle_mod, name):
    o = make(pcf, oracle_mod, cases(pcf)[name])
    rng = np.random.default_rng(8)
    for f in ("U", "V", "W"):
        o.set(f, o.get(f) + rng.standard_normal((o.nxx, o.nyy, o.nzz)))
    o.set_boundaries(0)
    o.compute_lhs()
    b = o.get_vector("RHS")
    w = inner(o, o.get("JAC")) * np.sqrt(inner(o, o.get("PBAR")) ** 2 / inner(o, o.get("RHOBAR")))
    assert abs((w * b).sum()) / np.abs(w * b).sum() < 1e-13


@pytest.mark.parametrize("name", ["mountain_wave", "mountain_wave_3d", "cold_bubble"])
@pytest.mark.parametrize("var", ["RHO", "RHOP"])
def test_flux_divergence_conserves_mass(pcf, oracle_mod, name, var):
    abi = pcf._abi
    o = make(pcf, oracle_mod, cases(pcf)[name])
    rng = np.random.default_rng(9)
    for f in ("RHOP", "U", "V", "W"):
        a = o.get(f)
        o.set(f, a + 0.1 * max(np.abs(a).max(), 1e-3) * rng.standard_normal(a.shape))
    o.set_boundaries(0)
    o.backup_predictands()
    jac = inner(o, o.get("JAC"))
    before = inner(o, o.get(var)).copy()
    o.compute_fluxes()
    o.update_lhs(0.05, 1, abi.VAR_RHO if var == "RHO" else abi.VAR_RHOP)
    delta = inner(o, o.get(var)) - before
    assert np.abs(delta).max() > 0
    assert abs((jac * delta).sum()) / np.abs(jac * delta).sum() < 1e-12


def test_free_stream_is_preserved(pcf, oracle_mod):
    o = make(pcf, oracle_mod, flat_boussinesq(pcf))
    before = {f: o.get(f) for f in ("RHOP", "U", "V", "W")}
    o.backup_predictands()
    for _ in range(3):
        o.explicit_integration_lhs(0.05, 0.0)
    for f, a in before.items():
        assert np.abs(o.get(f) - a).max() <= 1e-13 * max(np.abs(a).max(), 1.0), f


PHI = ["PHIRHOP_X", "PHIRHOP_Y", "PHIRHOP_Z", "PHIU_X", "PHIU_Y", "PHIU_Z",
       "PHIV_X", "PHIV_Y", "PHIV_Z", "PHIW_X", "PHIW_Y", "PHIW_Z"]


@pytest.mark.parametrize("flow,viscous", [
    # u = a z: stress (1,3) = a.  The meridional W flux of the reference uses
    # the (3,1) component (compute_fluxes!.jl:1351-1367) where (3,2) would be
    # expected, so it carries the shear too.
    ("U", {"PHIU_Z", "PHIW_X", "PHIW_Y"}),
    # v = a z: stress (2,3) = a; the meridional W flux sees (3,1) = 0.
    ("V", {"PHIV_Z"}),
])
def test_viscous_fluxes_of_linear_shear(pcf, oracle_mod, flow, viscous):
    """The viscous part of compute_fluxes! (compute_fluxes!.jl:694-781,
    998-1085, 1320-1391) for a linear shear on a flat grid: the flux of
    the sheared momentum component is -(1/Re) rhobar(k0) a on every interior
    face, every other flux is inviscid.  Two runs that differ only in the
    viscosity isolate the viscous part."""
    a = 0.3
    out, inv_re = [], []
    for nu_kin in (1.5e-5, 40.0):
        nl = flat_boussinesq(pcf, u0=0.0)
        nl.atmosphere.kinematic_viscosity = nu_kin
        o = make(pcf, oracle_mod, nl)
        zc = o.get("ZC")
        f = np.zeros((o.nxx, o.nyy, o.nzz))
        for name in ("U", "V", "W"):
            o.set(name, a * zc if name == flow else f)
        o.backup_predictands()
        o.compute_fluxes()
        out.append({n: o.get(n) for n in PHI})
        inv_re.append(1.0 / o.constants()["re"])
        rhobar_k0 = o.get("RHOBAR")[3, 3, 3]
    expected = -(inv_re[1] - inv_re[0]) * rhobar_k0 * a
    assert abs(expected) > 1e-6
    core = (slice(3, 3 + o.nx), slice(3, 3 + o.ny), slice(3, 3 + o.nz - 1))
    for n in PHI:
        d = (out[1][n] - out[0][n])[core]
        want = expected if n in viscous else 0.0
        assert np.abs(d - want).max() <= 1e-10 * abs(expected), (n, np.abs(d - want).max())


def test_multigrid_reference_accelerates_bicgstab(pcf, oracle_mod):
    """The FP64 numpy restatement of the multigrid preconditioner (used to
    check the GPU V-cycle) cuts BiCGSTAB's iterations against the
    reference's line-relaxation preconditioner on the same operator."""
    import mg_reference as mg

    nl = pcf.examples.mountain_wave_3d(x_size=32, y_size=32, z_size=16, iterations=1)
    o = make(pcf, oracle_mod, nl)
    o.compute_sponges(0.05, 0.1)
    o.compute_operator(0.05, 1.0)
    C = {n: o.get_vector(n) for n in mg.NAMES}
    c = o.constants()
    b = o.apply_operator(np.random.default_rng(12).standard_normal((o.nx, o.ny, o.nz)))
    s = np.random.default_rng(13).standard_normal(b.shape)
    assert np.abs(mg.apply(C, s) - o.apply_operator(s)).max() <= 1e-13 * np.abs(o.apply_operator(s)).max()
    lv = mg.levels(C, c["dx"], c["dy"])
    assert [L["C"]["AC"].shape[:2] for L in lv] == [(32, 32), (16, 16), (8, 8), (4, 4)]
    # one smoothing step from zero on the finest level is one pseudo-time
    # step of the reference preconditioner
    one = mg.smooth(lv[0], np.zeros_like(b), b, 2)
    assert np.abs(one - o.apply_preconditioner(b)).max() <= 1e-12 * np.abs(one).max()

    def bicgstab(M, tol=1e-8):
        rms = lambda v: np.sqrt(np.mean(v * v))  # noqa: E731
        x, r, p = np.zeros_like(b), b.copy(), b.copy()
        r0, rho = b.copy(), float(np.vdot(b, b))
        for k in range(1, 500):
            y = M(p)
            v = o.apply_operator(y)
            alpha = rho / float(np.vdot(v, r0))
            s_ = r - alpha * v
            z = M(s_)
            t = o.apply_operator(z)
            omega = float(np.vdot(t, s_)) / float(np.vdot(t, t))
            x, r = x + alpha * y + omega * z, s_ - omega * t
            if rms(r) / rms(b) <= tol:
                return x, k
            rho_n = float(np.vdot(r, r0))
            p = r + rho_n / rho * alpha / omega * (p - omega * v)
            rho = rho_n
        raise AssertionError("no convergence")

    _, k_line = bicgstab(o.apply_preconditioner)
    x, k_mg = bicgstab(lambda v: mg.vcycle(lv, v))
    assert np.sqrt(np.mean((b - o.apply_operator(x)) ** 2)) <= 1.1e-8 * np.sqrt(np.mean(b ** 2))
    assert k_mg * 2 < k_line, (k_mg, k_line)
