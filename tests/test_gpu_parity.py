"""Parity of the HIP library with the CPU oracle, kernel by kernel and step by step.

The library evaluates the reference expressions in the reference's operation
order with FP64 and without contraction, like the oracle; kernels without a
reduction are expected to agree to the last bit.  Tolerances below:

* per-kernel outputs (operator coefficients and applies, preconditioner,
  fluxes, updates, corrector): relative 1e-13 of the field's max-norm;
* anything downstream of a global reduction: the reduction order differs,
  so BiCGSTAB iterates differ in the last bits and a solve may stop a few
  iterations earlier or later.  A solve must meet the reference criterion
  ||b - A x|| / ||b|| <= tol, checked with the oracle's operator, and the two
  solutions may differ only by what both criteria allow:
  ||A (x_gpu - x_oracle)|| <= 2 tol ||b||.  Stages and whole steps are
  compared with the Poisson tolerance tightened to 1e-12 (relative), where
  the fields agree to 1e-9 relative.  Iteration counts may differ by at most
  iter_slack(n), the spread that reordering the oracle's own sums gives.
"""
import hipmem
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KERNEL_RTOL = 1e-13
SOLVE_RTOL = 1e-9
STEP_RTOL = 1e-9


def rel(a, b):
    if b.size == 0:  # e.g. interior W faces of a one-layer domain
        return 0.0
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def sl(a, b):
    """1-based inclusive index range -> 0-based slice."""
    return slice(a - 1, b)


def configs(pcf):
    ph = pcf.examples.periodic_hill(x_size=24, z_size=20)
    mw = pcf.examples.mountain_wave(x_size=12, y_size=10, z_size=9)
    mw3 = pcf.examples.mountain_wave_3d(x_size=16, y_size=14, z_size=12, iterations=2)
    cb = pcf.examples.cold_bubble(x_size=12, z_size=10)
    # odd x_size: the one-cell-per-thread stencil kernels (even x_size uses two)
    mwo = pcf.examples.mountain_wave(x_size=11, y_size=6, z_size=9)
    # the tracer: the vortex's (Boussinesq) and a dipole in the
    # pseudo-incompressible mountain wave, where chi carries rho + rhobar
    vx = pcf.examples.vortex(x_size=12, y_size=10)
    mwt = pcf.examples.mountain_wave_3d(x_size=16, y_size=14, z_size=12, iterations=2)
    mwt.tracer = pcf.TracerNamelist(tracer_setup="TracerOn", initial_tracer=pcf.TracerDipole(50000.0, 50000.0))
    # the compressible model: the hot bubble, and the mountain wave (terrain,
    # LHS sponge) with a conductivity large enough for the heat conduction to
    # matter and the tracer on
    hb = pcf.examples.hot_bubble(x_size=12, z_size=10)
    mwc = pcf.examples.mountain_wave(x_size=12, y_size=10, z_size=9)
    mwc.atmosphere.model = "Compressible"
    mwc.atmosphere.thermal_conductivity = 50.0
    mwc.tracer = pcf.TracerNamelist(tracer_setup="TracerOn", initial_tracer=pcf.TracerDipole(5000.0, 5000.0))
    # columns taller than the 80 levels the FP64 line solve keeps in LDS
    mwz = pcf.examples.mountain_wave_3d(x_size=16, y_size=14, z_size=136, iterations=2)
    # the two sponge options off in every example: the LHS sponge relaxing
    # the winds to their horizontal layer means (apply_lhs_sponge!.jl:365-530)
    # and the RHS sponge's Rayleigh damping of the horizontal winds in the
    # implicit updates and the operator (update!.jl:846-881,
    # compute_operator!.jl:283-378)
    mwr = pcf.examples.mountain_wave(x_size=12, y_size=10, z_size=9)
    mwr.sponge.relax_to_mean = True
    phd = pcf.examples.periodic_hill(x_size=24, z_size=20)
    phd.sponge.damp_horizontal_wind_on_rhs = True
    return {"periodic_hill": ph, "mountain_wave": mw, "mountain_wave_3d": mw3, "cold_bubble": cb,
            "mountain_wave_odd": mwo, "vortex": vx, "mountain_wave_tracer": mwt, "hot_bubble": hb,
            "mountain_wave_compressible": mwc, "mountain_wave_tall": mwz, "mountain_wave_relax": mwr,
            "periodic_hill_damp": phd}


CASES = ["periodic_hill", "mountain_wave", "mountain_wave_3d", "cold_bubble", "mountain_wave_odd", "vortex",
         "mountain_wave_tracer", "hot_bubble", "mountain_wave_compressible", "mountain_wave_relax",
         "periodic_hill_damp"]

# BiCGSTAB iterates of the device and the oracle differ only through the
# order of the dot products' sums (1e-15 relative); near the tolerance that
# moves the count at which a solve stops, by as much as reordering the
# oracle's own sums moves it (test_oracle_reduction_order.py measures the
# spread and sets the bound)
from test_oracle_reduction_order import iter_slack  # noqa: E402


def compressible(nl):
    return nl.atmosphere.model == "Compressible"


def predictands(nl, pip=True):
    """The predictands of the case (rho in the pseudo-incompressible model,
    chi with a tracer)."""
    names = ["RHOP", "U", "V", "W"] + (["PIP"] if pip else [])
    if nl.atmosphere.model != "Boussinesq":
        names.append("RHO")
    if nl.tracer.tracer_setup == "TracerOn":
        names.append("CHI")
    if compressible(nl):
        names.append("P")
    return names


def pair(pcf, abi, oracle_mod, name, tight=False, flux_precision="fp64"):
    nl = configs(pcf)[name]
    nl.discretization.flux_precision = flux_precision  # the oracle's arithmetic is FP64 whatever it says
    if tight:
        nl.poisson.tolerance = 1e-12
        nl.poisson.tolerance_is_relative = True
    gpu = pcf.State(nl)
    orc = oracle_mod.Oracle(pcf.to_config(nl), abi)
    return nl, gpu, orc


def perturb(gpu, orc, names, seed, amp=0.05):
    """Identical random perturbations of haloed fields on both sides."""
    rng = np.random.default_rng(seed)
    for n in names:
        a = orc.get(n)
        scale = max(float(np.max(np.abs(a))), 1e-3)
        a = a + amp * scale * rng.standard_normal(a.shape)
        gpu.set(n, a)
        orc.set(n, a)


@pytest.mark.parametrize("name", CASES)
def test_setup_fields(pcf, abi, oracle_mod, name):
    nl, gpu, orc = pair(pcf, abi, oracle_mod, name)
    tracer = ("CHI", "INITIALTRACER") if nl.tracer.tracer_setup == "TracerOn" else ()
    comp = ("P",) if compressible(nl) else ()
    for f in ("PBAR", "THETABAR", "RHOBAR", "N2", "JAC", "MET13", "MET23", "MET33", "ZC", "ZCTILDE",
              "RHO", "RHOP", "U", "V", "W", "PIP") + tracer + comp:
        a, b = gpu.get(f), orc.get(f)
        assert np.array_equal(a, b), (f, float(np.max(np.abs(a - b))))


@pytest.mark.parametrize("name", CASES)
def test_sponges(pcf, abi, oracle_mod, name):
    _, gpu, orc = pair(pcf, abi, oracle_mod, name)
    for s in (gpu, orc):
        s.compute_sponges(0.37, 5.0)
    for f in ("ALPHAR", "BETAR"):
        a, b = gpu.get(f), orc.get(f)
        assert np.array_equal(a, b), (f, float(np.max(np.abs(a - b))))


@pytest.mark.parametrize("name", CASES)
def test_compute_lhs(pcf, abi, oracle_mod, name):
    nl, gpu, orc = pair(pcf, abi, oracle_mod, name)
    perturb(gpu, orc, ["U", "V", "W"], seed=1)
    if compressible(nl):  # the heating term: conduction fluxes of a perturbed p0
        perturb(gpu, orc, ["RHO"], seed=11, amp=0.01)
        for s in (gpu, orc):
            s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
            s.backup_predictands()
            s.compute_theta_fluxes()
            s.set_boundaries(abi.BOUNDARY_FLUXES)
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
    t1, t2 = gpu.compute_lhs(), orc.compute_lhs()
    err = rel(gpu.get_vector("RHS"), orc.get_vector("RHS"))
    assert err <= KERNEL_RTOL, err
    assert t1 == pytest.approx(t2, rel=1e-12)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("dt,rf", [(1.0, 1.0), (0.37, 2.0)])
def test_compute_operator(pcf, abi, oracle_mod, name, dt, rf):
    nl, gpu, orc = pair(pcf, abi, oracle_mod, name)
    perturb(gpu, orc, ["RHO"] + (["P"] if compressible(nl) else []), seed=2, amp=0.01)
    for s in (gpu, orc):
        s.compute_sponges(0.5, 3.0)
        s.compute_operator(dt, rf)
    for v in abi.VECTORS[2:]:
        err = rel(gpu.get_vector(v), orc.get_vector(v))
        assert err <= KERNEL_RTOL, (v, err)


@pytest.mark.parametrize("name", CASES)
def test_apply_operator_and_preconditioner(pcf, abi, oracle_mod, name):
    _, gpu, orc = pair(pcf, abi, oracle_mod, name)
    for s in (gpu, orc):
        s.compute_sponges(0.5, 3.0)
        s.compute_operator(0.8, 1.0)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((gpu.nx, gpu.ny, gpu.nz))
    dx = hipmem.DeviceArray(x.shape).upload(x)
    dy = hipmem.DeviceArray(x.shape)
    for total in (True, False):
        gpu.apply_operator_device(dx.ptr, dy.ptr, total)
        err = rel(dy.download(), orc.apply_operator(x, total))
        assert err <= KERNEL_RTOL, (total, err)
    gpu.apply_preconditioner_device(dx.ptr, dy.ptr)
    err = rel(dy.download(), orc.apply_preconditioner(x))
    assert err <= KERNEL_RTOL, err
    # the line-solve factors follow coefficients changed through the ABI
    for vec in ("AC", "AU"):
        a = orc.get_vector(vec) * (1.0 + 0.1 * rng.standard_normal(x.shape))
        gpu.set_vector(vec, a)
        orc.set_vector(vec, a)
        gpu.apply_preconditioner_device(dx.ptr, dy.ptr)
        err = rel(dy.download(), orc.apply_preconditioner(x))
        assert err <= KERNEL_RTOL, (vec, err)
    # compute_global_dot_product on the device
    got = gpu.dot_device(dx.ptr, dx.ptr)
    assert got == pytest.approx(float(np.sum(x * x)), rel=1e-13)


def test_preconditioner_tall_columns(pcf, abi, oracle_mod):
    """136 levels: the line solve's upward sweep of levels 80 and above goes
    through HBM, the lower ones through LDS; both parts against the oracle."""
    _, gpu, orc = pair(pcf, abi, oracle_mod, "mountain_wave_tall")
    for s in (gpu, orc):
        s.compute_sponges(0.5, 3.0)
        s.compute_operator(0.8, 1.0)
    x = np.random.default_rng(11).standard_normal((gpu.nx, gpu.ny, gpu.nz))
    dx = hipmem.DeviceArray(x.shape).upload(x)
    dy = hipmem.DeviceArray(x.shape)
    gpu.apply_preconditioner_device(dx.ptr, dy.ptr)
    got, want = dy.download(), orc.apply_preconditioner(x)
    for half in (slice(0, 80), slice(80, None)):
        err = rel(got[:, :, half], want[:, :, half])
        assert err <= KERNEL_RTOL, (half, err)


def test_preconditioner_many_columns(pcf, abi, oracle_mod):
    """1032 blocks of 64 columns, more than four per CU: the line solve keeps
    64 levels in LDS (five blocks per CU) instead of 80; 72 levels, so both
    parts of a column are checked against the oracle."""
    nl = pcf.examples.mountain_wave_3d(x_size=64, y_size=1032, z_size=72, iterations=1)
    gpu = pcf.State(nl)
    orc = oracle_mod.Oracle(pcf.to_config(nl), abi)
    for s in (gpu, orc):
        s.compute_sponges(0.5, 3.0)
        s.compute_operator(0.8, 1.0)
    x = np.random.default_rng(12).standard_normal((gpu.nx, gpu.ny, gpu.nz))
    dx = hipmem.DeviceArray(x.shape).upload(x)
    dy = hipmem.DeviceArray(x.shape)
    gpu.apply_preconditioner_device(dx.ptr, dy.ptr)
    got, want = dy.download(), orc.apply_preconditioner(x)
    for half in (slice(0, 64), slice(64, None)):
        err = rel(got[:, :, half], want[:, :, half])
        assert err <= KERNEL_RTOL, (half, err)


@pytest.mark.parametrize("name", CASES)
def test_bicgstab(pcf, abi, oracle_mod, name):
    _, gpu, orc = pair(pcf, abi, oracle_mod, name)
    perturb(gpu, orc, ["U", "V", "W"], seed=4)
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
        s.compute_sponges(0.4, 1.0)
    tolref = gpu.compute_lhs()
    orc.compute_lhs()
    for s in (gpu, orc):
        s.compute_operator(0.4, 1.0)
    n1 = gpu.apply_bicgstab(tolref)
    rc, n2 = orc.apply_bicgstab(tolref)
    assert rc == 0
    assert abs(n1 - n2) <= iter_slack(n2), (n1, n2)
    tol = 1e-8 / tolref
    b = orc.get_vector("RHS")
    xg, xo = gpu.get_vector("SOLUTION"), orc.get_vector("SOLUTION")
    rms = lambda v: float(np.sqrt(np.mean(v * v)))
    true_res = rms(b - orc.apply_operator(xg)) / rms(b)
    assert true_res <= tol * (1 + 1e-6), (true_res, tol)
    assert gpu.last_residual() == pytest.approx(true_res, rel=1e-6)
    gap = rms(orc.apply_operator(xg - xo)) / rms(b)
    assert gap <= 2 * tol, (gap, tol)
    if n1 == n2:
        err = rel(xg, xo)
        assert err <= SOLVE_RTOL, err


@pytest.mark.parametrize("name", CASES)
def test_fluxes(pcf, abi, oracle_mod, name):
    nl, gpu, orc = pair(pcf, abi, oracle_mod, name)
    names = predictands(nl, pip=False)
    perturb(gpu, orc, names, seed=5)
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
        s.backup_predictands()
    perturb(gpu, orc, ["U", "V", "W"], seed=6, amp=0.02)  # current winds differ from p0
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
        s.compute_theta_fluxes()  # before compute_fluxes, which zeroes its wall faces
    # the reference's sequence (explicit_integration!.jl:53-58) on the device,
    # the oracle's compute_fluxes runs all three of its steps
    gpu.reconstruct()
    gpu.set_boundaries(abi.BOUNDARY_RECONSTRUCTIONS)
    for s in (gpu, orc):
        s.compute_fluxes()
    nb = 3
    i0, j0, k0 = nb + 1, nb + 1, nb + 1
    i1, j1, k1 = i0 + gpu.nx - 1, j0 + gpu.ny - 1, k0 + gpu.nz - 1
    # faces the updates read (update!.jl loop ranges)
    rng = {
        "X": (sl(i0 - 1, i1), sl(j0, j1), sl(k0, k1)),
        "Y": (sl(i0, i1), sl(j0 - 1, j1), sl(k0, k1)),
        "Z": (sl(i0, i1), sl(j0, j1), sl(k0 - 1, k1)),
    }
    special = {
        "PHIU_X": (sl(i0 - 2, i1), sl(j0, j1), sl(k0, k1)),
        "PHIU_Y": (sl(i0 - 1, i1), sl(j0 - 1, j1), sl(k0, k1)),
        "PHIU_Z": (sl(i0 - 1, i1), sl(j0, j1), sl(k0 - 1, k1)),
        "PHIV_X": (sl(i0 - 1, i1), sl(j0 - 1, j1), sl(k0, k1)),
        "PHIV_Y": (sl(i0, i1), sl(j0 - 2, j1), sl(k0, k1)),
        "PHIV_Z": (sl(i0, i1), sl(j0 - 1, j1), sl(k0 - 1, k1)),
        "PHIW_X": (sl(i0 - 1, i1), sl(j0, j1), sl(k0, k1 - 1)),
        "PHIW_Y": (sl(i0, i1), sl(j0 - 1, j1), sl(k0, k1 - 1)),
        "PHIW_Z": (sl(i0, i1), sl(j0, j1), sl(k0 - 1, k1)),
    }
    fl = ["PHIRHOP", "PHIU", "PHIV", "PHIW"] + (["PHIRHO"] if "RHO" in names else []) + \
        (["PHICHI"] if "CHI" in names else []) + (["PHIP", "PHITHETA"] if compressible(nl) else [])
    for f in fl:
        for d in "XYZ":
            key = f"{f}_{d}"
            r = special.get(key, rng[d])
            a, b = gpu.get(key)[r], orc.get(key)[r]
            err = rel(a, b)
            assert err <= KERNEL_RTOL, (key, err)


@pytest.mark.parametrize("name", CASES)
def test_explicit_stages(pcf, abi, oracle_mod, name):
    """explicit_integration! LHS (3 RK stages with sponges) and RHS on perturbed states."""
    nl, gpu, orc = pair(pcf, abi, oracle_mod, name)
    names = predictands(nl)
    perturb(gpu, orc, names, seed=7, amp=0.03)
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
        s.compute_sponges(0.3, 2.0)
        s.backup_predictands()
        s.compute_theta_fluxes()
        s.explicit_integration_lhs(0.3, 2.0)
    for f in names:
        err = rel(gpu.get(f), orc.get(f))
        assert err <= KERNEL_RTOL, ("lhs", f, err)
    for s in (gpu, orc):
        s.explicit_integration_rhs(0.3, 2.0)
    for f in names:
        err = rel(gpu.get(f), orc.get(f))
        assert err <= KERNEL_RTOL, ("rhs", f, err)


# FP32 fluxes (flux_precision = "mixed") against the FP64 oracle: one LHS
# integration leaves relative errors of at most 2.7e-8 in the fields and
# 7.3e-6 in the RK increments (FP32 rounding of the reconstruction, fluxes
# and divergences; measured over these cases).  The density flux's large
# background part, carrier x rhobar / pbar, stays FP64 (k_rho_background):
# rounded to FP32 with the rest it left errors of a few 1e-6 in RHO.  A
# wrong index, sign or term gives errors of 1e-3 and more.
FLUX32_FIELD_RTOL = 1e-7
FLUX32_INCR_RTOL = 3e-5
INCREMENTS = {"RHO": "DRHO", "RHOP": "DRHOP", "U": "DU", "V": "DV", "W": "DW", "CHI": "DCHI"}


@pytest.mark.parametrize("name", ["periodic_hill", "mountain_wave", "mountain_wave_3d", "cold_bubble", "vortex",
                                  "mountain_wave_tracer", "mountain_wave_relax", "periodic_hill_damp"])
def test_explicit_stage_fp32_fluxes(pcf, abi, oracle_mod, name):
    """explicit_integration! LHS with the RK stages' fluxes in FP32 against
    the FP64 oracle, fields and increments."""
    nl, gpu, orc = pair(pcf, abi, oracle_mod, name, flux_precision="mixed")
    names = predictands(nl)
    perturb(gpu, orc, names, seed=7, amp=0.03)
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
        s.compute_sponges(0.3, 2.0)
        s.backup_predictands()
        s.compute_theta_fluxes()
        s.explicit_integration_lhs(0.3, 2.0)
    for f in names:
        if f == "PIP":
            continue
        err = rel(gpu.get(f), orc.get(f))
        derr = rel(gpu.get(INCREMENTS[f]), orc.get(INCREMENTS[f]))
        print(name, f, f"field {err:.2e} increment {derr:.2e}")
        assert err <= FLUX32_FIELD_RTOL, (f, err)
        assert derr <= FLUX32_INCR_RTOL, (INCREMENTS[f], derr)


@pytest.mark.parametrize("name", ["mountain_wave_3d", "periodic_hill"])
@pytest.mark.parametrize("fluxes", ["fp64", "mixed"])
def test_stage_launch_shapes_give_the_same_bits(pcf, abi, oracle_mod, name, fluxes):
    """Every launch shape the autotuner of the fused stage kernels can pick
    (rows and levels per thread) gives the same fields, bit for bit, and the
    tuner records a choice for every variable it ran."""
    import ctypes
    results = []
    for shape in range(9):
        nl, gpu, orc = pair(pcf, abi, oracle_mod, name, flux_precision=fluxes)
        names = predictands(nl)
        perturb(gpu, orc, names, seed=7, amp=0.03)
        for slot in range(14):
            assert gpu.lib.pcf_set_stage_shape(gpu.h, slot, shape) == 0
        gpu.set_boundaries(abi.BOUNDARY_PREDICTANDS)
        gpu.compute_sponges(0.3, 2.0)
        gpu.backup_predictands()
        gpu.explicit_integration_lhs(0.3, 2.0)
        results.append({f: gpu.get(f) for f in names})
        gpu.close()
    for r in results[1:]:
        for f in r:
            assert np.array_equal(r[f], results[0][f]), f
    nl, gpu, orc = pair(pcf, abi, oracle_mod, name, flux_precision=fluxes)
    gpu.set_boundaries(abi.BOUNDARY_PREDICTANDS)
    gpu.compute_sponges(0.3, 2.0)
    gpu.backup_predictands()
    gpu.explicit_integration_lhs(0.3, 2.0)
    tuned = gpu.stage_tuning()
    print(name, fluxes, tuned)
    want = {"rhop", "u", "v", "w"} | ({"rho"} if "RHO" in predictands(nl) else set())
    assert {k.split("/")[0] for k in tuned} == want


@pytest.mark.parametrize("method", ["fp64", "mixed-line", "mixed-multigrid"])
def test_solver_graphs_give_the_eager_bits(pcf, method):
    """BiCGSTAB iterations replayed from the captured hipGraph (the default)
    against the same iterations launched kernel by kernel: the same
    iteration counts and the same fields, bit for bit, over steps whose
    solves replay the graph tens of times with host synchronisations in
    between."""
    runs = []
    for graphs in (True, False):
        nl = pcf.examples.mountain_wave_3d(x_size=32, y_size=32, z_size=16, iterations=4,
                                           precision="fp64" if method == "fp64" else "mixed")
        if method == "mixed-multigrid":
            nl.poisson.preconditioner_type = "multigrid"
            nl.poisson.initial_guess = "previous"
        st = pcf.State(nl)
        assert st.lib.pcf_set_solver_graphs(st.h, int(graphs)) == 0
        its = [st.initial_cleaning()]
        for _ in range(3):
            its.append(list(st.step().niter))
        runs.append((its, {f: st.get(f) for f in ("RHO", "RHOP", "U", "V", "W", "PIP")}))
        st.close()
    print(method, "iterations (graphs, eager):", runs[0][0], runs[1][0])
    assert runs[0][0] == runs[1][0]
    for f in runs[0][1]:
        assert np.array_equal(runs[0][1][f], runs[1][1][f]), f


@pytest.mark.parametrize("name", CASES)
def test_implicit_stage(pcf, abi, oracle_mod, name):
    nl, gpu, orc = pair(pcf, abi, oracle_mod, name, tight=True)
    names = predictands(nl)
    perturb(gpu, orc, names, seed=8, amp=0.03)
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
        s.compute_sponges(0.3, 2.0)
    n1 = gpu.implicit_integration(0.3, 2.0, 2.0)
    n2 = orc.implicit_integration(0.3, 2.0, 2.0)
    assert abs(n1 - n2) <= iter_slack(n2), (n1, n2)
    for f in names:
        err = rel(gpu.get(f), orc.get(f))
        assert err <= SOLVE_RTOL, (f, err)


@pytest.mark.parametrize("name", CASES)
def test_steps(pcf, abi, oracle_mod, name):
    """initial cleaning + three full time steps of integrate."""
    nl, gpu, orc = pair(pcf, abi, oracle_mod, name, tight=True)
    m1, m2 = gpu.initial_cleaning(), orc.initial_cleaning()
    assert abs(m1 - m2) <= iter_slack(m2), (m1, m2)
    for _ in range(3):
        a, b = gpu.step(), orc.step()
        assert a.dt == pytest.approx(b.dt, rel=1e-10)
        for q in range(2):
            assert abs(a.niter[q] - b.niter[q]) <= iter_slack(b.niter[q]), (a.niter[q], b.niter[q])
    names = predictands(nl)
    for f in names:
        err = rel(gpu.get(f), orc.get(f))
        assert err <= STEP_RTOL, (f, err)
