"""The reference's error paths on the device, against the oracle: the same
return codes and iteration counts, each reached through a bounded loop.

* BiCGSTAB that does not converge within poisson_iterations
  (apply_bicgstab!.jl:146-149, errflag = true): PCF_E_NOCONV with niter =
  poisson_iterations, from a too small iteration limit and from a right-hand
  side with a NaN in it (every comparison with NaN is false, so the solve
  never meets its criterion and stops at the limit, as the reference does);
* dt < dtmin (compute_time_step.jl:142-150): PCF_E_DTMIN;
* divl2_norm = 0 while divl2 != 0 (compute_lhs!.jl:149-157): PCF_E_DIVNORM,
  reached through underflow: a uniform divergence whose three components are
  each about 1e-162, so that each square rounds to zero while the square of
  their sum does not; and both zero (a flow at rest), where tolref = 1 and no
  iteration is needed.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def corrector(st, dt=1.0, rf=1.0):
    """pcf_apply_corrector through the C ABI: (return code, niter)."""
    n = ctypes.c_int32(-1)
    rc = st.lib.pcf_apply_corrector(st.h, ctypes.c_double(dt), ctypes.c_double(rf), ctypes.byref(n))
    return rc, n.value


def divergent_winds(gpu, orc, seed=3):
    rng = np.random.default_rng(seed)
    for f in ("U", "V", "W"):
        a = orc.get(f)
        a = a + 0.1 * max(float(np.max(np.abs(a))), 1.0) * rng.standard_normal(a.shape)
        gpu.set(f, a)
        orc.set(f, a)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_iteration_limit_is_noconv(pcf, abi, oracle_mod, precision):
    nl = pcf.examples.mountain_wave(x_size=12, y_size=10, z_size=9)
    nl.poisson.poisson_iterations = 1
    nl64 = pcf.examples.mountain_wave(x_size=12, y_size=10, z_size=9)
    nl64.poisson.poisson_iterations = 1
    nl.poisson.precision = precision
    if precision == "mixed":
        nl.poisson.preconditioner_type = "multigrid"
    gpu = pcf.State(nl)
    orc = oracle_mod.Oracle(pcf.to_config(nl64), abi)
    divergent_winds(gpu, orc)
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
    rc, n = corrector(gpu)
    orc_rc, orc_n = orc.apply_corrector(1.0, 1.0)
    assert (orc_rc, orc_n) == (abi.E_NOCONV, 1)
    assert (rc, n) == (abi.E_NOCONV, 1)
    # the state's own handle raises the reference's errflag as an exception
    with pytest.raises(pcf.PoissonNotConverged):
        gpu.apply_corrector(1.0, 1.0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nan_right_hand_side_stops_at_the_limit(pcf, abi, oracle_mod, precision):
    limit = 12
    nls = []
    for _ in range(2):
        nl = pcf.examples.mountain_wave(x_size=12, y_size=10, z_size=9)
        nl.poisson.poisson_iterations = limit
        nls.append(nl)
    nls[0].poisson.precision = precision
    if precision == "mixed":
        nls[0].poisson.preconditioner_type = "multigrid"
    gpu = pcf.State(nls[0])
    orc = oracle_mod.Oracle(pcf.to_config(nls[1]), abi)
    divergent_winds(gpu, orc)
    u = orc.get("U")
    u[7, 6, 5] = np.nan  # one interior face
    for s in (gpu, orc):
        s.set("U", u)
    assert orc.apply_corrector(1.0, 1.0) == (abi.E_NOCONV, limit)
    assert corrector(gpu) == (abi.E_NOCONV, limit)


def test_dt_below_dtmin(pcf, abi, oracle_mod):
    nl = pcf.examples.cold_bubble(x_size=12, z_size=10)
    nl.discretization.dtmin = 100.0  # above dtmax = 60 s: every step is too short
    gpu = pcf.State(nl)
    orc = oracle_mod.Oracle(pcf.to_config(nl), abi)
    info = pcf._abi.PcfStepInfo()
    assert gpu.lib.pcf_step(gpu.h, ctypes.byref(info)) == abi.E_DTMIN
    assert "dtmin" in gpu.lib.pcf_last_error().decode()
    with pytest.raises(oracle_mod.OracleError) as e:
        orc.step()
    assert e.value.rc == abi.E_DTMIN
    with pytest.raises(pcf.PinCFlowError):
        gpu.step()


def underflow_namelists(pcf):
    return pcf.Namelists(
        atmosphere=pcf.AtmosphereNamelist(model="Boussinesq", background="NeutralStratification"),
        domain=pcf.DomainNamelist(x_size=8, y_size=6, z_size=5, lx=8000.0, ly=6000.0, lz=5000.0))


def test_zero_divergence_norm(pcf, abi, oracle_mod):
    nl = underflow_namelists(pcf)
    gpu = pcf.State(nl)
    orc = oracle_mod.Oracle(pcf.to_config(nl), abi)
    # at rest: divl2 = divl2_norm = 0, tolref = 1, and the solve needs no iteration
    for s in (gpu, orc):
        for f in ("U", "V", "W"):
            s.set(f, np.zeros((gpu.nxx, gpu.nyy, gpu.nzz)))
    assert gpu.compute_lhs() == 1.0 == orc.compute_lhs()
    assert corrector(gpu) == (0, 0)
    assert orc.apply_corrector(1.0, 1.0) == (0, 0)
    # winds growing linearly in every direction (halo values included, no
    # boundaries set) give one divergence in every cell; scale them to
    # components of 1e-162 each
    nxx, nyy, nzz = gpu.nxx, gpu.nyy, gpu.nzz
    idx = np.meshgrid(np.arange(nxx), np.arange(nyy), np.arange(nzz), indexing="ij")
    coef = []
    for name, grow in zip(("U", "V", "W"), idx):
        for f in ("U", "V", "W"):
            orc.set(f, np.where(f == name, 1.0, 0.0) * grow)
        orc.compute_lhs()
        coef.append(float(np.mean(orc.get_vector("RHS"))))
    for name, grow, c in zip(("U", "V", "W"), idx, coef):
        a = grow * (1.0e-162 / c)
        gpu.set(name, a)
        orc.set(name, a)
    with pytest.raises(oracle_mod.OracleError) as e:
        orc.compute_lhs()
    assert e.value.rc == abi.E_DIVNORM
    t = ctypes.c_double(0.0)
    assert gpu.lib.pcf_compute_lhs(gpu.h, ctypes.byref(t)) == abi.E_DIVNORM
    assert "divl2_norm = 0" in gpu.lib.pcf_last_error().decode()
    assert corrector(gpu)[0] == abi.E_DIVNORM
    # the right-hand side itself is still computed, equal to the oracle's
    np.testing.assert_allclose(gpu.get_vector("RHS"), orc.get_vector("RHS"), rtol=1e-13, atol=0.0)
