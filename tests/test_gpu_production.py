"""The device against the oracle at the production size: the pseudo-
incompressible mountain wave at 256 x 256 x 128 (isothermal 250 K, 200 km x
200 km x 30 km with sinh stretching, one Gaussian hill), the workload the
throughput is measured on.  Every kernel without a reduction is compared at
KERNEL_RTOL of the field's max norm, as in test_gpu_parity.py:

* the grid and background the state is built with (bit for bit);
* compute_operator! (compute_operator!.jl:40-1606), all 25 coefficients, on
  a perturbed density with the LHS sponge active;
* apply_operator! Total and Horizontal (apply_operator!.jl:48-479) and
  apply_preconditioner! (apply_preconditioner!.jl:40-124) on a random vector,
  with the line solves' levels below and above the 80 the FP64 line solve
  keeps in LDS checked apart;
* compute_lhs! (compute_lhs!.jl:68-160) on perturbed winds;
* one explicit_integration!(LHS) (explicit_integration!.jl:42-101: three RK
  stages of MUSCL fluxes with viscous stress, updates and sponges).
"""
import hipmem
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KERNEL_RTOL = 1e-13


def rel(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


@pytest.fixture(scope="module")
def pair(pcf, abi, oracle_mod):
    nl = pcf.examples.mountain_wave_3d(x_size=256, y_size=256, z_size=128)
    print("building the 256 x 256 x 128 mountain wave on the device and the oracle", flush=True)
    gpu = pcf.State(nl)
    orc = oracle_mod.Oracle(pcf.to_config(nl), abi)
    yield nl, gpu, orc
    gpu.close()


def perturb(gpu, orc, names, seed, amp):
    rng = np.random.default_rng(seed)
    for n in names:
        a = orc.get(n)
        scale = max(float(np.max(np.abs(a))), 1e-3)
        a = a + amp * scale * rng.standard_normal(a.shape)
        gpu.set(n, a)
        orc.set(n, a)


def test_setup_fields(pair):
    _, gpu, orc = pair
    for f in ("PBAR", "THETABAR", "RHOBAR", "N2", "JAC", "MET13", "MET23", "MET33", "ZC", "RHO", "U", "W"):
        assert np.array_equal(gpu.get(f), orc.get(f)), f


def test_compute_operator(pair, abi):
    _, gpu, orc = pair
    perturb(gpu, orc, ["RHO"], seed=21, amp=0.01)
    for s in (gpu, orc):
        s.compute_sponges(0.02, 1.0)
        s.compute_operator(0.02, 2.0)
    for v in abi.VECTORS[2:]:
        err = rel(gpu.get_vector(v), orc.get_vector(v))
        assert err <= KERNEL_RTOL, (v, err)


def test_apply_operator_and_preconditioner(pair):
    _, gpu, orc = pair
    for s in (gpu, orc):
        s.compute_sponges(0.02, 1.0)
        s.compute_operator(0.02, 1.0)
    x = np.random.default_rng(22).standard_normal((gpu.nx, gpu.ny, gpu.nz))
    dx = hipmem.DeviceArray(x.shape).upload(x)
    dy = hipmem.DeviceArray(x.shape)
    for total in (True, False):
        gpu.apply_operator_device(dx.ptr, dy.ptr, total)
        err = rel(dy.download(), orc.apply_operator(x, total))
        assert err <= KERNEL_RTOL, (total, err)
    gpu.apply_preconditioner_device(dx.ptr, dy.ptr)
    got, want = dy.download(), orc.apply_preconditioner(x)
    for half in (slice(0, 80), slice(80, None)):
        err = rel(got[:, :, half], want[:, :, half])
        assert err <= KERNEL_RTOL, (half, err)
    assert gpu.dot_device(dx.ptr, dx.ptr) == pytest.approx(float(np.sum(x * x)), rel=1e-13)


def test_compute_lhs(pair, abi):
    _, gpu, orc = pair
    perturb(gpu, orc, ["U", "V", "W"], seed=23, amp=0.05)
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
    t1, t2 = gpu.compute_lhs(), orc.compute_lhs()
    assert rel(gpu.get_vector("RHS"), orc.get_vector("RHS")) <= KERNEL_RTOL
    assert t1 == pytest.approx(t2, rel=1e-12)


def test_explicit_lhs_stage(pair, abi):
    _, gpu, orc = pair
    names = ["RHO", "RHOP", "U", "V", "W"]
    perturb(gpu, orc, names, seed=24, amp=0.02)
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
        s.compute_sponges(0.02, 1.0)
        s.backup_predictands()
    perturb(gpu, orc, ["U", "V", "W"], seed=25, amp=0.01)  # transporting winds differ from the current ones
    for s in (gpu, orc):
        s.set_boundaries(abi.BOUNDARY_PREDICTANDS)
        s.explicit_integration_lhs(0.02, 1.0)
    for f in names + ["PIP"]:
        err = rel(gpu.get(f), orc.get(f))
        assert err <= KERNEL_RTOL, (f, err)
