"""ctypes front end of the CPU FP64 oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product (pincflow.jl_amd/)
never imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liborc.so"

FIELDS = None  # filled from the product's ABI mirror on first use
VECTORS = None


class StepInfo(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("niter", ctypes.c_int32 * 2), ("output", ctypes.c_int32), ("pad", ctypes.c_int32)]


def build(quiet: bool = True) -> Path:
    """Compile liborc.so with oracle/Makefile."""
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=quiet, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        P, D, I = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        PD = ctypes.POINTER(ctypes.c_double)
        PI32 = ctypes.POINTER(ctypes.c_int32)
        sig = {
            "orc_create": (P, [P, ctypes.POINTER(ctypes.c_int)]),
            "orc_destroy": (None, [P]),
            "orc_last_error": (ctypes.c_char_p, [P]),
            "orc_set_threads": (None, [I]),
            "orc_set_reduction_order": (None, [I, I]),
            "orc_dims": (None, [P, PI32]),
            "orc_initial_cleaning": (I, [P, PI32]),
            "orc_step": (I, [P, ctypes.POINTER(StepInfo)]),
            "orc_time": (D, [P]),
            "orc_set_time": (I, [P, D]),
            "orc_get_field": (I, [P, I, PD]),
            "orc_set_field": (I, [P, I, PD]),
            "orc_get_vector": (I, [P, I, PD]),
            "orc_set_vector": (I, [P, I, PD]),
            "orc_compute_lhs": (I, [P, PD]),
            "orc_compute_operator": (None, [P, D, D]),
            "orc_apply_operator": (None, [P, PD, PD, I]),
            "orc_apply_preconditioner": (None, [P, PD, PD]),
            "orc_apply_bicgstab": (I, [P, D, PI32]),
            "orc_last_residual": (D, [P]),
            "orc_apply_corrector": (I, [P, D, D, PI32]),
            "orc_correct": (None, [P, D, D]),
            "orc_solution_to_dpip": (None, [P, D]),
            "orc_muscl_line": (None, [PD, I, PD, PD]),
            "orc_compute_sponges": (None, [P, D, D]),
            "orc_set_boundaries": (None, [P, I]),
            "orc_backup_predictands": (None, [P]),
            "orc_compute_fluxes": (None, [P]),
            "orc_save_backups": (None, [P]),
            "orc_compute_theta_fluxes": (None, [P]),
            "orc_modify_compressible_wind": (None, [P, I]),
            "orc_synchronize_compressible_atmosphere": (None, [P]),
            "orc_update_lhs": (None, [P, D, I, I]),
            "orc_update_rhs": (None, [P, D, I, I, D]),
            "orc_apply_lhs_sponge": (None, [P, D, D, I]),
            "orc_explicit_integration_lhs": (I, [P, D, D]),
            "orc_explicit_integration_rhs": (I, [P, D, D]),
            "orc_implicit_integration": (I, [P, D, D, D, PI32]),
            "orc_compute_time_step": (I, [P, PD]),
            "orc_vertical_wind": (D, [P, I, I, I]),
            "orc_stress_tensor": (D, [P, I, I, I, I, I]),
            "orc_constants": (None, [P, PD]),
            "orc_grid_1d": (None, [P, PD, PD]),
            "orc_config_size": (I, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleError(RuntimeError):
    def __init__(self, rc: int, msg: str):
        super().__init__(f"oracle returned {rc}: {msg}")
        self.rc = rc


class Oracle:
    """One oracle State (State.jl:106-131) built from a pcf_config."""

    def __init__(self, cfg, abi, threads: int = 0):
        L = lib()
        if L.orc_config_size() != ctypes.sizeof(cfg):
            raise RuntimeError("pcf_config layout mismatch between the oracle and the ABI mirror")
        if threads:
            L.orc_set_threads(threads)
        rc = ctypes.c_int(0)
        self.abi = abi
        self.h = L.orc_create(ctypes.byref(cfg), ctypes.byref(rc))
        if rc.value:
            msg = L.orc_last_error(self.h).decode()
            L.orc_destroy(self.h)
            self.h = None
            raise OracleError(rc.value, msg)
        d = (ctypes.c_int32 * 6)()
        L.orc_dims(self.h, d)
        self.nx, self.ny, self.nz, self.nxx, self.nyy, self.nzz = list(d)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc:
            raise OracleError(rc, lib().orc_last_error(self.h).decode())

    # ---- fields ----
    def get(self, name: str) -> np.ndarray:
        a = np.empty((self.nxx, self.nyy, self.nzz), order="F")
        self._check(lib().orc_get_field(self.h, self.abi.FIELD[name], _ptr(a)))
        return a

    def set(self, name: str, a: np.ndarray):
        a = np.asfortranarray(a, dtype=np.float64)
        self._check(lib().orc_set_field(self.h, self.abi.FIELD[name], _ptr(a)))

    def get_vector(self, name: str) -> np.ndarray:
        a = np.empty((self.nx, self.ny, self.nz), order="F")
        self._check(lib().orc_get_vector(self.h, self.abi.VECTOR[name], _ptr(a)))
        return a

    def set_vector(self, name: str, a: np.ndarray):
        a = np.asfortranarray(a, dtype=np.float64)
        self._check(lib().orc_set_vector(self.h, self.abi.VECTOR[name], _ptr(a)))

    # ---- reference functions ----
    def initial_cleaning(self) -> int:
        n = ctypes.c_int32(0)
        self._check(lib().orc_initial_cleaning(self.h, ctypes.byref(n)))
        return n.value

    def step(self) -> StepInfo:
        info = StepInfo()
        self._check(lib().orc_step(self.h, ctypes.byref(info)))
        return info

    def time(self) -> float:
        return lib().orc_time(self.h)

    def restart(self, nl, output, time_s: float, fields: dict):
        """Continue from a stored record: read_input! (read_input!.jl:14-84)
        followed by integrate.jl:168-186.  fields are write_output's datasets
        rhop, us, vs, wts, pip (p in the compressible model, chi with a
        tracer), dimensional; the interiors of the predictands are replaced,
        rho follows rhop outside the Boussinesq model, P is p / rhoref /
        thetaref, chi is the dataset times rhobar + rho; then the predictand
        boundaries, pbar := P, and the model time."""
        c = output.constants(nl)
        d = nl.domain
        ii = slice(d.nbx, d.nbx + d.x_size)
        jj = slice(d.nby, d.nby + d.y_size)
        kk = slice(d.nbz, d.nbz + d.z_size)
        for name, key, scale in (("RHOP", "rhop", c["rhoref"]), ("U", "us", c["uref"]), ("V", "vs", c["uref"]),
                                 ("W", "wts", c["uref"]), ("PIP", "pip", 1.0)):
            a = self.get(name)
            a[ii, jj, kk] = np.asarray(fields[key], dtype=np.float64) / scale
            self.set(name, a)
            if name == "RHOP" and nl.atmosphere.model != "Boussinesq":
                r = self.get("RHO")
                r[ii, jj, kk] = a[ii, jj, kk]
                self.set("RHO", r)
        if nl.atmosphere.model == "Compressible":
            pp = self.get("P")
            pp[ii, jj, kk] = np.asarray(fields["p"], dtype=np.float64) / c["rhoref"] / c["thetaref"]
            self.set("P", pp)
        if nl.tracer.tracer_setup == "TracerOn":
            chi, rho, rhobar = self.get("CHI"), self.get("RHO"), self.get("RHOBAR")
            chi[ii, jj, kk] = np.asarray(fields["chi"], dtype=np.float64) * (rhobar[ii, jj, kk] + rho[ii, jj, kk])
            self.set("CHI", chi)
        self.set_boundaries(self.abi.BOUNDARY_PREDICTANDS)
        self.synchronize_compressible_atmosphere()
        self._check(lib().orc_set_time(self.h, time_s))

    def compute_lhs(self) -> float:
        t = ctypes.c_double(0.0)
        self._check(lib().orc_compute_lhs(self.h, ctypes.byref(t)))
        return t.value

    def compute_operator(self, dt: float, rf: float):
        lib().orc_compute_operator(self.h, dt, rf)

    def apply_operator(self, s: np.ndarray, total: bool = True) -> np.ndarray:
        s = np.asfortranarray(s, dtype=np.float64)
        out = np.empty_like(s, order="F")
        lib().orc_apply_operator(self.h, _ptr(s), _ptr(out), 0 if total else 1)
        return out

    def apply_preconditioner(self, s: np.ndarray) -> np.ndarray:
        s = np.asfortranarray(s, dtype=np.float64)
        out = np.empty_like(s, order="F")
        lib().orc_apply_preconditioner(self.h, _ptr(s), _ptr(out))
        return out

    def apply_bicgstab(self, tolref: float):
        n = ctypes.c_int32(0)
        rc = lib().orc_apply_bicgstab(self.h, tolref, ctypes.byref(n))
        return rc, n.value

    def last_residual(self) -> float:
        return lib().orc_last_residual(self.h)

    def apply_corrector(self, dt: float, rf: float):
        n = ctypes.c_int32(0)
        rc = lib().orc_apply_corrector(self.h, dt, rf, ctypes.byref(n))
        return rc, n.value

    def correct(self, dt: float, rf: float):
        lib().orc_correct(self.h, dt, rf)

    def solution_to_dpip(self, dt: float):
        """dpip = solution / (sqrt(pbar^2/rhobar) dt) plus its boundaries (the post-solve part of
        apply_corrector!); the SOLUTION vector is divided in place, as in the reference."""
        lib().orc_solution_to_dpip(self.h, dt)

    def compute_sponges(self, dt: float, time: float):
        lib().orc_compute_sponges(self.h, dt, time)

    def set_boundaries(self, which: int):
        lib().orc_set_boundaries(self.h, which)

    def backup_predictands(self):
        lib().orc_backup_predictands(self.h)

    def compute_fluxes(self):
        lib().orc_compute_fluxes(self.h)

    def save_backups(self):
        lib().orc_save_backups(self.h)

    def compute_theta_fluxes(self):
        """compute_fluxes!(state, p0, Theta()) with the p0 copy."""
        lib().orc_compute_theta_fluxes(self.h)

    def modify_compressible_wind(self, multiply: bool):
        lib().orc_modify_compressible_wind(self.h, int(multiply))

    def synchronize_compressible_atmosphere(self):
        lib().orc_synchronize_compressible_atmosphere(self.h)

    def update_lhs(self, dt: float, m: int, var: int):
        lib().orc_update_lhs(self.h, dt, m, var)

    def update_rhs(self, dt: float, var: int, implicit: bool, rf: float = 1.0):
        lib().orc_update_rhs(self.h, dt, var, int(implicit), rf)

    def apply_lhs_sponge(self, dt: float, time: float, var: int):
        lib().orc_apply_lhs_sponge(self.h, dt, time, var)

    def explicit_integration_lhs(self, dt: float, time: float):
        self._check(lib().orc_explicit_integration_lhs(self.h, dt, time))

    def explicit_integration_rhs(self, dt: float, time: float):
        self._check(lib().orc_explicit_integration_rhs(self.h, dt, time))

    def implicit_integration(self, dt: float, time: float, rf: float) -> int:
        n = ctypes.c_int32(0)
        self._check(lib().orc_implicit_integration(self.h, dt, time, rf, ctypes.byref(n)))
        return n.value

    def compute_time_step(self) -> float:
        dt = ctypes.c_double(0.0)
        self._check(lib().orc_compute_time_step(self.h, ctypes.byref(dt)))
        return dt.value

    def vertical_wind(self, i, j, k) -> float:
        return lib().orc_vertical_wind(self.h, i, j, k)

    def stress_tensor(self, i, j, k, mu, nu) -> float:
        return lib().orc_stress_tensor(self.h, i, j, k, mu, nu)

    def constants(self):
        a = np.empty(13)
        lib().orc_constants(self.h, _ptr(a))
        keys = ["lref", "tref", "uref", "rhoref", "thetaref", "pref", "re", "g_ndim", "kappa", "dx", "dy", "dz",
                "dzcmin"]
        return dict(zip(keys, a.tolist()))


def integrate(nl, abi, output, records: bool = True, threads: int = 0, max_steps: int = 1 << 30):
    """integrate(namelists) (integrate.jl:68-415) on the oracle; returns
    (oracle, list of output records, list of per-step infos)."""
    cfg = abi_config(nl, abi)
    o = Oracle(cfg, abi, threads)
    recs, steps = [], []
    if nl.poisson.initial_cleaning:
        o.initial_cleaning()
    if records:
        recs.append(output.record_datasets(nl, o.get, 0.0, True))
    n = nl.output.iterations if nl.output.output_steps else max_steps
    for itime in range(1, n + 1):
        info = o.step()
        steps.append((info.time, info.dt, info.niter[0], info.niter[1]))
        if records and ((nl.output.output_steps and itime % nl.output.nout == 0) or
                        (not nl.output.output_steps and info.output)):
            recs.append(output.record_datasets(nl, o.get, info.time, False))
        if not nl.output.output_steps and info.time >= nl.output.tmax:
            break
    return o, recs, steps


def abi_config(nl, abi):
    from importlib import import_module
    return import_module(abi.__name__.rsplit(".", 1)[0] + ".namelists").to_config(nl)


def set_reduction_order(block: int = 0, backwards: bool = False):
    """Summation order of the oracle's global sums: block = 0 is Julia's
    pairwise order (the default), block > 0 sums blocks of that many terms
    in order and adds the block sums forwards or backwards."""
    lib().orc_set_reduction_order(block, int(backwards))


def muscl_line(phi: np.ndarray):
    """apply_1d_muscl! on one line: (left-face, right-face) limited states of every cell;
    the two end cells keep the reference's placeholder 1000.0."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    left = np.empty_like(phi)
    right = np.empty_like(phi)
    lib().orc_muscl_line(_ptr(phi), len(phi), _ptr(left), _ptr(right))
    return left, right

