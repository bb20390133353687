"""Host handle of the device-resident model state and ``integrate``.

``State(namelists)`` mirrors ``State(namelists)`` (src/Types/State.jl:106-131)
and owns one ``pcf_state`` of libpincflow_hip.so; its methods are the module
functions of the reference on that state.  ``integrate(namelists)`` mirrors
src/Integration/integrate.jl:68-415 with the HDF5 output replaced by in-memory
records of the output datasets.

Errors follow the reference: a BiCGSTAB that does not converge raises
``PoissonNotConverged`` (errflag = true), the ``error(...)`` cases raise
``PinCFlowError`` with the library's message.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Tuple

import numpy as np

from . import _abi
from .namelists import Namelists, to_config


class PinCFlowError(RuntimeError):
    def __init__(self, rc: int, msg: str):
        super().__init__(f"pincflow error {rc}: {msg}")
        self.rc = rc


class PoissonNotConverged(PinCFlowError):
    pass


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class State:
    """Device-resident model state (one pcf_state)."""

    def __init__(self, namelists: Namelists):
        self.namelists = namelists
        self.lib = _abi.load_library()
        cfg = to_config(namelists)
        h = ctypes.c_void_p()
        rc = self.lib.pcf_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc:
            raise PinCFlowError(rc, self.lib.pcf_last_error().decode())
        self.h = h
        d = (ctypes.c_int32 * 6)()
        self.lib.pcf_dims(self.h, d)
        self.nx, self.ny, self.nz, self.nxx, self.nyy, self.nzz = list(d)

    def close(self):
        if getattr(self, "h", None):
            self.lib.pcf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc: int, niter: Optional[int] = None):
        if rc == _abi.E_NOCONV:
            raise PoissonNotConverged(rc, f"BiCGSTAB did not converge in {niter} iterations")
        if rc:
            raise PinCFlowError(rc, self.lib.pcf_last_error().decode())

    # ---- host copies ------------------------------------------------------
    def get(self, name: str) -> np.ndarray:
        a = np.empty((self.nxx, self.nyy, self.nzz), order="F")
        self._check(self.lib.pcf_get_field(self.h, _abi.FIELD[name], _ptr(a)))
        return a

    def set(self, name: str, a: np.ndarray):
        a = np.asfortranarray(a, dtype=np.float64)
        self._check(self.lib.pcf_set_field(self.h, _abi.FIELD[name], _ptr(a)))

    def get_vector(self, name: str) -> np.ndarray:
        a = np.empty((self.nx, self.ny, self.nz), order="F")
        self._check(self.lib.pcf_get_vector(self.h, _abi.VECTOR[name], _ptr(a)))
        return a

    def set_vector(self, name: str, a: np.ndarray):
        a = np.asfortranarray(a, dtype=np.float64)
        self._check(self.lib.pcf_set_vector(self.h, _abi.VECTOR[name], _ptr(a)))

    # ---- Integration --------------------------------------------------------
    def initial_cleaning(self) -> int:
        n = ctypes.c_int32(0)
        rc = self.lib.pcf_initial_cleaning(self.h, ctypes.byref(n))
        self._check(rc, n.value)
        return n.value

    def step(self) -> _abi.PcfStepInfo:
        info = _abi.PcfStepInfo()
        rc = self.lib.pcf_step(self.h, ctypes.byref(info))
        self._check(rc, max(info.niter))
        return info

    def solver_graphs(self, on: bool = True):
        """Replay BiCGSTAB iterations from a captured hipGraph (same kernels,
        same bits).  Only for a state stepped by the one thread that uses
        HIP: while a capture runs, HIP refuses other threads' stream
        synchronisations and allocations."""
        self._check(self.lib.pcf_set_solver_graphs(self.h, int(on)))

    def compute_time_step(self) -> float:
        dt = ctypes.c_double(0.0)
        self._check(self.lib.pcf_compute_time_step(self.h, ctypes.byref(dt)))
        return dt.value

    def compute_sponges(self, dt: float, time: float):
        self._check(self.lib.pcf_compute_sponges(self.h, dt, time))

    def backup_predictands(self):
        self._check(self.lib.pcf_backup_predictands(self.h))

    def explicit_integration_lhs(self, dt: float, time: float):
        self._check(self.lib.pcf_explicit_integration_lhs(self.h, dt, time))

    def explicit_integration_rhs(self, dt: float, time: float):
        self._check(self.lib.pcf_explicit_integration_rhs(self.h, dt, time))

    def implicit_integration(self, dt: float, time: float, rayleigh_factor: float) -> int:
        n = ctypes.c_int32(0)
        rc = self.lib.pcf_implicit_integration(self.h, dt, time, rayleigh_factor, ctypes.byref(n))
        self._check(rc, n.value)
        return n.value

    # ---- Boundaries / FluxCalculator / Update -------------------------------
    def set_boundaries(self, which: int):
        self._check(self.lib.pcf_set_boundaries(self.h, which))

    def reconstruct(self):
        """reconstruct!(state): the reconstruction inputs of every variable
        (the limited face values are evaluated inside compute_fluxes)."""
        self._check(self.lib.pcf_reconstruct(self.h))

    def compute_fluxes(self):
        """compute_fluxes!(state, p0) and set_boundaries!(state, BoundaryFluxes())."""
        self._check(self.lib.pcf_compute_fluxes(self.h))

    def compute_theta_fluxes(self):
        """compute_fluxes!(state, p0, Theta()): heat conduction (compressible model)."""
        self._check(self.lib.pcf_compute_theta_fluxes(self.h))

    def modify_compressible_wind(self, multiply: bool):
        """modify_compressible_wind!(state, * or /) (compressible model)."""
        self._check(self.lib.pcf_modify_compressible_wind(self.h, int(multiply)))

    def synchronize_compressible_atmosphere(self):
        """synchronize_compressible_atmosphere!(state, predictands): pbar := P."""
        self._check(self.lib.pcf_synchronize_compressible_atmosphere(self.h))

    def update_lhs(self, dt: float, m: int, var: int):
        self._check(self.lib.pcf_update_lhs(self.h, dt, m, var))

    def update_rhs(self, dt: float, var: int, implicit: bool, rayleigh_factor: float = 1.0):
        self._check(self.lib.pcf_update_rhs(self.h, dt, var, int(implicit), rayleigh_factor))

    def apply_lhs_sponge(self, dt: float, time: float, var: int):
        self._check(self.lib.pcf_apply_lhs_sponge(self.h, dt, time, var))

    # ---- PoissonSolver -------------------------------------------------------
    def apply_corrector(self, dt: float, rayleigh_factor: float) -> int:
        n = ctypes.c_int32(0)
        rc = self.lib.pcf_apply_corrector(self.h, dt, rayleigh_factor, ctypes.byref(n))
        self._check(rc, n.value)
        return n.value

    def compute_lhs(self) -> float:
        t = ctypes.c_double(0.0)
        self._check(self.lib.pcf_compute_lhs(self.h, ctypes.byref(t)))
        return t.value

    def compute_operator(self, dt: float, rayleigh_factor: float):
        self._check(self.lib.pcf_compute_operator(self.h, dt, rayleigh_factor))

    def apply_bicgstab(self, tolref: float) -> int:
        n = ctypes.c_int32(0)
        rc = self.lib.pcf_apply_bicgstab(self.h, tolref, ctypes.byref(n))
        self._check(rc, n.value)
        return n.value

    def last_residual(self) -> float:
        r = ctypes.c_double(0.0)
        self._check(self.lib.pcf_last_residual(self.h, ctypes.byref(r)))
        return r.value

    def apply_operator_device(self, d_in: int, d_out: int, total: bool = True):
        """apply_operator! on device buffers (interior-sized FP64, given as addresses)."""
        self._check(self.lib.pcf_apply_operator(self.h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out),
                                                _abi.OP_TOTAL if total else _abi.OP_HORIZONTAL))

    def apply_preconditioner_device(self, d_in: int, d_out: int):
        self._check(self.lib.pcf_apply_preconditioner(self.h, ctypes.c_void_p(d_in), ctypes.c_void_p(d_out)))

    # ---- kernel timing -------------------------------------------------------
    def profile(self, on: bool, groups=None):
        """Record HIP events around the kernel groups of the library (clears
        totals); groups: names from _abi.KERNEL_GROUPS to record (all if None)."""
        if groups is None:
            self._check(self.lib.pcf_profile_enable(self.h, int(on)))
            return
        mask = 0
        for name in groups:
            mask |= 1 << _abi.KERNEL_GROUPS.index(name)
        self._check(self.lib.pcf_profile_groups(self.h, mask if on else 0))

    def profile_report(self) -> dict:
        """{group: (launches, total device ms)} since profile(True)."""
        out = {}
        for gid, name in enumerate(_abi.KERNEL_GROUPS):
            n = ctypes.c_int64(0)
            ms = ctypes.c_double(0.0)
            self._check(self.lib.pcf_profile_query(self.h, gid, ctypes.byref(n), ctypes.byref(ms)))
            if n.value:
                out[name] = (int(n.value), float(ms.value))
        return out

    def guess_fallbacks(self) -> int:
        """Solves whose initial guess (previous or learned) left a residual no
        smaller than |b| and was replaced by the zero start."""
        n = ctypes.c_int32(0)
        self._check(self.lib.pcf_guess_fallbacks(self.h, ctypes.byref(n)))
        return n.value

    def stage_tuning(self) -> dict:
        """The launch shapes the autotuner chose for the fused RK-stage
        kernels: {"u/fp64": "2 rows x 8 levels (0.51 ms)", ...}."""
        shape = (ctypes.c_int32 * 14)()
        ms = (ctypes.c_float * (14 * 9))()
        self._check(self.lib.pcf_stage_tuning(self.h, shape, ms))
        names = ["rho", "rhop", "u", "v", "w", "chi", "p"]
        out = {}
        for q in range(14):
            s = shape[q]
            if s >= 0:
                out[f"{names[q // 2]}/{'fp32' if q % 2 else 'fp64'}"] = (
                    f"{(1, 2, 4)[s // 3]} rows x {(4, 8, 16)[s % 3]} levels ({ms[9 * q + s]:.3f} ms)")
        return out

    def stage_fallbacks(self) -> int:
        """Solves rerun without the learned preconditioning stage after they
        failed to converge with it."""
        n = ctypes.c_int32(0)
        self._check(self.lib.pcf_stage_fallbacks(self.h, ctypes.byref(n)))
        return n.value

    # ---- time and restart ------------------------------------------------------
    def time(self):
        """(model time in s, steps taken)."""
        t = ctypes.c_double(0.0)
        n = ctypes.c_int32(0)
        self._check(self.lib.pcf_get_time(self.h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def restart(self, time_s: float, fields: dict):
        """Continue from a stored record (integrate.jl:168-186, read_input!.jl):
        fields are the dimensional datasets rhop, us, vs, wts and pip of
        write_output (p in the compressible model, chi with a tracer); the
        interior of the predictands is replaced, rho follows rhop outside the
        Boussinesq model, chi is the dataset times rhobar + rho
        (read_input!.jl:79-84), then the boundaries are set and pbar follows P."""
        from . import output

        nl = self.namelists
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
        if nl.atmosphere.model == "Compressible":  # read_input!.jl:73-77
            if "p" not in fields:
                raise ValueError("the restart record has no p dataset, but the model is compressible")
            pp = self.get("P")
            pp[ii, jj, kk] = np.asarray(fields["p"], dtype=np.float64) / c["rhoref"] / c["thetaref"]
            self.set("P", pp)
        if nl.tracer.tracer_setup == "TracerOn":
            if "chi" not in fields:
                raise ValueError("the restart record has no chi dataset, but the tracer is on")
            chi, rho, rhobar = self.get("CHI"), self.get("RHO"), self.get("RHOBAR")
            chi[ii, jj, kk] = np.asarray(fields["chi"], dtype=np.float64) * (rhobar[ii, jj, kk] + rho[ii, jj, kk])
            self.set("CHI", chi)
        self.set_boundaries(_abi.BOUNDARY_PREDICTANDS)
        # integrate.jl:182-185: pbar follows the restored P
        self._check(self.lib.pcf_synchronize_compressible_atmosphere(self.h))
        self._check(self.lib.pcf_set_time(self.h, time_s))

    def dot_device(self, d_a: int, d_b: int) -> float:
        r = ctypes.c_double(0.0)
        self._check(self.lib.pcf_dot(self.h, ctypes.c_void_p(d_a), ctypes.c_void_p(d_b), ctypes.byref(r)))
        return r.value


def integrate(namelists: Namelists, records: bool = True, max_steps: int = 1 << 30, write: bool = False,
              batch=None, on_step=None) -> Tuple[State, List[dict], List[tuple]]:
    """integrate(namelists) (integrate.jl:68-415) on the device.

    Returns the state, the output records (datasets of write_output at the
    start time and at every output time) and per-step (time, dt, niter1,
    niter2).  With output.restart the run continues from record iin of
    output.input_file; with write=True the records are stored in
    output.output_file (output.write_file).  ``batch`` (an
    ``ensemble.SolveBatch``): the state's solves join the batch's, from its
    creation to the end (or error) of the run.  ``on_step(itime, info)``
    is called after every step."""
    nl = namelists
    st = State(nl)
    if batch is None:
        return _integrate(st, records, max_steps, write, on_step)
    batch.attach(st)
    try:
        return _integrate(st, records, max_steps, write, on_step)
    finally:
        batch.detach(st)


def _integrate(st: State, records: bool, max_steps: int, write: bool, on_step
               ) -> Tuple[State, List[dict], List[tuple]]:
    from . import output

    nl = st.namelists
    recs: List[dict] = []
    steps: List[tuple] = []
    if nl.poisson.initial_cleaning:
        st.initial_cleaning()
    if nl.output.restart:
        time_s, fields = output.read_input(nl.output.input_file, nl.output.iin)
        st.restart(time_s, fields)
    if records:
        recs.append(output.record_datasets(nl, st.get, st.time()[0], True))
    n = nl.output.iterations if nl.output.output_steps else max_steps
    for itime in range(1, n + 1):
        info = st.step()
        steps.append((info.time, info.dt, info.niter[0], info.niter[1]))
        if on_step is not None:
            on_step(itime, info)
        if records and ((nl.output.output_steps and itime % nl.output.nout == 0) or
                        (not nl.output.output_steps and info.output)):
            recs.append(output.record_datasets(nl, st.get, info.time, False))
        if not nl.output.output_steps and info.time >= nl.output.tmax:
            break
    if write and records:
        output.write_file(nl.output.output_file, recs)
    return st, recs, steps
