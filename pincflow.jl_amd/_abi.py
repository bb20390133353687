"""ctypes mirror of include/pincflow_hip.h and the loader of libpincflow_hip.so.

The structure layout below must match ``pcf_config`` field for field; the
library reports ``sizeof(pcf_config)`` through ``pcf_config_size`` and the
CPU tests compare it with ``ctypes.sizeof(PcfConfig)``.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libpincflow_hip.so"

ABI_VERSION = 8
MAX_HILLS = 32

# return codes
OK = 0
E_NOCONV = 1
E_DT_ZERO = -2
E_DIVNORM = -3
E_DTMIN = -4
E_OPTION = -5
E_CONFIG = -6
E_HIP = -7
E_ARG = -8

# enums
BOUSSINESQ, PSEUDO_INCOMPRESSIBLE, COMPRESSIBLE = 0, 1, 2
NEUTRAL_STRATIFICATION, STABLE_STRATIFICATION, ISOTHERMAL, ISENTROPIC, REALISTIC, LAPSE_RATES = range(6)
TOPO_FLAT, TOPO_COSINE_HILL, TOPO_AGNESI, TOPO_GAUSSIANS = range(4)
STRETCH_NONE, STRETCH_SINH = 0, 1
INIT_RHOP, INIT_THETAP = 0, 1
PERT_NONE, PERT_COSINE_BUBBLE, PERT_VORTEX = 0, 1, 2
LHS_SPONGE_NONE, LHS_SPONGE_SIN2 = 0, 1
RHS_SPONGE_NONE, RHS_SPONGE_SIN2_TOP = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1
GUESS_ZERO, GUESS_PREVIOUS, GUESS_LEARNED, GUESS_EXTRAPOLATED = 0, 1, 2, 3
PC_LINE_RELAXATION, PC_MULTIGRID = 0, 1
STORAGE_FP32, STORAGE_BF16 = 0, 1

FIELDS = [
    "RHO", "RHOP", "U", "V", "W", "PIP",
    "PBAR", "THETABAR", "RHOBAR", "N2",
    "JAC", "MET13", "MET23", "MET33", "ZC", "ZCTILDE",
    "ALPHAR", "BETAR",
    "DPIP", "CORX", "CORY",
    "DRHO", "DRHOP", "DU", "DV", "DW",
    "RHOOLD", "RHOPOLD", "UOLD", "VOLD", "WOLD",
    "P0_RHO", "P0_RHOP", "P0_U", "P0_V", "P0_W",
    "PHIRHO_X", "PHIRHO_Y", "PHIRHO_Z",
    "PHIRHOP_X", "PHIRHOP_Y", "PHIRHOP_Z",
    "PHIU_X", "PHIU_Y", "PHIU_Z",
    "PHIV_X", "PHIV_Y", "PHIV_Z",
    "PHIW_X", "PHIW_Y", "PHIW_Z",
    "CHI", "DCHI", "INITIALTRACER", "CHI0",
    "PHICHI_X", "PHICHI_Y", "PHICHI_Z",
    "P", "DP", "P0_P",
    "PHIP_X", "PHIP_Y", "PHIP_Z",
    "PHITHETA_X", "PHITHETA_Y", "PHITHETA_Z",
]
FIELD = {name: i for i, name in enumerate(FIELDS)}

VECTORS = [
    "RHS", "SOLUTION",
    "AC", "AL", "AR", "AB", "AF", "AD", "AU",
    "ALD", "ALU", "ARD", "ARU", "ABD", "ABU",
    "AFD", "AFU", "ADD", "AUU", "ALDD", "ALUU",
    "ARDD", "ARUU", "ABDD", "ABUU", "AFDD", "AFUU",
]
VECTOR = {name: i for i, name in enumerate(VECTORS)}

BOUNDARY_PREDICTANDS, BOUNDARY_RECONSTRUCTIONS, BOUNDARY_FLUXES = 0, 1, 2
VAR_RHO, VAR_RHOP, VAR_U, VAR_V, VAR_W, VAR_PIP, VAR_CHI, VAR_P = range(8)
NO_TRACER, TRACER_ON = 0, 1
TRACER_ZERO, TRACER_DIPOLE = 0, 1
OP_TOTAL, OP_HORIZONTAL = 0, 1


class PcfHill(ctypes.Structure):
    _fields_ = [("height", ctypes.c_double), ("x0", ctypes.c_double),
                ("y0", ctypes.c_double), ("width", ctypes.c_double)]


class PcfConfig(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("x_size", ctypes.c_int32), ("y_size", ctypes.c_int32), ("z_size", ctypes.c_int32),
        ("nbx", ctypes.c_int32), ("nby", ctypes.c_int32), ("nbz", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("cfl_number", ctypes.c_double), ("dtmin", ctypes.c_double), ("dtmax", ctypes.c_double),
        ("adaptive_time_step", ctypes.c_int32),
        ("poisson_iterations", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("dtau", ctypes.c_double),
        ("preconditioner", ctypes.c_int32),
        ("preconditioner_iterations", ctypes.c_int32),
        ("initial_cleaning", ctypes.c_int32),
        ("tolerance_is_relative", ctypes.c_int32),
        ("model", ctypes.c_int32),
        ("background", ctypes.c_int32),
        ("specify_reynolds_number", ctypes.c_int32),
        ("buoyancy_initialization", ctypes.c_int32),
        ("inverse_reynolds_number", ctypes.c_double),
        ("kinematic_viscosity", ctypes.c_double),
        ("buoyancy_frequency", ctypes.c_double),
        ("potential_temperature", ctypes.c_double),
        ("temperature", ctypes.c_double),
        ("ground_pressure", ctypes.c_double),
        ("coriolis_frequency", ctypes.c_double),
        ("tropopause_height", ctypes.c_double),
        ("troposphere_lapse_rate", ctypes.c_double),
        ("stratosphere_lapse_rate", ctypes.c_double),
        ("initial_u", ctypes.c_double), ("initial_v", ctypes.c_double), ("initial_w", ctypes.c_double),
        ("perturbation", ctypes.c_int32), ("pad0", ctypes.c_int32),
        ("pert_amplitude", ctypes.c_double),
        ("pert_x0", ctypes.c_double), ("pert_y0", ctypes.c_double), ("pert_z0", ctypes.c_double),
        ("pert_rx", ctypes.c_double), ("pert_ry", ctypes.c_double), ("pert_rz", ctypes.c_double),
        ("stretching", ctypes.c_int32), ("topography", ctypes.c_int32),
        ("stretching_parameter", ctypes.c_double),
        ("topo_height", ctypes.c_double), ("topo_width", ctypes.c_double),
        ("n_hills", ctypes.c_int32), ("pad1", ctypes.c_int32),
        ("hills", PcfHill * MAX_HILLS),
        ("damp_horizontal_wind_on_rhs", ctypes.c_int32),
        ("relax_to_mean", ctypes.c_int32),
        ("lhs_sponge", ctypes.c_int32),
        ("rhs_sponge", ctypes.c_int32),
        ("lhs_alpharmax", ctypes.c_double),
        ("lhs_dxr", ctypes.c_double), ("lhs_dyr", ctypes.c_double), ("lhs_dzr", ctypes.c_double),
        ("lhs_divisor", ctypes.c_double),
        ("rhs_zr", ctypes.c_double),
        ("rhs_amplitude", ctypes.c_double),
        ("relaxed_u", ctypes.c_double), ("relaxed_v", ctypes.c_double), ("relaxed_w", ctypes.c_double),
        ("output_steps", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("nout", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("output_interval", ctypes.c_double),
        ("tmax", ctypes.c_double),
        ("initial_guess", ctypes.c_int32),
        ("preconditioner_type", ctypes.c_int32),
        ("mg_smoothing_steps", ctypes.c_int32),
        ("mg_coarse_steps", ctypes.c_int32),
        ("preconditioner_storage", ctypes.c_int32),
        ("tracer_setup", ctypes.c_int32),
        ("initial_tracer", ctypes.c_int32),
        ("tracer_rx", ctypes.c_double),
        ("tracer_ry", ctypes.c_double),
        ("thermal_conductivity", ctypes.c_double),
        ("flux_precision", ctypes.c_int32),
        ("pad2", ctypes.c_int32),
    ]


class PcfStepInfo(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("niter", ctypes.c_int32 * 2), ("output", ctypes.c_int32), ("pad", ctypes.c_int32)]


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PD = ctypes.POINTER(ctypes.c_double)

# (name, restype, argtypes) of every symbol include/pincflow_hip.h declares
SIGNATURES = [
    ("pcf_abi_version", _I, []),
    ("pcf_last_error", ctypes.c_char_p, []),
    ("pcf_config_default", None, [ctypes.POINTER(PcfConfig)]),
    ("pcf_config_size", _I, []),
    ("pcf_create", _I, [ctypes.POINTER(PcfConfig), ctypes.POINTER(_P)]),
    ("pcf_destroy", _I, [_P]),
    ("pcf_dims", _I, [_P, _PI32]),
    ("pcf_get_time", _I, [_P, _PD, _PI32]),
    ("pcf_set_time", _I, [_P, _D]),
    ("pcf_initial_cleaning", _I, [_P, _PI32]),
    ("pcf_step", _I, [_P, ctypes.POINTER(PcfStepInfo)]),
    ("pcf_compute_time_step", _I, [_P, _PD]),
    ("pcf_compute_sponges", _I, [_P, _D, _D]),
    ("pcf_explicit_integration_lhs", _I, [_P, _D, _D]),
    ("pcf_explicit_integration_rhs", _I, [_P, _D, _D]),
    ("pcf_implicit_integration", _I, [_P, _D, _D, _D, _PI32]),
    ("pcf_backup_predictands", _I, [_P]),
    ("pcf_set_boundaries", _I, [_P, _I]),
    ("pcf_reconstruct", _I, [_P]),
    ("pcf_compute_fluxes", _I, [_P]),
    ("pcf_compute_theta_fluxes", _I, [_P]),
    ("pcf_modify_compressible_wind", _I, [_P, _I]),
    ("pcf_synchronize_compressible_atmosphere", _I, [_P]),
    ("pcf_update_lhs", _I, [_P, _D, ctypes.c_int32, _I]),
    ("pcf_update_rhs", _I, [_P, _D, _I, _I, _D]),
    ("pcf_apply_lhs_sponge", _I, [_P, _D, _D, _I]),
    ("pcf_apply_corrector", _I, [_P, _D, _D, _PI32]),
    ("pcf_compute_lhs", _I, [_P, _PD]),
    ("pcf_compute_operator", _I, [_P, _D, _D]),
    ("pcf_apply_operator", _I, [_P, _P, _P, _I]),
    ("pcf_apply_operator_transpose", _I, [_P, _P, _P]),
    ("pcf_apply_preconditioner", _I, [_P, _P, _P]),
    ("pcf_apply_bicgstab", _I, [_P, _D, _PI32]),
    ("pcf_dot", _I, [_P, _P, _P, _PD]),
    ("pcf_dot_noalloc", _I, [_P, _P, ctypes.c_int64, _P, _P]),
    ("pcf_get_field", _I, [_P, _I, _PD]),
    ("pcf_set_field", _I, [_P, _I, _PD]),
    ("pcf_get_vector", _I, [_P, _I, _PD]),
    ("pcf_set_vector", _I, [_P, _I, _PD]),
    ("pcf_last_residual", _I, [_P, _PD]),
    ("pcf_net_param_count", _I, []),
    ("pcf_net_create", _I, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_P)]),
    ("pcf_net_destroy", _I, [_P]),
    ("pcf_net_set_params", _I, [_P, ctypes.POINTER(ctypes.c_float)]),
    ("pcf_net_get_params", _I, [_P, ctypes.POINTER(ctypes.c_float)]),
    ("pcf_net_forward", _I, [_P, _P, _P]),
    ("pcf_net_gradients", _I, [_P, _P, _P, ctypes.POINTER(ctypes.c_float), _PD]),
    ("pcf_net_train_step", _I, [_P, _P, _P, ctypes.c_float, _PD]),
    ("pcf_net_train_step_dout", _I, [_P, _P, _P, ctypes.c_float]),
    ("pcf_guess_input", _I, [_P, _P, _P, _P, _PD]),
    ("pcf_set_guess_network", _I, [_P, _P]),
    ("pcf_guess_fallbacks", _I, [_P, _PI32]),
    ("pcf_set_stage_network", _I, [_P, _P, ctypes.c_int32]),
    ("pcf_stage_fallbacks", _I, [_P, _PI32]),
    ("pcf_net_train_residual", _I, [_P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_float, _PD, _PD]),
    ("pcf_profile_groups", _I, [_P, ctypes.c_uint32]),
    ("pcf_profile_enable", _I, [_P, _I]),
    ("pcf_profile_query", _I, [_P, _I, ctypes.POINTER(ctypes.c_int64), _PD]),
    ("pcf_stage_tuning", _I, [_P, _PI32, ctypes.POINTER(ctypes.c_float)]),
    ("pcf_set_stage_shape", _I, [_P, ctypes.c_int32, ctypes.c_int32]),
    ("pcf_set_solver_graphs", _I, [_P, ctypes.c_int32]),
    ("pcf_ensemble_create", _I, [ctypes.POINTER(ctypes.c_void_p)]),
    ("pcf_ensemble_destroy", _I, [_P]),
    ("pcf_ensemble_attach", _I, [_P, _P]),
    ("pcf_ensemble_detach", _I, [_P]),
    ("pcf_ensemble_stats", _I, [_P, ctypes.POINTER(ctypes.c_int64)]),
]

KERNEL_GROUPS = ["apply_total_dot", "pc_thomas", "pc_rhs", "compute_operator", "bicg_vector", "fluxes", "updates",
                 "net"]

_lib = None


def load_library(path: os.PathLike | str | None = None) -> ctypes.CDLL:
    """Load libpincflow_hip.so.  There is no fallback: a missing library is an error."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(str(p))
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pcf_abi_version() != ABI_VERSION:
        raise RuntimeError("libpincflow_hip.so ABI version mismatch")
    if lib.pcf_config_size() != ctypes.sizeof(PcfConfig):
        raise RuntimeError("pcf_config layout mismatch between the header and _abi.py")
    if path is None:
        _lib = lib
    return lib
