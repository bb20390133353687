"""Output datasets and their norms (src/Output/write_output.jl:73-379,
test/TestTools/compute_norms.jl:10-28).

``record_datasets`` builds, from host copies of the fields, the datasets
write_output stores for one record (FP32, as the reference's HDF5 files);
``norms`` reduces the records the way compute_norms reduces the file: the L2
and L-infinity norm of every dataset, time-dependent ones over all records.
"""
from __future__ import annotations

import math
from typing import Callable, Dict, List

import numpy as np

from .namelists import Namelists


def constants(nl: Namelists) -> Dict[str, float]:
    """Reference quantities of Constants.jl:108-182."""
    gamma = 1.4
    rsp, g = 287.0, 9.81
    rhoref, pref = 1.184, 101325.0
    aref = math.sqrt(pref / rhoref)
    uref = aref
    lref = pref / rhoref / g
    tref = lref / aref
    thetaref = aref * aref / rsp
    return dict(gamma=gamma, rsp=rsp, g=g, rhoref=rhoref, pref=pref, uref=uref, lref=lref, tref=tref,
                thetaref=thetaref)


def _cartesian_w(u, v, w, jac, met13, met23):
    """compute_vertical_wind (compute_vertical_wind.jl:27-63, transform.jl:86-122)
    at every (i, j, k) with i, j >= 2 and k <= nzz - 1 (others left 0)."""
    out = np.zeros_like(w)
    jc, ju = jac[1:, 1:, :-1], jac[1:, 1:, 1:]
    jacedgeu = 2.0 * jc * ju / (jc + ju)
    uc = 0.5 * (u[1:, 1:, :-1] + u[:-1, 1:, :-1])
    uu = 0.5 * (u[1:, 1:, 1:] + u[:-1, 1:, 1:])
    vc = 0.5 * (v[1:, 1:, :-1] + v[1:, :-1, :-1])
    vu = 0.5 * (v[1:, 1:, 1:] + v[1:, :-1, 1:])
    out[1:, 1:, :-1] = jacedgeu * (
        -(ju * (met13[1:, 1:, :-1] * uc + met23[1:, 1:, :-1] * vc)
          + jc * (met13[1:, 1:, 1:] * uu + met23[1:, 1:, 1:] * vu)) / (jc + ju) + w[1:, 1:, :-1])
    return out


def record_datasets(nl: Namelists, get: Callable[[str], np.ndarray], time_s: float, first: bool) -> Dict[str, np.ndarray]:
    """Datasets of one write_output record.  ``get(name)`` returns a haloed
    field (Fortran-ordered nxx x nyy x nzz) by its pcf_field name."""
    c = constants(nl)
    d = nl.domain
    nbx, nby, nbz = d.nbx, d.nby, d.nbz
    ii = slice(nbx, nbx + d.x_size)
    jj = slice(nby, nby + d.y_size)
    kk = slice(nbz, nbz + d.z_size)
    out: Dict[str, np.ndarray] = {"t": np.array([time_s], dtype=np.float32)}
    model = nl.atmosphere.model
    if first:
        lx, ly = d.lx / c["lref"], d.ly / c["lref"]
        dx, dy = lx / d.x_size, ly / d.y_size
        x = np.array([-lx / 2 + (i - nbx) * dx + dx / 2 for i in range(d.x_size + 2 * nbx)])
        y = np.array([-ly / 2 + (j - nby) * dy + dy / 2 for j in range(d.y_size + 2 * nby)])
        out["x"] = (x[ii] * c["lref"]).astype(np.float32)
        out["y"] = (y[jj] * c["lref"]).astype(np.float32)
        out["z"] = (get("ZC")[ii, jj, kk] * c["lref"]).astype(np.float32)
        out["ztilde"] = (get("ZCTILDE")[ii, jj, nbz - 1:nbz + d.z_size] * c["lref"]).astype(np.float32)
        if model != "Boussinesq":
            out["rhobar"] = (get("RHOBAR")[ii, jj, kk] * c["rhoref"]).astype(np.float32)
            out["thetabar"] = (get("THETABAR")[ii, jj, kk] * c["thetaref"]).astype(np.float32)
            out["n2"] = (get("N2")[ii, jj, kk] / c["tref"] ** 2).astype(np.float32)
            if model != "Compressible":
                out["p"] = (get("PBAR")[ii, jj, kk] * c["rhoref"] * c["thetaref"]).astype(np.float32)
    if model == "Compressible":  # P is predicted: written with every record
        out["p"] = (get("P")[ii, jj, kk] * c["rhoref"] * c["thetaref"]).astype(np.float32)
    o = nl.output
    rho, rhop = get("RHO"), get("RHOP")
    if o.prepare_restart or "rhop" in o.output_variables:
        src = rhop if model == "Boussinesq" else rho
        out["rhop"] = (src[ii, jj, kk] * c["rhoref"]).astype(np.float32)
    u, v, w = get("U"), get("V"), get("W")
    if o.prepare_restart or "us" in o.output_variables:
        out["us"] = (u[ii, jj, kk] * c["uref"]).astype(np.float32)
    if o.prepare_restart or "vs" in o.output_variables:
        out["vs"] = (v[ii, jj, kk] * c["uref"]).astype(np.float32)
    if "w" in o.output_variables:
        wc = _cartesian_w(u, v, w, get("JAC"), get("MET13"), get("MET23"))
        k0, k1 = nbz, nbz + d.z_size
        out["w"] = ((wc[ii, jj, k0:k1] + wc[ii, jj, k0 - 1:k1 - 1]) / 2 * c["uref"]).astype(np.float32)
    if o.prepare_restart or "wts" in o.output_variables:
        out["wts"] = (w[ii, jj, kk] * c["uref"]).astype(np.float32)
    if "thetap" in o.output_variables:
        pbar, rhobar, thetabar = get("PBAR"), get("RHOBAR"), get("THETABAR")
        src = rhop if model == "Boussinesq" else rho
        out["thetap"] = ((pbar[ii, jj, kk] / (rhobar[ii, jj, kk] + src[ii, jj, kk]) - thetabar[ii, jj, kk])
                         * c["thetaref"]).astype(np.float32)
    if o.prepare_restart or "pip" in o.output_variables:
        out["pip"] = get("PIP")[ii, jj, kk].astype(np.float32)
    if nl.tracer.tracer_setup == "TracerOn":  # write_output.jl:274-286
        rhobar = get("RHOBAR")
        out["chi"] = (get("CHI")[ii, jj, kk] / (rhobar[ii, jj, kk] + rho[ii, jj, kk])).astype(np.float32)
    return out


def norms(records: List[Dict[str, np.ndarray]]):
    """(L2, Linf) of every dataset over all records (compute_norms.jl:10-28)."""
    keys = set()
    for r in records:
        keys.update(r)
    l2, linf = {}, {}
    for key in sorted(keys):
        parts = [r[key].astype(np.float64).ravel() for r in records if key in r]
        a = np.concatenate(parts)
        l2[key] = float(np.sqrt(np.sum(a * a)))
        linf[key] = float(np.max(np.abs(a))) if a.size else 0.0
    return l2, linf


# datasets written once, with the first record (write_output.jl:73-165);
# "p" is time-dependent in the compressible model
STATIC = ("x", "y", "z", "ztilde", "rhobar", "thetabar", "n2", "p")


def write_file(path: str, records: List[Dict[str, np.ndarray]]) -> None:
    """Store output records the way create_output / write_output lay out the
    reference's HDF5 file (create_output.jl, write_output.jl:73-379): the
    grid and background datasets once, "t" as a vector over records and every
    time-dependent field as an (x, y, z, t) array.  The container is a numpy
    .npz archive (no HDF5 library in this environment); dataset names and
    units are the reference's."""
    out: Dict[str, np.ndarray] = {}
    keys = []
    for r in records:
        keys += [k for k in r if k not in keys]
    for k in keys:
        parts = [r[k] for r in records if k in r]
        if k in STATIC and len(parts) == 1:
            out[k] = parts[0]
        elif k == "t":
            out[k] = np.concatenate(parts).astype(np.float32)
        else:
            out[k] = np.stack(parts, axis=-1)
    with open(path, "wb") as f:
        np.savez(f, **out)


def read_input(path: str, iin: int = -1):
    """read_input! (read_input!.jl:13-105) for the datasets this library
    integrates: returns (time in s, {rhop, us, vs, wts, pip[, p][, chi]}) of record
    iin (1-based as in the reference; -1 = the last record)."""
    with np.load(path, allow_pickle=False) as data:
        idx = -1 if iin == -1 else iin - 1
        time_s = float(data["t"][idx])
        fields = {k: np.array(data[k][..., idx]) for k in ("rhop", "us", "vs", "wts", "pip", "chi") if k in data}
        if "p" in data:  # the compressible model's P (read_input!.jl:73-77); one record is stored 3-D
            fields["p"] = np.array(data["p"][..., idx] if data["p"].ndim == 4 else data["p"])
    missing = {"rhop", "us", "vs", "wts", "pip"} - set(fields)
    if missing:
        raise ValueError(f"restart file {path} lacks {sorted(missing)} (write it with prepare_restart = true)")
    return time_s, fields
