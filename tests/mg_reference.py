"""FP64 numpy restatement of the multigrid preconditioner of the mixed solve
(test infrastructure: the GPU V-cycle runs in FP32 and is checked against
this one).

The operator is given as the 25 coefficient arrays of compute_operator!
(x fastest, periodic in x and y, zero beyond the walls in z).  Level l+1 is
the Galerkin product R A P of level l with piecewise-constant aggregation of
fx x fy columns (a dimension is halved while it is even and at least 8).
The smoother is the pseudo-time step of apply_preconditioner!
(apply_preconditioner!.jl:58-111): (1 - deta L_v) s' = s + deta (L_h s - b),
with deta scaled per level by the aggregation of the couplings.  V-cycle:
2 pre-smoothing steps from 0, residual restriction, coarse correction with
weight 1.5, 2 post-smoothing steps; 20 steps from 0 on the coarsest level
(the defaults of mg_smoothing_steps and mg_coarse_steps).  round_apply
emulates a reduced storage of the coefficients where the device's row
kernels apply them: the horizontal part of the smoother on the levels with
more than row_columns columns, and their residual where it is restricted by
the fused 2 x 2 kernel (nx % 4 == 0); the line solve keeps the unrounded
coefficients.
"""
import numpy as np

NAMES = ["AC", "AL", "AR", "AB", "AF", "AD", "AU", "ALD", "ALU", "ARD", "ARU", "ABD", "ABU", "AFD", "AFU",
         "ADD", "AUU", "ALDD", "ALUU", "ARDD", "ARUU", "ABDD", "ABUU", "AFDD", "AFUU"]
OFF = dict(zip(NAMES, [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
                       (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1), (0, -1, -1), (0, -1, 1), (0, 1, -1),
                       (0, 1, 1), (0, 0, -2), (0, 0, 2), (-1, 0, -2), (-1, 0, 2), (1, 0, -2), (1, 0, 2),
                       (0, -1, -2), (0, -1, 2), (0, 1, -2), (0, 1, 2)]))
KEY = {v: k for k, v in OFF.items()}
NU, OMEGA, COARSE_STEPS, MIN_SIZE = 2, 1.5, 20, 8


def shift(s, di, dj, dk):
    """s at (i+di, j+dj, k+dk): periodic in x and y, zero beyond the walls."""
    out = np.roll(s, (-di, -dj), axis=(0, 1))
    if dk:
        z = np.zeros_like(out)
        if dk > 0:
            z[:, :, :-dk] = out[:, :, dk:]
        else:
            z[:, :, -dk:] = out[:, :, :dk]
        out = z
    return out


def apply(C, s, horizontal=False):
    v = np.zeros_like(s)
    for n, c in C.items():
        if horizontal and n in ("AC", "AD", "AU"):
            continue
        v += c * shift(s, *OFF[n])
    return v


def line_solve(C, deta, rhs):
    """(1 - deta tridiag(ad, ac, au)) s = rhs per column (Thomas)."""
    ac, ad, au = C["AC"], C["AD"], C["AU"]
    nz = rhs.shape[2]
    s = np.empty_like(rhs)
    q = np.empty_like(rhs)
    p = 1 - deta * ac[:, :, 0]
    q[:, :, 0] = deta * au[:, :, 0] / p
    s[:, :, 0] = rhs[:, :, 0] / p
    for k in range(1, nz):
        p = 1 / (1 - deta * ac[:, :, k] - deta * ad[:, :, k] * q[:, :, k - 1])
        q[:, :, k] = deta * au[:, :, k] * p
        s[:, :, k] = (rhs[:, :, k] + deta * ad[:, :, k] * s[:, :, k - 1]) * p
    for k in range(nz - 2, -1, -1):
        s[:, :, k] += q[:, :, k] * s[:, :, k + 1]
    return s


def coarsen(C, fx, fy):
    nx, ny, nz = C["AC"].shape
    out = {n: np.zeros((nx // fx, ny // fy, nz)) for n in NAMES}
    for n, c in C.items():
        di, dj, dk = OFF[n]
        for a in range(fx):
            for b in range(fy):
                out[KEY[((a + di) // fx, (b + dj) // fy, dk)]] += c[a::fx, b::fy, :]
    return out


def levels(C, dx, dy, dtau=1.0, round_apply=None, row_columns=1024):
    cx, cy = 1.0 / dx ** 2, 1.0 / dy ** 2
    lv = [{"C": C, "deta": dtau / (2 * (cx + cy))}]
    while True:
        nx, ny, _ = lv[-1]["C"]["AC"].shape
        fx = 2 if nx % 2 == 0 and nx >= MIN_SIZE else 1
        fy = 2 if ny % 2 == 0 and ny >= MIN_SIZE else 1
        if fx == 1 and fy == 1:
            break
        lv[-1]["f"] = (fx, fy)
        cx, cy = cx * fy, cy * fx
        lv.append({"C": coarsen(lv[-1]["C"], fx, fy), "deta": dtau / (2 * (cx + cy))})
    if round_apply is not None:
        for L in lv:
            nx, ny, _ = L["C"]["AC"].shape
            if nx * ny > row_columns:
                L["A"] = {n: round_apply(c) for n, c in L["C"].items()}
                if L.get("f") == (2, 2) and nx % 4 == 0:
                    L["R"] = L["A"]
    return lv


def smooth(L, s, b, steps):
    A = L.get("A", L["C"])
    for _ in range(steps):
        s = line_solve(L["C"], L["deta"], s + L["deta"] * (apply(A, s, True) - b))
    return s


def vcycle(lv, b, l=0, nu=NU, coarse_steps=COARSE_STEPS):
    L = lv[l]
    if l == len(lv) - 1:
        return smooth(L, np.zeros_like(b), b, coarse_steps)
    fx, fy = L["f"]
    s = smooth(L, np.zeros_like(b), b, nu)
    r = b - apply(L.get("R", L["C"]), s)
    rc = sum(r[a::fx, c::fy, :] for a in range(fx) for c in range(fy))
    e = vcycle(lv, rc, l + 1, nu, coarse_steps)
    s = s + OMEGA * np.repeat(np.repeat(e, fx, axis=0), fy, axis=1)
    return smooth(L, s, b, nu)


def bf16(c):
    """FP64 -> FP32 -> bf16 (round to nearest even), returned as FP64."""
    u = np.asarray(c, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
