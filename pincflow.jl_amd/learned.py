"""The learned initial guess of the mixed-precision solve.

A small 3D convolutional network (six 3x3x3 layers, 32 channels, GELU; the
bf16 matrix-core kernels of csrc/net.hip) corrects the warm start: from the
Poisson right-hand side b, the previous time step's solution of the same
implicit stage rescaled to this step (prev) and the terrain Jacobian it
gives x0 = prev + |b| net(b/|b|, prev/|b|, jac-1) (|b| the RMS norm; prev = 0
on a run's first step).  An untrained network (last layer zero) proposes the
warm start itself.  BiCGSTAB then starts from x0; its residual, measured
against |b| as for a zero start, still decides convergence, so the guess can
only change the iteration count, never the accuracy.

Training data are the solver's own solves: after every time step the RHS
vector and the solution of the step's second implicit stage are read back
(the solution times sqrt(pbar^2/rhobar) undoes the scaling solve_poisson!
applies in place, solve_poisson!.jl:69).  ``train`` fits the solution change
(mean squared error); ``train_residual`` fits the Poisson residual of the
guess, with the gradient through the operator's transpose.
"""
from __future__ import annotations

import ctypes
from typing import List, Tuple

import numpy as np

from . import _abi
from .state import PinCFlowError, State


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class DeviceBuffer:
    """A device allocation through the HIP runtime the library links."""

    _hip = None

    def __init__(self, nbytes: int):
        hip = DeviceBuffer.hip()
        p = ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise PinCFlowError(_abi.E_HIP, "hipMalloc failed")
        self.ptr, self.nbytes = p.value, nbytes

    @classmethod
    def hip(cls):
        if cls._hip is None:
            for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
                try:
                    cls._hip = ctypes.CDLL(name)
                    break
                except OSError:
                    continue
            if cls._hip is None:
                raise PinCFlowError(_abi.E_HIP, "libamdhip64.so not found")
            cls._hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            cls._hip.hipFree.argtypes = [ctypes.c_void_p]
            cls._hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        return cls._hip

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        if DeviceBuffer.hip().hipMemcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p), self.nbytes,
                                        1) != 0:
            raise PinCFlowError(_abi.E_HIP, "hipMemcpy failed")
        return self

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes == self.nbytes
        if DeviceBuffer.hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), self.nbytes,
                                        2) != 0:
            raise PinCFlowError(_abi.E_HIP, "hipMemcpy failed")
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            DeviceBuffer.hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None


class GuessNetwork:
    """One pcf_net handle (a grid size); parameters as one FP32 vector."""

    def __init__(self, nx: int, ny: int, nz: int, training: bool = False, params: np.ndarray | None = None,
                 seed: int = 0):
        self.lib = _abi.load_library()
        self.shape = (nx, ny, nz)
        self.h = ctypes.c_void_p()
        rc = self.lib.pcf_net_create(nx, ny, nz, int(training), ctypes.byref(self.h))
        if rc:
            raise PinCFlowError(rc, self.lib.pcf_last_error().decode())
        self.n_params = self.lib.pcf_net_param_count()
        self.set_params(params if params is not None else init_params(seed))

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {p.size}")
        rc = self.lib.pcf_net_set_params(self.h, _fptr(p))
        if rc:
            raise PinCFlowError(rc, self.lib.pcf_last_error().decode())

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        rc = self.lib.pcf_net_get_params(self.h, _fptr(out))
        if rc:
            raise PinCFlowError(rc, self.lib.pcf_last_error().decode())
        return out

    def train_step(self, d_input: DeviceBuffer, d_target: DeviceBuffer, lr: float) -> float:
        loss = ctypes.c_double(0.0)
        rc = self.lib.pcf_net_train_step(self.h, ctypes.c_void_p(d_input.ptr), ctypes.c_void_p(d_target.ptr),
                                         ctypes.c_float(lr), ctypes.byref(loss))
        if rc:
            raise PinCFlowError(rc, self.lib.pcf_last_error().decode())
        return loss.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.pcf_net_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def init_params(seed: int = 0) -> np.ndarray:
    """He-scaled weights, zero biases, as one FP32 vector (layer by layer
    W[27][cout][cin] then b[cout]); the last layer starts at zero, so an
    untrained network proposes x0 = 0."""
    rng = np.random.default_rng(seed)
    chans = (3, 32, 32, 32, 32, 32, 1)
    parts = []
    for n, (cin, cout) in enumerate(zip(chans[:-1], chans[1:])):
        w = rng.standard_normal(27 * cout * cin) * np.sqrt(2.0 / (27 * cin))
        if n == len(chans) - 2:
            w[:] = 0.0
        parts += [w, np.zeros(cout)]
    return np.concatenate(parts).astype(np.float32)


def _interior_fcscal(st: State) -> np.ndarray:
    d = st.namelists.domain
    ii, jj, kk = (slice(d.nbx, d.nbx + d.x_size), slice(d.nby, d.nby + d.y_size), slice(d.nbz, d.nbz + d.z_size))
    pbar, rhobar = st.get("PBAR")[ii, jj, kk], st.get("RHOBAR")[ii, jj, kk]
    return np.sqrt(pbar * pbar / rhobar)


def collect_samples(st: State, steps: int) -> List[Tuple[np.ndarray, np.ndarray]]:
    """Run `steps` time steps and return, per step after the first, the
    network input and target of the step's second implicit-stage solve:
    (input n x 3 FP32 channels-last, target n FP32 = (x - prev) / |b|)."""
    fcs = _interior_fcscal(st)
    n = st.nx * st.ny * st.nz
    d_b, d_prev = DeviceBuffer(8 * n), DeviceBuffer(8 * n)
    d_in = DeviceBuffer(12 * n)
    out = []
    prev, prev_dt = None, None
    for _ in range(steps):
        info = st.step()
        b = st.get_vector("RHS")
        x = st.get_vector("SOLUTION") * fcs
        if prev is not None:
            warm = np.asfortranarray(prev * (info.dt / prev_dt)).ravel(order="F")
            d_b.upload(np.asfortranarray(b).ravel(order="F"))
            d_prev.upload(warm)
            bnorm = ctypes.c_double(0.0)
            rc = st.lib.pcf_guess_input(st.h, ctypes.c_void_p(d_b.ptr), ctypes.c_void_p(d_prev.ptr),
                                        ctypes.c_void_p(d_in.ptr), ctypes.byref(bnorm))
            if rc:
                raise PinCFlowError(rc, st.lib.pcf_last_error().decode())
            inp = d_in.download(np.float32, (n * 3,))
            target = ((np.asfortranarray(x).ravel(order="F") - warm) / bnorm.value).astype(np.float32)
            out.append((inp, target))
        prev, prev_dt = x, info.dt
    return out


def train(net: GuessNetwork, samples, epochs: int, lr: float = 3e-4, log=None) -> List[float]:
    """Adam on the mean squared error over the samples, one sample per step.

    The targets are divided by their RMS over all samples, so the loss starts
    near 1 for the zero output of an untrained network; the scale is folded
    into the (linear) last layer afterwards, so the network's output is the
    unscaled x / |b|.  Returns the normalised loss per epoch."""
    n = samples[0][1].size
    scale = float(np.sqrt(np.mean([np.mean(t.astype(np.float64) ** 2) for _, t in samples]))) or 1.0
    d_in, d_t = DeviceBuffer(12 * n), DeviceBuffer(4 * n)
    losses = []
    for e in range(epochs):
        tot = 0.0
        for inp, target in samples:
            d_in.upload(inp)
            d_t.upload((target / scale).astype(np.float32))
            tot += net.train_step(d_in, d_t, lr)
        losses.append(tot / len(samples))
        if log:
            log(f"epoch {e + 1}: loss {losses[-1]:.4e}")
    p = net.params()
    p[-(27 * 32 + 1):] *= scale  # last layer: W[27][1][32] and b[1]
    net.set_params(p)
    return losses


def train_residual(net: GuessNetwork, states, steps: int, inner: int = 6, lr: float = 1e-3, log=None
                   ) -> List[Tuple[float, float]]:
    """Online training on the residual of the guess the network proposes,
    the quantity BiCGSTAB's stopping test and iteration count follow:

        L = |A x0 - b|^2 / |b|^2,  x0 = warm + |b| net(b / |b|, warm / |b|, J - 1)

    with A and b those of the solve.  Every time step of each state (after
    the first, which has no warm start) gives the solve of the step's second
    implicit stage, whose operator and right-hand side stay in the state, and
    `inner` Adam steps on it; the gradient with respect to the network output
    is dL/dnet = 2 |b| A^T (A x0 - b) / |b|^2 (A^T: the transposed apply).
    Returns, per solve, (L of the warm start, L after training on it)."""
    lib = states[0].lib
    out = []
    for st in states:
        n = st.nx * st.ny * st.nz
        fcs = _interior_fcscal(st).ravel(order="F")
        d_b, d_prev, d_x, d_ax = DeviceBuffer(8 * n), DeviceBuffer(8 * n), DeviceBuffer(8 * n), DeviceBuffer(8 * n)
        d_in, d_y = DeviceBuffer(12 * n), DeviceBuffer(4 * n)
        prev, prev_dt = None, None

        def loss_and_residual(x0, b, bb):
            d_x.upload(x0)
            st.apply_operator_device(d_x.ptr, d_ax.ptr, True)
            r = d_ax.download(np.float64, (n,)) - b
            return float(np.dot(r, r)) / bb, r

        for _ in range(steps):
            info = st.step()
            b = np.asfortranarray(st.get_vector("RHS")).ravel(order="F")
            x = np.asfortranarray(st.get_vector("SOLUTION")).ravel(order="F") * fcs
            if prev is not None:
                warm = prev * (info.dt / prev_dt)
                bb = float(np.dot(b, b))
                d_b.upload(b)
                d_prev.upload(warm)
                bnorm = ctypes.c_double(0.0)
                rc = lib.pcf_guess_input(st.h, ctypes.c_void_p(d_b.ptr), ctypes.c_void_p(d_prev.ptr),
                                         ctypes.c_void_p(d_in.ptr), ctypes.byref(bnorm))
                if rc:
                    raise PinCFlowError(rc, lib.pcf_last_error().decode())
                l_warm, _ = loss_and_residual(warm, b, bb)
                l = l_warm
                for _ in range(inner):
                    rc = net.lib.pcf_net_forward(net.h, ctypes.c_void_p(d_in.ptr), ctypes.c_void_p(d_y.ptr))
                    if rc:
                        raise PinCFlowError(rc, net.lib.pcf_last_error().decode())
                    y = d_y.download(np.float32, (n,)).astype(np.float64)
                    l, r = loss_and_residual(warm + bnorm.value * y, b, bb)
                    d_x.upload(r)
                    rc = lib.pcf_apply_operator_transpose(st.h, ctypes.c_void_p(d_x.ptr), ctypes.c_void_p(d_ax.ptr))
                    if rc:
                        raise PinCFlowError(rc, lib.pcf_last_error().decode())
                    g = d_ax.download(np.float64, (n,)) * (2.0 * bnorm.value / bb)
                    d_y.upload(g.astype(np.float32))
                    rc = net.lib.pcf_net_train_step_dout(net.h, ctypes.c_void_p(d_in.ptr), ctypes.c_void_p(d_y.ptr),
                                                         ctypes.c_float(lr))
                    if rc:
                        raise PinCFlowError(rc, net.lib.pcf_last_error().decode())
                out.append((l_warm, l))
                if log:
                    log(f"solve {len(out)}: residual loss of the warm start {l_warm:.4e}, of the guess {l:.4e}")
            prev, prev_dt = x, info.dt
    return out


OBJECTIVES = {"residual": 0, "energy": 1}


def residual_step(net: GuessNetwork, st: State, d_v: DeviceBuffer, d_base: DeviceBuffer | None, inner: int,
                  lr: float, objective: str = "residual") -> Tuple[float, float]:
    """`inner` Adam steps on the device (pcf_net_train_residual) for
    x = base + |v| net(v / |v|, base / |v|, J - 1), A the state's current
    operator, on the residual |A x - v|^2 / |v|^2 or the energy norm of the
    error (objective).  Returns (residual loss of base, residual loss with
    the network after the steps); inner = 0 only evaluates."""
    lb, ln = ctypes.c_double(0.0), ctypes.c_double(0.0)
    rc = st.lib.pcf_net_train_residual(st.h, net.h, ctypes.c_void_p(d_v.ptr),
                                       ctypes.c_void_p(d_base.ptr) if d_base is not None else None,
                                       OBJECTIVES[objective], int(inner), ctypes.c_float(lr), ctypes.byref(lb),
                                       ctypes.byref(ln))
    if rc:
        raise PinCFlowError(rc, st.lib.pcf_last_error().decode())
    return lb.value, ln.value


def train_online(states, steps: int, guess: GuessNetwork | None = None, stage: GuessNetwork | None = None,
                 inner: int = 4, lr: float = 1e-4, stage_lr: float | None = None, objective: str = "residual",
                 log=None):
    """Online training on the solver's own solves, all on the device.  Every
    time step of each state after the first gives the solve of the step's
    second implicit stage (its operator A and right-hand side b stay in the
    state) and `inner` Adam steps for each network given:

    * the initial guess: v = b, base = the warm start (the previous step's
      solution rescaled to this dt), so x0 = warm + |b| net(b/|b|, warm/|b|, J-1);
    * the preconditioning stage: v = b - A warm, the residual the solve's first
      preconditioner application sees, base = M^-1 v (the state's configured
      preconditioner), so the stage learns to correct what M^-1 leaves.

    `objective` is "residual" or "energy" (pcf_net_train_residual).  Returns
    per solve {"guess": (L warm, L learned), "stage": (L M^-1, L M^-1 + net)},
    the residual losses."""
    hist = []
    stage_lr = lr if stage_lr is None else stage_lr
    for st in states:
        n = st.nx * st.ny * st.nz
        fcs = _interior_fcscal(st).ravel(order="F")
        d_b, d_w, d_r, d_y = (DeviceBuffer(8 * n) for _ in range(4))
        prev, prev_dt = None, None
        for _ in range(steps):
            info = st.step()
            b = np.asfortranarray(st.get_vector("RHS")).ravel(order="F")
            x = np.asfortranarray(st.get_vector("SOLUTION")).ravel(order="F") * fcs
            if prev is not None:
                warm = prev * (info.dt / prev_dt)
                rec = {}
                d_b.upload(b)
                d_w.upload(warm)
                if guess is not None:
                    rec["guess"] = residual_step(guess, st, d_b, d_w, inner, lr, objective)
                if stage is not None:
                    st.apply_operator_device(d_w.ptr, d_r.ptr, True)
                    d_r.upload(b - d_r.download(np.float64, (n,)))
                    rc = st.lib.pcf_apply_preconditioner(st.h, ctypes.c_void_p(d_r.ptr), ctypes.c_void_p(d_y.ptr))
                    if rc:
                        raise PinCFlowError(rc, st.lib.pcf_last_error().decode())
                    rec["stage"] = residual_step(stage, st, d_r, d_y, inner, stage_lr, objective)
                hist.append(rec)
                if log:
                    log(f"solve {len(hist)}: " + ", ".join(f"{k} {a:.4e} -> {c:.4e}" for k, (a, c) in rec.items()))
            prev, prev_dt = x, info.dt
    return hist


def attach_stage(st: State, net: GuessNetwork, iterations: int):
    """Use net as the learned flexible-BiCGSTAB preconditioning stage of this
    state's mixed solves, in the first `iterations` iterations of each."""
    rc = st.lib.pcf_set_stage_network(st.h, net.h, int(iterations))
    if rc:
        raise PinCFlowError(rc, st.lib.pcf_last_error().decode())
    st._stage_net = net


def residual_loss(net: GuessNetwork, st: State, steps: int) -> List[Tuple[float, float]]:
    """Per solve of `steps` time steps of st (after the first): the residual
    loss |A x0 - b|^2 / |b|^2 of the warm start and of the network's guess."""
    n = st.nx * st.ny * st.nz
    lib = st.lib
    fcs = _interior_fcscal(st).ravel(order="F")
    d_b, d_prev, d_x, d_ax = DeviceBuffer(8 * n), DeviceBuffer(8 * n), DeviceBuffer(8 * n), DeviceBuffer(8 * n)
    d_in, d_y = DeviceBuffer(12 * n), DeviceBuffer(4 * n)
    out, prev, prev_dt = [], None, None
    for _ in range(steps):
        info = st.step()
        b = np.asfortranarray(st.get_vector("RHS")).ravel(order="F")
        x = np.asfortranarray(st.get_vector("SOLUTION")).ravel(order="F") * fcs
        if prev is not None:
            warm = prev * (info.dt / prev_dt)
            bb = float(np.dot(b, b))
            d_b.upload(b)
            d_prev.upload(warm)
            bnorm = ctypes.c_double(0.0)
            rc = lib.pcf_guess_input(st.h, ctypes.c_void_p(d_b.ptr), ctypes.c_void_p(d_prev.ptr),
                                     ctypes.c_void_p(d_in.ptr), ctypes.byref(bnorm))
            if rc:
                raise PinCFlowError(rc, lib.pcf_last_error().decode())
            rc = net.lib.pcf_net_forward(net.h, ctypes.c_void_p(d_in.ptr), ctypes.c_void_p(d_y.ptr))
            if rc:
                raise PinCFlowError(rc, net.lib.pcf_last_error().decode())
            y = d_y.download(np.float32, (n,)).astype(np.float64)
            res = []
            for x0 in (warm, warm + bnorm.value * y):
                d_x.upload(x0)
                st.apply_operator_device(d_x.ptr, d_ax.ptr, True)
                r = d_ax.download(np.float64, (n,)) - b
                res.append(float(np.dot(r, r)) / bb)
            out.append(tuple(res))
        prev, prev_dt = x, info.dt
    return out


def evaluate(net: GuessNetwork, samples) -> float:
    """Mean squared error of the network's output against the targets,
    relative to the zero output's (1 = no better than the warm start)."""
    n = samples[0][1].size
    d_in, d_out = DeviceBuffer(12 * n), DeviceBuffer(4 * n)
    num = den = 0.0
    for inp, target in samples:
        d_in.upload(inp)
        rc = net.lib.pcf_net_forward(net.h, ctypes.c_void_p(d_in.ptr), ctypes.c_void_p(d_out.ptr))
        if rc:
            raise PinCFlowError(rc, net.lib.pcf_last_error().decode())
        out = d_out.download(np.float32, (n,)).astype(np.float64)
        t = target.astype(np.float64)
        num += float(np.sum((out - t) ** 2))
        den += float(np.sum(t * t))
    return num / den


def attach(st: State, net: GuessNetwork):
    """Use net for initial_guess = "learned" of this state (same grid)."""
    rc = st.lib.pcf_set_guess_network(st.h, net.h)
    if rc:
        raise PinCFlowError(rc, st.lib.pcf_last_error().decode())
    st._guess_net = net  # keep the network alive with the state
