"""The initial-guess network on the device against its numpy restatement
(tests/cnn_reference.py, itself pinned by finite differences).

The device keeps the weights, the input, the activations and the backward
signal of the matrix-core layers in bf16 and accumulates in FP32.  The
forward check emulates the bf16 rounding of input, weights and activations
in numpy; FP32 summation order then still flips the odd activation by one bf16
step (0.4 %), which FORWARD_RTOL allows for.  The
gradient check compares against the exact FP64 gradient: bf16 operands in
the backward GEMMs leave a relative error of a few 1e-3 (GRAD_RTOL, in the
2-norm per layer).
"""
import ctypes

import hipmem
import numpy as np
import pytest

import cnn_reference as cnn

pytestmark = pytest.mark.gpu

FORWARD_RTOL = 1e-2
GRAD_RTOL = 3e-2


def bf16(a):
    """Round FP32 values to bf16 (round to nearest even), returned as FP32."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def flatten(params):
    return np.concatenate([np.concatenate([w.reshape(-1), b.reshape(-1)]) for w, b in params]).astype(np.float32)


def unflatten(vec):
    out, o = [], 0
    for cin, cout in zip(cnn.CHANNELS[:-1], cnn.CHANNELS[1:]):
        nw = 27 * cout * cin
        out.append((vec[o:o + nw].reshape(27, cout, cin).astype(np.float64), vec[o + nw:o + nw + cout].astype(np.float64)))
        o += nw + cout
    return out


def to_device_layout(x):
    """(C, nx, ny, nz) -> flat channels-last, voxel-major with x fastest."""
    return np.ascontiguousarray(np.transpose(x, (3, 2, 1, 0))).reshape(-1).astype(np.float32)


def from_device_layout(v, shape):
    nx, ny, nz = shape
    return np.transpose(v.reshape(nz, ny, nx), (2, 1, 0))


def forward_bf16(params, x):
    """cnn_reference.forward with the device's bf16 storage emulated (the
    input, every layer's weights and the activations; biases stay FP32)."""
    a = bf16(x).astype(np.float64)
    for n, (w, b) in enumerate(params):
        w = bf16(w).astype(np.float64)
        z = cnn.conv(a, w, b)
        a = z if n == len(params) - 1 else bf16(cnn.gelu(z)).astype(np.float64)
    return a


class Net:
    def __init__(self, lib, shape, training):
        self.lib = lib
        self.h = ctypes.c_void_p()
        rc = lib.pcf_net_create(*shape, int(training), ctypes.byref(self.h))
        assert rc == 0, lib.pcf_last_error()
        self.n = lib.pcf_net_param_count()

    def set_params(self, vec):
        vec = np.ascontiguousarray(vec, dtype=np.float32)
        assert self.lib.pcf_net_set_params(self.h, vec.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0

    def get_params(self):
        out = np.empty(self.n, dtype=np.float32)
        assert self.lib.pcf_net_get_params(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
        return out

    def close(self):
        self.lib.pcf_net_destroy(self.h)


def make_problem(shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((3,) + shape)
    target = rng.standard_normal((1,) + shape) * 0.5
    params = cnn.init_params(seed=seed)
    params = [(w, rng.standard_normal(b.shape) * 0.1) for w, b in params]
    return x, target, params


@pytest.mark.parametrize("shape", [(32, 4, 5), (64, 3, 4)])
def test_net_forward_matches_restatement(pcf, shape):
    lib = pcf._abi.load_library()
    x, _, params = make_problem(shape, 1)
    net = Net(lib, shape, False)
    assert net.n == flatten(params).size
    net.set_params(flatten(params))
    xin = to_device_layout(x)
    dx = hipmem.DeviceArray(xin.shape, dtype=np.float32).upload(xin)
    dy = hipmem.DeviceArray((xin.size // 3,), dtype=np.float32)
    assert lib.pcf_net_forward(net.h, ctypes.c_void_p(dx.ptr), ctypes.c_void_p(dy.ptr)) == 0
    got = from_device_layout(dy.download(), shape)
    want = forward_bf16(unflatten(flatten(params)), x)[0]
    err = float(np.max(np.abs(got - want))) / float(np.max(np.abs(want)))
    print("forward rel err", err)
    assert err <= FORWARD_RTOL
    net.close()


def test_net_gradients_match_restatement(pcf):
    lib = pcf._abi.load_library()
    shape = (32, 4, 5)
    x, target, params = make_problem(shape, 2)
    p32 = unflatten(flatten(params))
    net = Net(lib, shape, True)
    net.set_params(flatten(params))
    xin = to_device_layout(x)
    tin = to_device_layout(target)
    dx = hipmem.DeviceArray(xin.shape, dtype=np.float32).upload(xin)
    dt = hipmem.DeviceArray(tin.shape, dtype=np.float32).upload(tin)
    g = np.empty(net.n, dtype=np.float32)
    loss = ctypes.c_double(0.0)
    assert lib.pcf_net_gradients(net.h, ctypes.c_void_p(dx.ptr), ctypes.c_void_p(dt.ptr),
                                 g.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(loss)) == 0
    want_loss, want = cnn.loss_and_grads(p32, x, target)
    assert loss.value == pytest.approx(want_loss, rel=2e-2)
    got = unflatten(g)
    for n, ((gw, gb), (ww, wb)) in enumerate(zip(got, want)):
        for a, b in ((gw, ww), (gb, wb)):
            err = float(np.linalg.norm(a - b)) / max(float(np.linalg.norm(b)), 1e-30)
            print("layer", n, "grad rel err", err)
            assert err <= GRAD_RTOL, (n, err)
    net.close()


def test_training_reduces_the_loss(pcf):
    lib = pcf._abi.load_library()
    shape = (32, 8, 6)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((3,) + shape)
    # a target the network can represent: a smoothed copy of the first input
    target = sum(cnn.shift(x[:1], dx, dy, dz) for dx, dy, dz in cnn.TAPS) / 27.0
    params = cnn.init_params(seed=4)
    net = Net(lib, shape, True)
    net.set_params(flatten(params))
    xin = to_device_layout(x)
    tin = to_device_layout(target)
    dx = hipmem.DeviceArray(xin.shape, dtype=np.float32).upload(xin)
    dt = hipmem.DeviceArray(tin.shape, dtype=np.float32).upload(tin)
    losses = []
    for _ in range(60):
        loss = ctypes.c_double(0.0)
        assert lib.pcf_net_train_step(net.h, ctypes.c_void_p(dx.ptr), ctypes.c_void_p(dt.ptr), ctypes.c_float(3e-3),
                                      ctypes.byref(loss)) == 0
        losses.append(loss.value)
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < 0.2 * losses[0]
    net.close()
