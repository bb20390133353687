"""pcf_dot_noalloc, the allocation-free variant of pcf_dot (no reference
counterpart): its block partials live in a file-scope device array, so it
must give pcf_dot's bits for the same vectors, on two calls in a row (the
second smaller, so stale partials of the first would show), and refuse sizes
outside its range with PCF_E_ARG and a message."""
import ctypes

import hipmem
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dot_pair(pcf, abi, nl, seed):
    """(pcf_dot, pcf_dot_noalloc) of two seeded interior-sized vectors of the state nl builds."""
    lib = abi.load_library()
    st = pcf.State(nl)
    try:
        n = st.nx * st.ny * st.nz
        rng = np.random.default_rng(seed)
        a = hipmem.DeviceArray((n,)).upload(rng.standard_normal(n))
        b = hipmem.DeviceArray((n,)).upload(rng.standard_normal(n) * 1e3)
        ref = ctypes.c_double()
        assert lib.pcf_dot(st.h, ctypes.c_void_p(a.ptr), ctypes.c_void_p(b.ptr), ctypes.byref(ref)) == 0
        out = hipmem.DeviceArray((1,)).upload(np.array([np.nan]))
        rc = lib.pcf_dot_noalloc(ctypes.c_void_p(a.ptr), ctypes.c_void_p(b.ptr), n, ctypes.c_void_p(out.ptr), None)
        assert rc == 0, lib.pcf_last_error()
        got = float(out.download()[0])
        return n, ref.value, got
    finally:
        st.close()


def test_noalloc_dot_matches_pcf_dot_on_two_calls(pcf, abi):
    big = pcf.examples.mountain_wave_3d(x_size=64, y_size=48, z_size=32)
    small = pcf.examples.periodic_hill(x_size=40, z_size=40, iterations=1)
    n1, ref1, got1 = dot_pair(pcf, abi, big, seed=5)
    print(f"first call: n = {n1}, pcf_dot {ref1!r}, pcf_dot_noalloc {got1!r}", flush=True)
    n2, ref2, got2 = dot_pair(pcf, abi, small, seed=6)
    print(f"second call: n = {n2}, pcf_dot {ref2!r}, pcf_dot_noalloc {got2!r}", flush=True)
    assert n2 < n1
    assert got1 == ref1  # the same kernels and launch shape: bit for bit
    assert got2 == ref2


def test_noalloc_dot_refuses_sizes_out_of_range(abi):
    lib = abi.load_library()
    a = hipmem.DeviceArray((4,)).upload(np.ones(4))
    out = hipmem.DeviceArray((1,))
    for n in (-1, (1 << 34) + 1):
        rc = lib.pcf_dot_noalloc(ctypes.c_void_p(a.ptr), ctypes.c_void_p(a.ptr), n, ctypes.c_void_p(out.ptr), None)
        assert rc == abi.E_ARG
        assert b"outside" in lib.pcf_last_error()
    rc = lib.pcf_dot_noalloc(None, ctypes.c_void_p(a.ptr), 4, ctypes.c_void_p(out.ptr), None)
    assert rc == abi.E_ARG
