"""pp_pom_pmt_hits_noalloc, the entry point that allocates nothing (its
scratch is a static device array), against pp_pom_pmt_hits on the same hits:
identical PMT ids and detected counts, on two calls in a row, the second
smaller than the first. The inputs and outputs are device buffers the test
allocates with the HIP runtime the library itself links."""
import ctypes
from ctypes import POINTER, c_int, c_int64, c_uint64, c_void_p

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(P):
    P._lib.load_library()  # loads libamdhip64 first, so this is the same runtime
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [POINTER(c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [c_void_p]
    h.hipMemcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def device_copy(hip, a):
    p = c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
    assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
    return p


def test_noalloc_equals_allocating_entry(P, hip):
    from photonprop import pom
    lib = P._lib.load_library()
    f = lib.pp_pom_pmt_hits_noalloc
    acc = pom.default_acceptance()
    grid = np.ascontiguousarray(acc.wavelengths, np.float32)
    a1 = np.ascontiguousarray(acc.acc_1, np.float32)
    a2 = np.ascontiguousarray(acc.acc_2, np.float32)
    orient = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    model = pom._model(acc, orient)
    rng = np.random.default_rng(11)
    for n, seed in ((600_000, 21), (70_000, 22)):
        rel = rng.normal(size=(n, 3))
        rel = (rel / np.linalg.norm(rel, axis=1, keepdims=True)).astype(np.float32)
        wl = rng.uniform(280, 820, n).astype(np.float32)
        pid = rng.integers(0, 1 << 40, n).astype(np.uint64)
        src = rng.integers(0, 300, n).astype(np.int32)
        print(f"noalloc call: n={n} seed={seed}", flush=True)
        bufs = [device_copy(hip, x) for x in (rel, wl, pid, src)]
        out_d = c_void_p()
        assert hip.hipMalloc(ctypes.byref(out_d), n * 4) == 0
        nd = c_int64(-1)
        rc = f(0, ctypes.byref(model), grid.ctypes.data, a1.ctypes.data, a2.ctypes.data, n,
               *bufs, c_uint64(seed), out_d, ctypes.byref(nd), None)
        assert rc == 0, P._lib.last_error()
        got = np.empty(n, np.int32)
        assert hip.hipMemcpy(got.ctypes.data, out_d, n * 4, 2) == 0  # device to host
        for b in (*bufs, out_d):
            hip.hipFree(b)
        ref = np.zeros(n, np.int32)
        t = {}
        ref = pom._pmt_ids(rel, wl, acc, orient, pid, src, seed, 0, timing=t)
        assert np.array_equal(got, ref)
        assert nd.value == t["n_detected"] == int(np.sum(ref > 0)) > 0


def test_noalloc_rejects_oversized_tables(P):
    from photonprop import pom
    lib = P._lib.load_library()
    f = lib.pp_pom_pmt_hits_noalloc
    acc = pom.default_acceptance()
    model = pom._model(acc, None)
    model.n_wl = 1025
    one = np.zeros(4, np.float32)
    rc = f(0, ctypes.byref(model), one.ctypes.data, one.ctypes.data, one.ctypes.data, 4,
           one.ctypes.data, one.ctypes.data, one.ctypes.data, one.ctypes.data, c_uint64(1),
           one.ctypes.data, None, None)
    assert rc == P._lib.PP_ERR_INVALID
    assert "n_wl" in P._lib.last_error()
