"""The key sort that allocates nothing (GPU), through the C ABI.

mpm_sort_keys_noalloc takes device positions and writes device outputs,
using only the library's fixed device scratch.  Its keys and permutation must
equal mpm_engine_get_sort's (the reference's setKeys + stable key sort,
kernels.cu:361-371, engine.cu:92-102) bit for bit, on two calls in a row --
the second with fewer particles, so that what the first left in the shared
scratch must not leak into it.  Sizes above its stated maximum are refused
before anything runs.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class DeviceBuffer:
    def __init__(self, lib, nbytes):
        self.lib, self.ptr = lib, ctypes.c_void_p()
        assert lib.mpm_device_alloc(nbytes, ctypes.byref(self.ptr)) == 0, lib.mpm_last_error()

    def close(self):
        if self.ptr.value:
            assert self.lib.mpm_device_free(self.ptr) == 0
            self.ptr = ctypes.c_void_p()


def positions(n, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros((26, n), dtype=np.float32)
    p[0:3] = rng.uniform(-10.0, 700.0, size=(3, n))
    t = n // 4
    p[0:3, :t] = np.floor(rng.uniform(0, 6, size=(3, t))) + 0.25  # many equal keys
    return p


def engine_sort(mpm, p):
    cfg = mpm.make_config(SIZE_X=64, SIZE_Y=64, SIZE_Z=64)
    e = mpm.Engine(cfg)
    e.init(p.shape[1], p)
    keys, perm = e.sortParticles()
    e.close()
    return keys, perm


def noalloc_sort(lib, p, bufs):
    n = p.shape[1]
    pos = np.ascontiguousarray(p[0:3])
    px, keys_d, perm_d = bufs
    assert lib.mpm_copy_to_device(px.ptr, pos.ctypes.data, pos.nbytes) == 0
    base = px.ptr.value
    rc = lib.mpm_sort_keys_noalloc(ctypes.c_void_p(base), ctypes.c_void_p(base + 4 * n),
                                   ctypes.c_void_p(base + 8 * n), n, keys_d.ptr, perm_d.ptr, None)
    assert rc == 0, lib.mpm_last_error()
    keys = np.empty(n, dtype=np.int32)
    perm = np.empty(n, dtype=np.int32)
    assert lib.mpm_copy_to_host(keys.ctypes.data, keys_d.ptr, keys.nbytes) == 0  # waits for the null stream
    assert lib.mpm_copy_to_host(perm.ctypes.data, perm_d.ptr, perm.nbytes) == 0
    return keys, perm


def test_noalloc_sort_matches_engine_sort_twice(gpu):
    lib = gpu.load_library()
    n1, n2 = 300_000, 7_001
    bufs = [DeviceBuffer(lib, 3 * 4 * n1), DeviceBuffer(lib, 4 * n1), DeviceBuffer(lib, 4 * n1)]
    try:
        for n, seed in ((n1, 31), (n2, 32)):
            say(f"no-alloc sort: n={n} seed={seed}")
            p = positions(n, seed)
            k_ref, p_ref = engine_sort(gpu, p)
            k, perm = noalloc_sort(lib, p, bufs)
            np.testing.assert_array_equal(perm, p_ref)
            np.testing.assert_array_equal(k, k_ref)
            assert (np.diff(k.astype(np.int64)) >= 0).all()
    finally:
        for b in bufs:
            b.close()


def test_noalloc_sort_edge_sizes(gpu):
    lib = gpu.load_library()
    # n = 0 does nothing; n above the maximum is refused before any launch
    assert lib.mpm_sort_keys_noalloc(None, None, None, 0, None, None, None) == 0
    rc = lib.mpm_sort_keys_noalloc(None, None, None, gpu.SORT_NOALLOC_MAX + 1, None, None, None)
    assert rc == gpu.MPM_ELIMIT and b"exceed" in lib.mpm_last_error()
    assert lib.mpm_sort_keys_noalloc(None, None, None, -1, None, None, None) != 0
    # a single particle
    bufs = [DeviceBuffer(lib, 12), DeviceBuffer(lib, 4), DeviceBuffer(lib, 4)]
    try:
        p = positions(1, 5)
        k, perm = noalloc_sort(lib, p, bufs)
        k_ref, p_ref = engine_sort(gpu, p)
        assert list(perm) == [0] and list(k) == list(k_ref)
    finally:
        for b in bufs:
            b.close()
