"""ecm_mlp_forward_noalloc -- the proposal network's forward pass on device
buffers with its scratch in the library's static per-GPU array, no
allocation -- against ecm_mlp_forward (host buffers, workspace allocated by
the handle) on the same samples: two calls in a row, the second with fewer
samples, each equal bit for bit; a call beyond the scratch's capacity is
rejected before anything is launched."""
import ctypes as C

import numpy as np
import pytest

from ecmgenerator_amd import _native as N
from ecmgenerator_amd.learned import ProposalNet

pytestmark = pytest.mark.gpu

MAX_SAMPLES = 32768          # ECM_MLP_NOALLOC_MAX_SAMPLES


class DeviceArray:
    """A float32 device buffer through the HIP runtime (test plumbing)."""

    def __init__(self, hip, n_floats):
        self.hip = hip
        self.nbytes = 4 * n_floats
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def download(self, shape):
        out = np.empty(shape, np.float32)
        assert out.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def samples(seed, n):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 64), np.float32)
    x[:, :61] = rng.normal(0, 1.5, (n, 61)).astype(np.float32)
    return x


def test_noalloc_forward_equals_forward():
    hip = C.CDLL("libamdhip64.so")
    net = ProposalNet.load()
    dx = DeviceArray(hip, MAX_SAMPLES * 64)
    dy = DeviceArray(hip, MAX_SAMPLES * 16)
    try:
        for seed, n in ((1, 20000), (2, 777)):
            x = samples(seed, n)
            want = net.forward(x)
            dx.upload(x)
            net.forward_device(dx.ptr.value, n, dy.ptr.value)
            got = dy.download((n, 16))
            bad = int(np.sum(~((got == want) | (np.isnan(got) & np.isnan(want))).all(axis=1)))
            print(f"{n} samples: {bad} rows differ, max |diff| {np.abs(got - want).max():.3e}", flush=True)
            assert bad == 0
        with pytest.raises(N.NativeError, match="exceed"):
            net.forward_device(dx.ptr.value, MAX_SAMPLES + 1, dy.ptr.value)
    finally:
        dx.free()
        dy.free()
        net.close()
