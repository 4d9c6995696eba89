"""qms_gram_schmidt_noalloc, the allocation-free variant of qms_gram_schmidt
(no counterpart in the reference), through the C ABI on the GPU.

The caller holds every buffer on the device (allocated here through the HIP
runtime the library is linked against); the entry point allocates nothing and keeps its partial sums
in the library's static device scratch.  Bar: on buffers of a handle's size
(n = 4 n1) the results are bit for bit those of qms_gram_schmidt on the same
inputs, over two calls in a row, the second with a smaller vector and fewer
basis vectors, so the scratch left by the first call cannot leak into the
second.  A length that is not a multiple of 4 (one segment) is checked
against numpy FP64 to the Gram-Schmidt bars of test_gpu_krylov.py."""
import numpy as np
import pytest

from alpineqms_amd import configs, native
from tests.test_gpu_parity import solver_for

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers of the caller, through the HIP runtime libqms.so itself
    is linked against (the process's libamdhip64.so.7)."""

    def __init__(self):
        import ctypes
        native.load()
        self.c = ctypes
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.c.c_void_p()
        assert self.L.hipMalloc(self.c.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0        # host to device
        return p.value

    def get(self, p, n):
        out = np.empty(n, dtype=np.float64)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0    # device to host
        return out

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)


def _noalloc(basis, w):
    nv, n = basis.shape
    hip = _Hip()
    try:
        bd, wd = hip.put(basis), hip.put(w)
        hd = hip.put(np.full(nv + 1, np.nan))
        native.gram_schmidt_noalloc(n, nv, bd, wd, hd, 0)
        return hip.get(hd, nv + 1), hip.get(wd, n)
    finally:
        hip.close()


@pytest.mark.parametrize("cells, nv", [(17, 31), (9, 6)])
def test_noalloc_matches_gram_schmidt_bit_for_bit(cells, nv):
    # the two parameter cases run in this order: the second call is smaller
    # in both the vector length and the basis size
    d = configs.small_box_design(cells + 1, 7, nbox=2, ramp=False)
    with solver_for(d) as s:
        nx, ny, nz = s.nodes
        n = 4 * nx * ny * nz
        rng = np.random.default_rng(100 + nv)
        basis = rng.normal(size=(nv, n)) * rng.uniform(0.1, 10.0, size=(nv, 1))
        w = rng.normal(size=n) * 3.0
        h_ref, w_ref = s.gram_schmidt(basis, w)
        h, w2 = _noalloc(basis, w)
        assert np.array_equal(h, h_ref)
        assert np.array_equal(w2, w_ref)
        # and again on the same buffers' values: the scratch carries nothing over
        h3, w3 = _noalloc(basis, w)
        assert np.array_equal(h3, h_ref) and np.array_equal(w3, w_ref)


def test_noalloc_one_segment_matches_numpy():
    n, nv = 100_003, 9
    rng = np.random.default_rng(5)
    basis = rng.normal(size=(nv, n))
    w = rng.normal(size=n)
    h, w2 = _noalloc(basis, w)
    h_ref = basis @ w
    scale = np.linalg.norm(basis, axis=1) * np.linalg.norm(w)
    assert np.all(np.abs(h[:nv] - h_ref) <= 1e-13 * scale)
    w_ref = w - h[:nv] @ basis
    assert np.all(np.abs(w2 - w_ref) <= 1e-13 * (np.abs(w) + np.abs(h[:nv]) @ np.abs(basis)))
    assert h[nv] == pytest.approx(float(w2 @ w2), rel=1e-12)


def test_noalloc_refuses_out_of_range_arguments():
    # refused on the host before any GPU work
    with pytest.raises(native.QmsError):
        native.gram_schmidt_noalloc(1000, native.NOALLOC_MAX_VECTORS + 1, 1, 1, 1, 0)
    with pytest.raises(native.QmsError):
        native.gram_schmidt_noalloc(0, 4, 1, 1, 1, 0)
    with pytest.raises(native.QmsError):
        native.gram_schmidt_noalloc(1000, 4, 0, 1, 1, 0)
