"""The ctypes wrapper validates caller-supplied outputs before any library call: an output of the
other kind (host vs device) than the inputs, or of the wrong shape, raises instead of letting a
kernel write through a host pointer or past the end of a buffer."""
import numpy as np
import pytest

from mcpnet_amd import _lib


class FakeDeviceTensor:
    """Stands in for a torch tensor on the GPU (the wrapper reads only these attributes)."""
    __module__ = "torch.fake"

    def __init__(self, shape, dtype):
        import torch
        self.shape = tuple(shape)
        self.dtype = dtype
        self.is_cuda = True
        self._strides = (shape[1], 1) if len(shape) == 2 else (1,)

    def dim(self):
        return len(self.shape)

    def stride(self, i):
        return self._strides[i]

    def data_ptr(self):
        return 0x7F0000000000


class RecordingLib:
    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        def fn(*args):
            self.calls.append(name)
            return 0
        return fn


@pytest.fixture
def h():
    hd = object.__new__(_lib.Handle)
    hd.lib = RecordingLib()
    hd.h = None
    hd.device = 0
    return hd


def test_wrong_shape_out_is_rejected(h):
    A = np.ones((4, 5), np.float32)
    M = np.ones((4, 5), np.float32)
    with pytest.raises(ValueError, match="shape"):
        h.ratio(A, M, out=np.empty((4, 4), np.float32))
    with pytest.raises(ValueError, match="shape"):
        h.maxmin(A, M, out=np.empty((4, 5), np.float32))     # M x N = 4 x 4
    Sq = np.ones((6, 6), np.float32)
    with pytest.raises(ValueError, match="shape"):
        h.mcp(Sq, 3, out=np.empty((6, 6), np.float32), maxmin_out=np.empty((5, 6), np.float32))
    with pytest.raises(ValueError, match="shape"):
        h.bspline_mi(np.ones((8, 30), np.float32), out=np.empty((8, 7), np.float32))
    assert h.lib.calls == []


def test_wrong_dtype_out_is_rejected(h):
    A = np.ones((4, 5), np.float32)
    with pytest.raises(TypeError):
        h.ratio(A, A, out=np.empty((4, 5), np.float64))
    with pytest.raises(TypeError):
        h.rank(A, out=np.empty((4, 5), np.float32))          # ranks are int32
    assert h.lib.calls == []


def test_host_inputs_with_device_out_are_rejected(h):
    import torch
    A = np.ones((4, 5), np.float32)
    with pytest.raises(TypeError, match="output"):
        h.ratio(A, A, out=FakeDeviceTensor((4, 5), torch.float32))
    with pytest.raises(TypeError, match="output"):
        h.mcp(np.ones((6, 6), np.float32), 2, maxmin_out=FakeDeviceTensor((6, 6), torch.float32))
    assert h.lib.calls == []


def test_device_inputs_with_host_out_are_rejected(h):
    import torch
    A = FakeDeviceTensor((6, 6), torch.float32)
    with pytest.raises(TypeError, match="output"):
        h.mcp(A, 3, out=np.empty((6, 6), np.float32))
    with pytest.raises(TypeError, match="output"):
        h.bspline_mi(FakeDeviceTensor((6, 40), torch.float32), out=np.empty((6, 6), np.float32))
    with pytest.raises(TypeError, match="mixing"):
        h.maxmin(A, np.ones((6, 6), np.float32))
    with pytest.raises(TypeError, match="mixing"):
        h.mcpnet(FakeDeviceTensor((6, 20), torch.float32), tfs=np.ones(6, np.float32))
    assert h.lib.calls == []


def test_matching_out_reaches_the_library(h):
    A = np.ones((4, 5), np.float32)
    out = np.empty((4, 5), np.float32)
    assert h.ratio(A, A, out=out) is out
    assert h.lib.calls == ["mcpg_ratio_f32"]


def test_block_outside_matrix_is_rejected(h):
    X = np.ones((10, 20), np.float32)
    R = np.zeros((10, 20), np.int32)
    for args in ((0, 11, 0, 10), (5, 6, 0, 1), (0, 1, 9, 2), (-1, 2, 0, 2), (0, 2, 0, -1)):
        with pytest.raises(ValueError):
            h.bspline_mi_block(X, *args)
        with pytest.raises(ValueError):
            h.ap_mi_block(R, *args)
    with pytest.raises(ValueError, match="shape"):
        h.bspline_mi_block(X, 2, 3, 4, 5, out=np.empty((3, 4), np.float32))
    assert h.lib.calls == []
    h.bspline_mi_block(X, 2, 3, 4, 5)
    h.ap_mi_block(R, 0, 10, 0, 10)
    assert h.lib.calls == ["mcpg_bspline_mi_block_f32", "mcpg_ap_mi_block_i32"]
