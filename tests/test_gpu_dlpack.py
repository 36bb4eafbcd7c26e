"""DLPack link on the HIP device (the role of tomosipo's CuPy link,
links/cupy.py:25-159): a foreign array type is projected in place through a
zero-copy view, outputs come back as the same type, and results equal the
torch path bitwise."""
import numpy as np
import pytest

import tomosipo_amd as ts

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_dlpack_arrays_project_like_torch_tensors():
    from tests import dlpack_producer as dp

    vg = ts.volume(shape=24, size=1)
    pg = ts.cone(angles=20, shape=(24, 36), size=(1.4, 2.1), cone_angle=0.5)
    A = ts.operator(vg, pg)
    x = torch.from_numpy(ts.phantom.ellipsoids(24, 5, seed=9)).cuda()
    y_ref, b_ref = A(x), A.T(A(x))
    xa = dp.Array(x.clone())
    y = A(xa)
    assert isinstance(y, dp.Array) and torch.equal(y.to_torch(), y_ref)
    b = A.T(y)
    assert isinstance(b, dp.Array) and torch.equal(b.to_torch(), b_ref)
    out = dp.Array(torch.full(A.range_shape, 3.0, device="cuda"))
    A(xa, out=out)
    assert torch.equal(out.to_torch(), y_ref)
    # mixed with a torch tensor on the same device
    yt = torch.empty(A.range_shape, device="cuda")
    A(xa, out=yt)
    assert torch.equal(yt, y_ref)
    # a foreign array in host memory is staged like a numpy array
    xh = dp.Array(x.cpu())
    yh = A(xh)
    np.testing.assert_array_equal(yh.to_torch().numpy(), y_ref.cpu().numpy())
