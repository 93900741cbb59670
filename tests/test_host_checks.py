"""Host-side argument checks of the Python mirror (no GPU needed): tensors
handed to device entry points are rejected before any C call when their
dtype, size, layout or device does not match what the kernels assume."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_dev_ptrs_rejects_bad_tensors():
    from giem2g_amd.operator import dev_ptrs
    n = 24
    with pytest.raises(ValueError, match="complex128"):
        dev_ptrs(n, "c128", torch.zeros(n, dtype=torch.complex64))
    with pytest.raises(ValueError, match="complex64"):
        dev_ptrs(n, "c64", torch.zeros(n, dtype=torch.complex128))
    with pytest.raises(ValueError, match="24 elements"):
        dev_ptrs(n, "c128", torch.zeros(n - 1, dtype=torch.complex128))
    with pytest.raises(ValueError, match="contiguous"):
        dev_ptrs(n, "c128", torch.zeros(2 * n, dtype=torch.complex128)[::2])
    with pytest.raises(ValueError, match="on the GPU"):
        dev_ptrs(n, "c128", torch.zeros(n, dtype=torch.complex128))
    with pytest.raises(ValueError, match="torch tensor"):
        dev_ptrs(n, "c128", np.zeros(n, np.complex128))
    with pytest.raises(ValueError, match="must be a tensor"):
        dev_ptrs(n, "c128", None)
    # a nullable position passes None through (the check of the other argument still runs)
    with pytest.raises(ValueError, match="on the GPU"):
        dev_ptrs(n, "c128", torch.zeros(n, dtype=torch.complex128), None, allow_none=(1,))


def test_precision_of_batch_vectors():
    from giem2g_amd.operator import _prec_of
    assert _prec_of(torch.zeros(2, dtype=torch.complex128)) == (0, "c128")
    assert _prec_of(torch.zeros(2, dtype=torch.complex64)) == (1, "c64")
    with pytest.raises(ValueError, match="complex128 or complex64"):
        _prec_of(torch.zeros(2, dtype=torch.float64))
