"""Host-to-device uploads through the pinned staging ring (``_lib.upload``).

The ring is reused once the device has copied a region, so a run of uploads
whose bytes pass the ring's capacity while the device is still busy (the
copies queued behind a long product) must come out whole: each upload equals
its host array, in the requested dtype and shape.  The empty array, a 0-d
array, dtype conversion on the way in and an upload too large for the ring
(its own pinned copy) are covered too.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_uploads_survive_ring_wrap_while_device_busy():
    import torch
    from mbem_amd import _lib
    dev = _lib.device()
    rng = np.random.default_rng(7)
    m = torch.randn(4096, 4096, dtype=torch.float64, device=dev)
    for _ in range(4):                      # the copies queue behind these
        m = m @ m * 1e-2
    chunk = 5 << 20                         # 5 MiB per upload, 120 MiB in all
    arrays = [rng.standard_normal(chunk // 8) for _ in range(24)]
    outs = [_lib.upload(a, torch.float64) for a in arrays]
    assert sum(a.nbytes for a in arrays) > 1.5 * _lib._StagingRing.CAPACITY
    torch.cuda.synchronize()
    for a, o in zip(arrays, outs):
        assert o.dtype == torch.float64 and o.shape == a.shape
        assert np.array_equal(o.cpu().numpy(), a)


def test_upload_dtypes_shapes_and_sizes():
    import torch
    from mbem_amd import _lib
    pts = np.arange(60, dtype=np.float64).reshape(5, 4, 3)
    idx32 = np.arange(1000, dtype=np.int32)[::-1]
    cases = [(pts, torch.float64), (idx32, torch.int64), ([1, 2, 3], torch.int64),
             (np.float64(2.5), torch.float64), (np.zeros((0, 3)), torch.float64),
             (np.linspace(0, 1, 17), torch.float32)]
    outs = [_lib.upload(a, dt) for a, dt in cases]
    big = np.arange(_lib._StagingRing.CAPACITY // 8 + 5, dtype=np.int64)
    out_big = _lib.upload(big, torch.int64)            # larger than half the ring
    torch.cuda.synchronize()
    for (a, dt), o in zip(cases, outs):
        want = torch.as_tensor(np.ascontiguousarray(a), dtype=dt)
        assert o.dtype == dt and tuple(o.shape) == tuple(want.shape)
        assert torch.equal(o.cpu(), want)
    assert torch.equal(out_big.cpu(), torch.as_tensor(big))
