"""The FGMRES matvec mbem_gemv through the C-ABI against float64 products on
the host (its aligned form stages x through LDS 1,024 columns at a time)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n", [(1, 1), (7, 300), (45, 1024), (33, 2049), (130, 5003),
                                 (5, 9217)])
def test_gemv_matches_host_product(m, n):
    """Column counts with no tile, one, two and nine tiles, odd counts (the
    last column alone), row counts that leave waves past the last row; beta
    != 0; a second call with x one element off 16-byte alignment (the
    unvectorised form)."""
    import torch
    from mbem_amd._lib import call, ptr, stream
    rng = np.random.default_rng(m * 10007 + n)
    lda = n + (n & 1) + 2                       # even: 16-byte rows
    A = rng.standard_normal((m, lda))
    x = rng.standard_normal(n + 1)
    y0 = rng.standard_normal(m)
    dev = torch.device("cuda")
    Ad = torch.as_tensor(A, device=dev)
    xd = torch.as_tensor(x, device=dev)
    for off in (0, 1):
        yd = torch.as_tensor(y0, device=dev)
        call("mbem_gemv", ptr(Ad), lda, m, n, ptr(xd[off:]), ptr(yd), 0.75, -1.5, stream())
        xs = x[off:off + n]
        ref = 0.75 * (A[:, :n] @ xs) - 1.5 * y0
        scale = np.abs(A[:, :n]) @ np.abs(xs) + np.abs(y0)
        err = np.abs(yd.cpu().numpy() - ref) / scale
        assert err.max() < 1e-13, (off, err.max())


def test_gemv_zero_columns_leave_the_sums_bit_identical():
    """The staged form keeps each lane's column order: two matrices whose
    rows differ only past the first tile give equal sums over that tile."""
    import torch
    from mbem_amd._lib import call, ptr, stream
    rng = np.random.default_rng(3)
    m, n = 64, 4096
    A = rng.standard_normal((m, n))
    x = rng.standard_normal(n)
    x[1024:] = 0.0
    dev = torch.device("cuda")
    y_full = torch.empty(m, dtype=torch.float64, device=dev)
    y_head = torch.empty(m, dtype=torch.float64, device=dev)
    Ad = torch.as_tensor(A, device=dev)
    xd = torch.as_tensor(x, device=dev)
    call("mbem_gemv", ptr(Ad), n, m, n, ptr(xd), ptr(y_full), 1.0, 0.0, stream())
    Ah = torch.as_tensor(np.ascontiguousarray(A[:, :1024]), device=dev)
    call("mbem_gemv", ptr(Ah), 1024, m, 1024, ptr(xd), ptr(y_head), 1.0, 0.0, stream())
    # the zero columns add +0.0 terms, which leave every partial sum unchanged
    assert torch.equal(y_full, y_head)
