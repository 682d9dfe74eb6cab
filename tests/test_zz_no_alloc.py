"""The allocation-free dot (ein_dot_f64_noalloc) against ein_dot_f64 through the C ABI.

Both entry points run the same two kernels (per-workgroup partial sums, then a fixed-order
sum of the partials), so their results must agree bit for bit. Two calls in a row on one
stream, the second with a smaller input; then the size limit and a strided case."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def _dot(lib, fn, n, alpha, x, incx, y, incy, beta, r, stream):
    # argument types come from the library's ctypes signatures (einsums_amd._lib)
    return getattr(lib, fn)(n, alpha, x.data_ptr(), incx, y.data_ptr(), incy, beta, r.data_ptr(), stream)


def test_noalloc_dot_matches_dot(core):
    lib = core.lib
    s = torch.cuda.current_stream()
    h = core.stream_handle(s)
    g = torch.Generator(device="cuda").manual_seed(77)
    # 40M elements (4883 partials), then 1.5M (184 partials): the second call reuses the
    # scratch the first one filled
    for n, beta in ((40_000_003, 0.0), (1_500_001, 0.5)):
        x = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) - 0.5
        y = torch.rand(n, dtype=torch.float64, device="cuda", generator=g) - 0.5
        r_ref = torch.full((1,), 3.0, dtype=torch.float64, device="cuda")
        r_new = r_ref.clone()
        assert _dot(lib, "ein_dot_f64", n, 0.25, x, 1, y, 1, beta, r_ref, h) == 0
        assert _dot(lib, "ein_dot_f64_noalloc", n, 0.25, x, 1, y, 1, beta, r_new, h) == 0
        torch.cuda.synchronize()
        assert torch.equal(r_ref, r_new), (n, r_ref.item(), r_new.item())
        exact = 0.25 * float(np.dot(x.cpu().numpy(), y.cpu().numpy())) + beta * 3.0
        assert abs(r_new.item() - exact) <= 1e-12 * abs(exact)
    # a strided and a reversed case (grid-stride kernel)
    x = torch.rand(3 * 100_001, dtype=torch.float64, device="cuda", generator=g)
    y = torch.rand(2 * 100_001, dtype=torch.float64, device="cuda", generator=g)
    r_ref = torch.zeros(1, dtype=torch.float64, device="cuda")
    r_new = torch.zeros(1, dtype=torch.float64, device="cuda")
    assert _dot(lib, "ein_dot_f64", 100_001, 1.0, x, 3, y, -2, 0.0, r_ref, h) == 0
    assert _dot(lib, "ein_dot_f64_noalloc", 100_001, 1.0, x, 3, y, -2, 0.0, r_new, h) == 0
    torch.cuda.synchronize()
    assert torch.equal(r_ref, r_new)


def test_noalloc_dot_size_limit(core):
    lib = core.lib
    limit = 65536 * 8192  # EIN_DOT_NOALLOC_MAX_N
    r = torch.zeros(1, dtype=torch.float64, device="cuda")
    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    # rejected on the host before any launch (the operands are far too short to be read)
    rc = _dot(lib, "ein_dot_f64_noalloc", limit + 1, 1.0, x, 1, x, 1, 0.0, r, core.stream_handle(torch.cuda.current_stream()))
    assert rc == 1005  # EIN_ERROR_UNSUPPORTED
    assert b"static scratch" in lib.ein_last_error_string()
    # the largest accepted length still runs (8 elements, far below the limit, on the same scratch)
    x8 = torch.arange(8, dtype=torch.float64, device="cuda")
    assert _dot(lib, "ein_dot_f64_noalloc", 8, 1.0, x8, 1, x8, 1, 0.0, r, core.stream_handle(torch.cuda.current_stream())) == 0
    torch.cuda.synchronize()
    assert r.item() == 140.0
