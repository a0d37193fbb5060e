"""svdh_spmv_noalloc: the allocation-free SpMV (partials in a static device
scratch) against svdh_spmv with device vectors on the same operator, bit for
bit (the same kernels in the same order), on two calls in a row, the second
with a smaller operator; plus its two error paths (an operator not prepared
for the product, and a product whose partials exceed the scratch)."""
import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import single_svdlib
    from single_svdlib import _ffi
    _ffi.context(0)
    return single_svdlib


def noalloc():
    from single_svdlib import _ffi
    return _ffi.lib().svdh_spmv_noalloc


def matrix(m, n, per_row, long_rows, seed):
    """Random CSR with per_row entries per row plus a few rows of 6,000
    entries (chunked rows: several partials per output row)."""
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(m), per_row)
    cols = rng.integers(0, n, size=rows.size)
    extra_r = np.repeat(rng.choice(m, long_rows, replace=False), 6000)
    extra_c = rng.integers(0, n, size=extra_r.size)
    r = np.concatenate([rows, extra_r])
    c = np.concatenate([cols, extra_c])
    a = sp.csr_matrix((rng.normal(size=r.size), (r, c)), shape=(m, n))
    a.sum_duplicates()
    return a


def run_both(lib, a, dtype):
    from single_svdlib import _ffi
    ctx = _ffi.context(0)
    op = lib.CsrMatrix.from_scipy(a)
    code = _ffi.SVDH_F64 if dtype == np.float64 else _ffi.SVDH_F32
    for transposed in (0, 1):
        nin, nout = (a.shape[0], a.shape[1]) if transposed else (a.shape[1], a.shape[0])
        x = _ffi.DeviceArray((nin,), dtype).upload(np.random.default_rng(transposed).normal(size=nin))
        y1 = _ffi.DeviceArray((nout,), dtype)
        y2 = _ffi.DeviceArray((nout,), dtype)
        # the allocating call builds the schedule and value layouts
        _ffi.check(_ffi.lib().svdh_spmv(ctx, op._handle, transposed, x.ptr, y1.ptr, code, 1))
        _ffi.check(noalloc()(ctx, op._handle, transposed, x.ptr, y2.ptr, code))
        _ffi.check(_ffi.lib().svdh_ctx_synchronize(ctx))
        want, got = y1.numpy(), y2.numpy()
        np.testing.assert_array_equal(got, want)
        ref = (a.T if transposed else a) @ x.numpy().astype(np.float64)
        assert np.max(np.abs(want - ref)) <= 1e-3 * np.max(np.abs(ref))
    return op


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_spmv_equals_spmv_twice(lib, dtype):
    # first: 200,000 x 40,000 with 1.3 M nonzeros (source-tiled layout, several
    # source blocks, so partials in both directions); second: a smaller
    # operator (streamed tasks with chunked long rows)
    print(f"case noalloc {np.dtype(dtype).name} big", flush=True)
    run_both(lib, matrix(200_000, 40_000, 6, 20, seed=1), dtype)
    print(f"case noalloc {np.dtype(dtype).name} small", flush=True)
    run_both(lib, matrix(5_000, 3_000, 8, 4, seed=2), dtype)


def test_noalloc_spmv_errors(lib):
    from single_svdlib import _ffi
    ctx = _ffi.context(0)
    a = matrix(3_000, 2_000, 5, 0, seed=3)
    op = lib.CsrMatrix.from_scipy(a)
    x = _ffi.DeviceArray((2_000,)).upload(np.ones(2_000))
    y = _ffi.DeviceArray((3_000,))
    st = noalloc()(ctx, op._handle, 0, x.ptr, y.ptr, _ffi.SVDH_F64)
    assert st == 10 and b"not prepared" in _ffi.lib().svdh_last_error()
    # 800,000 x 40,000: the forward product has 3 source blocks of 800,000
    # partials each, above the scratch's 2^21
    big = matrix(800_000, 40_000, 2, 0, seed=4)
    opb = lib.CsrMatrix.from_scipy(big)
    xb = _ffi.DeviceArray((40_000,)).upload(np.ones(40_000))
    yb = _ffi.DeviceArray((800_000,))
    _ffi.check(_ffi.lib().svdh_spmv(ctx, opb._handle, 0, xb.ptr, yb.ptr, _ffi.SVDH_F64, 1))
    st = noalloc()(ctx, opb._handle, 0, xb.ptr, yb.ptr, _ffi.SVDH_F64)
    assert st == 10 and b"at most" in _ffi.lib().svdh_last_error()
