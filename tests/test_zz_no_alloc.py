"""The allocation-free CSR product (launch_spmv_chunked_noalloc, the chunk partial sums in the library's own device
scratch) against launch_spmv_chunked with a caller workspace, through the C ABI: bit for bit on two calls in a row,
the second with a smaller input (the scratch keeps the first call's partials, which must not leak into the second),
and a size beyond the scratch refused before any launch."""
import ctypes as ct

import numpy as np
import pytest
import torch

from flyhip import hip
from helpers import random_csr

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _pointers(tensors):
    return (ct.c_void_p * len(tensors))(*(t.data_ptr() for t in tensors))


def _product(lib, csr, x, noalloc, scale=1.0, accumulate=False, y0=None):
    B = x.shape[0]
    y = torch.zeros(B, csr.shape[0], device=DEV) if y0 is None else y0.clone()
    base = [csr.ptr, csr.chunk_row, csr.chunk_ptr, csr.idx, csr.values, x, y]
    stream = torch.cuda.current_stream().cuda_stream
    if noalloc:
        err = lib.launch_spmv_chunked_noalloc(_pointers(base), csr.shape[0], csr.shape[1], csr.nchunks, B,
                                              float(scale), int(accumulate), stream)
    else:
        work = torch.empty(max(1, B * csr.nchunks), device=DEV)
        err = lib.launch_spmv_chunked(_pointers(base + [work]), csr.shape[0], csr.shape[1], csr.nchunks, B,
                                      float(scale), int(accumulate), stream)
    assert err == 0, err
    torch.cuda.synchronize()
    return y


def test_noalloc_product_equals_the_workspace_product():
    lib = hip.lib(DEV)
    rng = np.random.default_rng(21)
    cases = [(random_csr(rng, 3000, 4000, 0.01, long_rows=(5, 17), long_len=900), 3, 1.0, False),
             (random_csr(rng, 500, 700, 0.03), 1, 0.5, True)]            # the second call is smaller
    for M, B, scale, acc in cases:
        csr = hip.CSR(M, DEV, kernel="chunked")
        x = torch.as_tensor(rng.normal(0, 1, (B, M.shape[1])).astype(np.float32), device=DEV)
        y0 = torch.as_tensor(rng.normal(0, 1, (B, M.shape[0])).astype(np.float32), device=DEV)
        want = _product(lib, csr, x, False, scale, acc, y0)
        got = _product(lib, csr, x, True, scale, acc, y0)
        assert torch.equal(got, want)
        ref = (M.astype(np.float64) @ x.double().cpu().numpy().T).T * scale + (y0.double().cpu().numpy() if acc else 0)
        np.testing.assert_allclose(got.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)


def test_noalloc_product_refuses_sizes_beyond_its_scratch():
    lib = hip.lib(DEV)
    dummy = torch.zeros(8, device=DEV)
    arr = _pointers([dummy] * 7)
    # 2 * 2,500,000 chunks > 4,194,304 floats: refused on the host (FH_ERR_SIZE), nothing launched
    err = lib.launch_spmv_chunked_noalloc(arr, 10, 10, 2_500_000, 2, 1.0, 0, torch.cuda.current_stream().cuda_stream)
    assert err == -2
