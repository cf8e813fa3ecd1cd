"""The allocation-free weighted Gram entry point against the one that takes a workspace.

``cdft_wgram_noalloc_f64`` keeps its block partials in the library's static device array; its results
must be bitwise those of ``cdft_wgram_f64`` (same kernels, same block order), on two calls in a row
with the second on a smaller input (the array keeps the first call's partials, which the second must
not read), with and without weights, and with the square-root post-processing.
"""

from __future__ import annotations

import pytest
import torch

pytestmark = pytest.mark.gpu


def _both(pkg, a, b, q, scale, post):
    from cdft_amd import ops

    n = a.shape[1]
    ka, kb = a.shape[0], b.shape[0]
    ref = torch.empty((ka, kb), dtype=torch.float64, device=a.device)
    ws = torch.empty(int(pkg._lib.load().cdft_reduction_workspace_doubles(n)), dtype=torch.float64, device=a.device)
    s = ops.stream_of(a)
    qp = None if q is None else q.data_ptr()
    pkg._lib.call("cdft_wgram_f64", a.data_ptr(), ka, b.data_ptr(), kb, qp, n, scale, post, ws.data_ptr(), ws.numel(),
                  ref.data_ptr(), s)
    out = torch.full((ka, kb), float("nan"), dtype=torch.float64, device=a.device)
    pkg._lib.call("cdft_wgram_noalloc_f64", a.data_ptr(), ka, b.data_ptr(), kb, qp, n, scale, post, out.data_ptr(), s)
    torch.cuda.synchronize()
    return out, ref


def test_noalloc_gram_is_the_workspace_gram(pkg, device):
    gen = torch.Generator().manual_seed(11)
    cases = [(5, 5, 274625, True, 1.0, 0), (3, 2, 35937, False, 0.5, 0), (1, 1, 4097, True, 1.0, 1),
             (12, 9, 70001, True, 2.0, 0), (2, 2, 100, False, 1.0, 1)]
    for ka, kb, n, weighted, scale, post in cases:  # calls in a row, each smaller than the first
        a = (torch.rand((ka, n), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
        b = a if ka == kb and post == 1 else (torch.rand((kb, n), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
        q = torch.rand(n, generator=gen, dtype=torch.float64).to(device) if weighted else None
        out, ref = _both(pkg, a, b, q, scale, post)
        assert torch.equal(out, ref), (ka, kb, n)


def test_noalloc_gram_rejects_bad_arguments(pkg, device):
    a = torch.ones((2, 10), dtype=torch.float64, device=device)
    out = torch.empty((2, 2), dtype=torch.float64, device=device)
    with pytest.raises(pkg._lib.HipLibraryError, match="sizes must be positive"):
        pkg._lib.call("cdft_wgram_noalloc_f64", a.data_ptr(), 0, a.data_ptr(), 2, None, 10, 1.0, 0, out.data_ptr(), 0)
    with pytest.raises(pkg._lib.HipLibraryError, match="null pointer"):
        pkg._lib.call("cdft_wgram_noalloc_f64", a.data_ptr(), 2, a.data_ptr(), 2, None, 10, 1.0, 0, None, 0)
