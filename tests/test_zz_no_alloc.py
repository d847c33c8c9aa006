"""lumen_ex_photometric_loss_noalloc through the C ABI: bit for bit what
lumen_ex_photometric_loss (which holds its partial sums in a context) gives,
on two calls in a row with the second image smaller (the fixed device array
keeps the first call's contents), and a clear error past its maximum size."""
import ctypes
import sys

import numpy as np
import pytest

from lumen_amd import _lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(W, H, seed):
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    pred = torch.rand((H, W, 3), generator=g, device="cuda")
    target = torch.rand((H, W, 3), generator=g, device="cuda")
    return pred, target


def _call(fn, ctx, pred, target, W, H, lam):
    import torch
    out = torch.full((3,), float("nan"), device="cuda")
    grad = torch.full((H, W, 3), float("nan"), device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    if ctx is None:
        rc = fn(pred.data_ptr(), target.data_ptr(), W, H, lam, out.data_ptr(), grad.data_ptr(), st)
    else:
        rc = fn(ctx, pred.data_ptr(), target.data_ptr(), W, H, lam, out.data_ptr(),
                grad.data_ptr(), st)
    _lib.check(fn.__name__, rc)
    torch.cuda.synchronize()
    return out, grad


def test_noalloc_loss_equals_context_loss():
    import torch
    L = _lib.lib()
    ctx = L.lumen_ex_context_create(0)
    assert ctx
    try:
        for W, H, seed in ((1920, 1080, 1), (333, 217, 2), (8192, 8192, 3), (7, 5, 4)):
            say(f"[noalloc] {W}x{H}")
            pred, target = _pair(W, H, seed)
            a_out, a_grad = _call(L.lumen_ex_photometric_loss_noalloc, None, pred, target, W, H, 0.2)
            b_out, b_grad = _call(L.lumen_ex_photometric_loss, ctx, pred, target, W, H, 0.2)
            assert torch.equal(a_out, b_out), (W, H)
            assert torch.equal(a_grad, b_grad), (W, H)
            assert bool(torch.isfinite(a_out).all()) and bool(torch.isfinite(a_grad).all())
            del pred, target, a_grad, b_grad
    finally:
        L.lumen_ex_context_destroy(ctx)


def test_noalloc_loss_rejects_images_past_its_array():
    import torch
    L = _lib.lib()
    dummy = torch.zeros(16, device="cuda")
    p = dummy.data_ptr()
    rc = L.lumen_ex_photometric_loss_noalloc(p, p, 16384, 8200, 0.2, p, None,
                                             torch.cuda.current_stream().cuda_stream)
    assert rc == -1
    assert "8192 x 8192" in L.lumen_ex_last_error().decode()
