"""cipher_wrms_norm_noalloc (the norm without a context, partial sums in a
static device array of the library) against cipher_wrms_norm on the same
vectors: bit-identical on two calls in a row, the second smaller, on the
default stream and on torch's current stream."""
import ctypes as C
import math

import numpy as np
import pytest

from cipher_amd import abi

pytestmark = pytest.mark.gpu


def _noalloc(lib, U, Y, atol, rtol, stream=None):
    out = C.c_double(-1.0)
    rc = lib.cipher_wrms_norm_noalloc(C.c_void_p(U.data_ptr()), C.c_void_p(Y.data_ptr()), U.numel(), atol, rtol,
                                      C.c_void_p(stream), C.byref(out))
    abi.check(rc, "cipher_wrms_norm_noalloc")
    return out.value


def test_noalloc_norm_matches_context_norm():
    import torch
    from cipher_amd import workload as W
    from cipher_amd.runtime import Context
    lib = abi.load_library()
    wl = W.Workload(W.grain_growth_yaml(res=(16, 16), ngrains=4), "noalloc")
    ctx = Context(wl.cfg)
    gen = torch.Generator(device="cuda").manual_seed(3)
    atol, rtol = 1e-4, 1e-3
    # first call: more elements than the partition's 1024 partial sums cover
    # one each (grid-stride blocks); second call: a few thousand elements,
    # 2 partial sums, after the first call left 1024 in the array
    for n in (6_000_011, 5_000):
        U = torch.randn(n, dtype=torch.float64, device="cuda", generator=gen)
        Y = U + 1e-3 * torch.randn(n, dtype=torch.float64, device="cuda", generator=gen)
        print(f"[noalloc] n={n}", flush=True)
        ref = ctx.wrms_norm(U, Y, atol, rtol)
        got = _noalloc(lib, U, Y, atol, rtol)
        got_s = _noalloc(lib, U, Y, atol, rtol, torch.cuda.current_stream().cuda_stream)
        assert got == ref and got_s == ref, (n, got, got_s, ref)
        u, y = U.cpu().numpy(), Y.cpu().numpy()
        tol = atol + rtol * np.maximum(np.abs(u), np.abs(y))
        exact = math.sqrt(np.mean(((u - y) / tol) ** 2))
        assert got == pytest.approx(exact, rel=1e-12, abs=0.0)
    # n = 0 on a live buffer: the norm of nothing is 0, as for the context
    buf = torch.ones(8, dtype=torch.float64, device="cuda")
    out = C.c_double(-1.0)
    abi.check(lib.cipher_wrms_norm_noalloc(C.c_void_p(buf.data_ptr()), C.c_void_p(buf.data_ptr()), 0, atol, rtol,
                                           None, C.byref(out)), "cipher_wrms_norm_noalloc")
    assert out.value == 0.0
    ctx.close()


def test_noalloc_norm_argument_errors():
    lib = abi.load_library()
    out = C.c_double(0.0)
    assert lib.cipher_wrms_norm_noalloc(None, None, 4, 1e-4, 1e-4, None, C.byref(out)) == abi.ERR_ARG
    buf = (C.c_double * 4)()
    assert lib.cipher_wrms_norm_noalloc(buf, buf, -1, 1e-4, 1e-4, None, C.byref(out)) == abi.ERR_ARG
