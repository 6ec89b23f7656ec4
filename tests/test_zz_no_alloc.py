"""thcgw_primary_to_aux_noalloc (the entry point that allocates no device memory) against
thcgw_primary_to_aux, through the C ABI.

Two calls in a row on each shape, the second with a smaller batch, both equal to the
allocating entry point's result (same GEMM kernels on the same inputs: identical values).
Shapes cover both chunkings of the 64 MiB staging buffer: whole blocks of nk matrices per
chunk (small N_mu N_orb) and k ranges within a block (N_mu = 2048, N_orb = 40, N_k = 64:
51 of the 1.3 MB intermediates fit).  A shape whose single intermediate exceeds the buffer
is refused with THCGW_ERR_ARG before any work.
"""
import ctypes

import numpy as np
import pytest

from coqui_amd import IAFT, ThcGW

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a HIP device"
    return _t


def cgauss(torch, gen, *shape):
    re = torch.randn(*shape, dtype=torch.float64, device="cuda", generator=gen)
    im = torch.randn(*shape, dtype=torch.float64, device="cuda", generator=gen)
    return torch.complex(re, im)


def call(ctx, name, nb, X, O, out, torch):
    fn = getattr(ctx.lib, name)
    rc = fn(ctx.h, ctypes.c_long(nb), ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(O.data_ptr()),
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc


@pytest.mark.parametrize("kgrid,no,nmu,nbs", [((2, 2, 2), 6, 40, (3, 1)), ((4, 4, 4), 40, 2048, (2, 1))])
def test_noalloc_matches_allocating_entry(torch, kgrid, no, nmu, nbs):
    print("case: noalloc primary_to_aux kgrid=%s no=%d nmu=%d nb=%s" % (kgrid, no, nmu, nbs), flush=True)
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    ctx = ThcGW(kgrid, no, nmu, ia)
    nk = ctx.nk
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    X = cgauss(torch, gen, nk, nmu, no)
    try:
        for nb in nbs:
            O = cgauss(torch, gen, nb, nk, no, no)
            ref = torch.empty((nb, nk, nmu, nmu), dtype=torch.complex128, device="cuda")
            got = torch.full((nb, nk, nmu, nmu), float("nan"), dtype=torch.complex128, device="cuda")
            assert call(ctx, "thcgw_primary_to_aux", nb, X, O, ref, torch) == 0
            assert call(ctx, "thcgw_primary_to_aux_noalloc", nb, X, O, got, torch) == 0
            assert bool(torch.isfinite(got).all())
            assert float((got - ref).abs().max()) == 0.0
            del O, ref, got
            torch.cuda.empty_cache()
    finally:
        ctx.close()


def test_noalloc_refuses_oversize(torch):
    """One 2048 x 2100 complex<double> intermediate (68.8 MB) exceeds the 64 MiB buffer."""
    ia = IAFT.from_table(1000.0, "1e2.1e-15")
    no, nmu = 2100, 2048
    ctx = ThcGW((1, 1, 1), no, nmu, ia)
    try:
        X = torch.zeros((1, nmu, no), dtype=torch.complex128, device="cuda")
        O = torch.zeros((1, 1, no, no), dtype=torch.complex128, device="cuda")
        out = torch.zeros((1, 1, nmu, nmu), dtype=torch.complex128, device="cuda")
        rc = call(ctx, "thcgw_primary_to_aux_noalloc", 1, X, O, out, torch)
        assert rc == 1  # THCGW_ERR_ARG
        assert "exceeds" in ctx.lib.thcgw_last_error(ctx.h).decode()
    finally:
        ctx.close()
