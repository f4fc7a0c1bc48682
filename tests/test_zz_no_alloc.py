"""wl_apply_noalloc (wl_apply with its staging buffers in the library's static
per-device scratch array) against wl_apply through the C ABI: the same
kernels on the same inputs, so every operator's output must be bit-identical,
on two calls in a row (the second on a smaller input, which reuses the
scratch array holding the first call's contents), on a second context of a
smaller lattice, and in a fresh process or after other tests.  A call whose
staging does not fit the array is refused before any launch."""

from ctypes import c_void_p

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _apply(lib, fn, ctx, gauge, op, psi, nrhs):
    from jaxqft_amd import _native as nat

    out = psi.new_empty(psi.shape)
    nat.check(getattr(lib, fn)(ctx.h, gauge.h, op, c_void_p(psi.data_ptr()), c_void_p(out.data_ptr()), nrhs,
                               nat.WL_C128 if psi.dtype.itemsize == 16 else nat.WL_C64, None))
    return out


def _fields(L, nrhs, seed, nc=3):
    import torch

    from jaxqft_amd import fields

    gen = torch.Generator(device="cuda").manual_seed(seed)
    psi = torch.complex(torch.randn((nrhs, *L, 4, nc), generator=gen, device="cuda", dtype=torch.float64),
                        torch.randn((nrhs, *L, 4, nc), generator=gen, device="cuda", dtype=torch.float64))
    if nc == 3:
        U = fields.su3_hot_start_device(seed + 1, 1, L, 0.5)
    else:
        U = torch.from_numpy(fields.u1_hot_start(fields.rng_for(seed + 1), 1, L, 0.5)).cuda()
    return U, psi


def test_noalloc_matches_wl_apply_two_calls():
    print("[gpu] wl_apply_noalloc vs wl_apply", flush=True)
    import torch

    from jaxqft_amd import _native as nat

    lib = nat.load()
    bc = (1, 1, 1, -1)
    L = (8, 8, 8, 16)
    U, psi = _fields(L, 6, 810)
    ctx = nat.Context(L, 3, 1, 0.07, 1.0, bc)
    gauge = nat.Gauge(ctx, U, nat.WL_LAYOUT_BMXYIJ)
    ops = (nat.WL_OP_D, nat.WL_OP_DDAG, nat.WL_OP_G5D, nat.WL_OP_K, nat.WL_OP_KDAG, nat.WL_OP_DDAGD)
    for op in ops:
        # first call on 6 columns, the second on 2 (a smaller input): the scratch
        # array still holds the first call's staging when the second runs
        for n in (6, 2):
            x = psi[:n].contiguous()
            ref = _apply(lib, "wl_apply", ctx, gauge, op, x, n)
            got = _apply(lib, "wl_apply_noalloc", ctx, gauge, op, x, n)
            torch.cuda.synchronize()
            assert torch.equal(got, ref), (op, n)
            assert float(torch.linalg.norm(ref)) > 0
    # complex64 input and a second, smaller lattice (a U(1) context) on the same scratch array
    got = _apply(lib, "wl_apply_noalloc", ctx, gauge, nat.WL_OP_D, psi[:1].to(torch.complex64), 1)
    ref = _apply(lib, "wl_apply", ctx, gauge, nat.WL_OP_D, psi[:1].to(torch.complex64), 1)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    L2 = (4, 4, 4, 8)
    U2, psi2 = _fields(L2, 3, 820, nc=1)
    ctx2 = nat.Context(L2, 1, 1, 0.1, 1.0, bc)
    gauge2 = nat.Gauge(ctx2, U2, nat.WL_LAYOUT_BMXYIJ)
    for op in (nat.WL_OP_D, nat.WL_OP_DDAGD):
        got = _apply(lib, "wl_apply_noalloc", ctx2, gauge2, op, psi2, 3)
        ref = _apply(lib, "wl_apply", ctx2, gauge2, op, psi2, 3)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), op


def test_noalloc_refuses_staging_beyond_the_scratch_array():
    print("[gpu] wl_apply_noalloc size limit", flush=True)
    import torch

    from jaxqft_amd import _native as nat

    lib = nat.load()
    L = (16, 16, 16, 16)
    U, psi = _fields(L, 2, 830)
    ctx = nat.Context(L, 3, 1, 0.05, 1.0, (1, 1, 1, -1))
    gauge = nat.Gauge(ctx, U, nat.WL_LAYOUT_BMXYIJ)
    col = 192 * (int(np.prod(L)) + 288)
    assert 2 * 2 * col <= nat.WL_NOALLOC_SCRATCH_BYTES < 3 * 2 * col
    # two columns of D fit (2 buffers); two columns of D^dag D (3 buffers) do not
    got = _apply(lib, "wl_apply_noalloc", ctx, gauge, nat.WL_OP_D, psi, 2)
    ref = _apply(lib, "wl_apply", ctx, gauge, nat.WL_OP_D, psi, 2)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    out = torch.zeros_like(psi)
    rc = lib.wl_apply_noalloc(ctx.h, gauge.h, nat.WL_OP_DDAGD, c_void_p(psi.data_ptr()), c_void_p(out.data_ptr()), 2,
                              nat.WL_C128, None)
    assert rc == nat.WL_ERR_BAD_ARG
    assert "scratch" in lib.wl_last_error().decode()
    torch.cuda.synchronize()
    assert float(torch.linalg.norm(out)) == 0.0  # nothing was launched
